"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares, and the ctypes binding table covers each of them
(nothing here calls into the device)."""
import pathlib
import re

from astroemperor_amd import _abi

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(emp_\w+)\s*\(", txt, flags=re.M))
    return names


def test_library_exports_every_declared_function():
    names = _declared()
    assert {"emp_model_create", "emp_pt_run", "emp_pt_set_moves", "emp_systems_run"} <= names
    L = _abi.lib()
    missing = sorted(n for n in names if not hasattr(L, n))
    assert not missing, missing


def test_binding_table_covers_the_header():
    names = _declared()
    unbound = sorted(names - set(_abi._SIGS))
    assert not unbound, unbound
    stale = sorted(set(_abi._SIGS) - names)
    assert not stale, stale
