"""Builds libemp_mi355x.so (HIP, gfx950) in-tree with hipcc.

    python -m astroemperor_amd.build [--verbose]
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libemp_mi355x.so")
SOURCES = ["emp_model.hip", "emp_pt.hip", "emp_grad.hip", "emp_systems.hip"]
# Per-source extra flags (none at present).
EXTRA: dict = {}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
         # HIP's default (fast) ignores `#pragma clang fp contract(off)`, which
         # the prior needs to reproduce numpy's unfused arithmetic bit for bit
         "-ffp-contract=fast-honor-pragmas",
         # no packed-FP32 instructions in device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # The likelihood runs inside loops over tiles, walkers and systems;
         # machine LICM hoists the 64-bit constants of every inlined log/exp
         # out of them and keeps them live in ~30-60 VGPRs through the hot
         # planet loop (the batched-systems sweep reached 240+ VGPRs, or
         # scratch spills at 1024 threads).  Left in place they cost a move
         # where they are used; measured on the half-step kernel: -17% time
         # (mixed) and -10% (FP64) at 3 planets x 1530 epochs.
         "-mllvm", "-disable-machine-licm",
         "-Wno-unused-command-line-argument"]


def _clean(stderr: str) -> str:
    noise = "'-packed-fp32-ops' is not a recognized feature"
    return "\n".join(l for l in (stderr or "").splitlines() if noise not in l)


def build(verbose: bool = False) -> str:
    """Compiles every source to an object (in parallel), then links the .so."""
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), "-c", path, "-o", obj]
        cmd.remove("-shared")
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    objs, errors = [], []
    for src, obj, proc in jobs:
        _, err = proc.communicate()
        err = _clean(err)
        if proc.returncode != 0:
            errors.append(f"{src}: hipcc failed ({proc.returncode}):\n{err}")
        elif verbose and err.strip():
            print(err)
        objs.append(obj)
    if errors:
        raise RuntimeError("\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc link failed ({r.returncode}):\n{_clean(r.stderr)}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
