"""Descriptor validation of the C ABI (no GPU needed: the library checks the
descriptor before it touches the device): sizes beyond the kernels' limits
are refused with EMP_EUNSUP, malformed descriptors with EMP_EINVAL."""
import ctypes as C

import numpy as np
import pytest

from astroemperor_amd import _abi, synth


def _create(desc):
    h = C.c_void_p()
    return _abi.lib().emp_model_create(C.byref(desc), C.byref(h))


def test_limits_refused():
    m = synth._multi(95, 9, [40, 40], 1000.0, 1, (1, 2, 1, 1), ma=False)["model"]    # 9 > 8 planets
    desc, keep = m.descriptor()
    assert _create(desc) == -4
    m = synth._multi(96, 1, [6] * 17, 1000.0, 1, (1, 2, 1, 1), ma=False)["model"]   # 17 > 16 instruments
    desc, keep = m.descriptor()
    assert _create(desc) == -4


def test_malformed_descriptors_refused():
    m = synth._multi(97, 1, [30, 30], 1000.0, 1, (1, 2, 1, 1), ma=True)["model"]
    desc, keep = m.descriptor()
    good = (desc.abi_version, desc.off_offset, desc.n_epochs)
    desc.abi_version = 999
    assert _create(desc) == -1
    desc.abi_version = good[0]
    desc.off_offset = desc.ndim_full            # block beyond the vector
    assert _create(desc) == -1
    desc.off_offset = good[1]
    t = np.ascontiguousarray(np.array(keep[0])[::-1])                 # epochs out of time order
    desc.t = t.ctypes.data_as(C.POINTER(C.c_double))
    assert _create(desc) == -1
    desc.t = keep[0].ctypes.data_as(C.POINTER(C.c_double))
    desc.off_sinusoid = desc.ndim_full - 1                            # sinusoid block beyond the vector
    assert _create(desc) == -1
    desc.off_sinusoid = -1
    desc.n_sai = 2                                                    # activity columns without arrays
    desc.off_sai = 0
    assert _create(desc) == -1
    desc.n_sai, desc.off_sai = 0, -1
    desc.ma_order = 5                                 # an MA order the block's length does not hold
    assert _create(desc) == -1
    assert b"MA block" in _abi.lib().emp_last_error()


def test_ma_order_above_four_refused():
    from tests.conftest import make_sim
    m = make_sim("GJ876", 1, parameterisation=1, moav={"order": 5, "global": True}).model
    desc, keep = m.descriptor()
    assert _create(desc) == -4
    assert b"orders 1 to 4" in _abi.lib().emp_last_error()


def test_ladder_mode_values():
    from astroemperor_amd.sampler import ladder_mode
    assert [ladder_mode(m) for m in (None, -1, 0, 1, 2)] == [0, -1, 0, 1, 2]
    for bad in (3, -2):
        with pytest.raises(NotImplementedError):
            ladder_mode(bad)
