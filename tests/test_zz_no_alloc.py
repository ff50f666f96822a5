"""The allocation-free adjoint entry point (emp_lnl_grad_batch_noalloc)
against emp_lnl_grad_batch, both called through the C ABI on device
pointers: identical lnL and gradients, bit for bit, on two calls in a row
(the second smaller), with and without the moving average, in FP64 and with
the mixed Kepler solve; lnL against the oracle; a batch beyond the static
scratch and host pointers are refused before anything is launched."""
import ctypes as C

import numpy as np
import pytest
import torch

from tests.conftest import make_sim, oracle_for

pytestmark = pytest.mark.gpu


def _draws(model, n, rng):
    lim = model.limits_free()
    mid, half = lim.mean(1), (lim[:, 1] - lim[:, 0]) / 2 * 0.6
    th = mid + half * rng.uniform(-1, 1, size=(n, len(mid)))
    for k in range(model.kplan__):
        s = th[:, 5 * k + 3] ** 2 + th[:, 5 * k + 4] ** 2
        th[:, 5 * k + 3:5 * k + 5] *= np.minimum(1.0, np.sqrt(0.8 / s))[:, None]
    return th


def _call(name, model, th, flags):
    from astroemperor_amd import _abi
    B, D = th.shape
    lnl = torch.full((B,), np.nan, dtype=torch.float64, device="cuda")
    grad = torch.full((B, D), np.nan, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    fn = getattr(_abi.lib(), name)
    _abi.check(fn(model.device.handle, C.c_void_p(th.data_ptr()), B, D, C.c_void_p(lnl.data_ptr()),
                  C.c_void_p(grad.data_ptr()), flags, C.c_void_p(stream)))
    torch.cuda.synchronize()
    return lnl.cpu().numpy(), grad.cpu().numpy()


@pytest.mark.parametrize("ma", [True, False])
@pytest.mark.parametrize("mixed", [False, True])
def test_noalloc_matches_allocating_entry(ma, mixed):
    from astroemperor_amd import _abi
    kw = {"moav": {"order": 1, "global": False}, "semantics": "intended"} if ma else {}
    m = make_sim("GJ876", 2, parameterisation=1, **kw).model
    flags = _abi.EMP_GRAD_MIXED if mixed else 0
    o = oracle_for(m)
    rng = np.random.default_rng(17 + 2 * ma + mixed)
    for B in (700, 150):
        th = torch.tensor(_draws(m, B, rng), dtype=torch.float64, device="cuda")
        got = _call("emp_lnl_grad_batch_noalloc", m, th, flags)     # before the model grows its scratch
        want = _call("emp_lnl_grad_batch", m, th, flags)
        assert np.isfinite(got[0]).all() and np.isfinite(got[1]).all(), B
        assert np.array_equal(got[0], want[0]), B
        assert np.array_equal(got[1], want[1]), B
        x = th.cpu().numpy()
        for b in (0, B // 2, B - 1):
            ref = o.my_likelihood(x[b])
            assert abs(got[0][b] - ref) <= (1e-6 if mixed else 1e-10 * abs(ref)), (B, b, got[0][b], ref)


def test_noalloc_refuses_what_it_cannot_hold():
    from astroemperor_amd import _abi
    m = make_sim("GJ876", 1, parameterisation=1, moav={"order": 1, "global": False}, semantics="intended").model
    N = m.data.ndat
    B = int(_abi.EMP_NOALLOC_SCRATCH_DOUBLES // N + 1)
    th = torch.tensor(_draws(m, B, np.random.default_rng(5)), dtype=torch.float64, device="cuda")
    with pytest.raises(_abi.EmpError, match="scratch"):
        _call("emp_lnl_grad_batch_noalloc", m, th, 0)
    small = th[:4].contiguous()
    with pytest.raises(_abi.EmpError, match="device pointers"):
        _call("emp_lnl_grad_batch_noalloc", m, small, _abi.EMP_HOST_POINTERS)
    # the largest batch it holds runs and agrees with the allocating entry point
    top = th[:B - 1].contiguous()
    got = _call("emp_lnl_grad_batch_noalloc", m, top, 0)
    want = _call("emp_lnl_grad_batch", m, top, 0)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
