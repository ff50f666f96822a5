"""The single-pass adjoint kernel (MA(1) carried forward through per-tile
(A, B) maps, no eps scratch) against the two-pass kernel on the same models
and draws, and both against the FP64 torch autograd reference: every model
feature the single-pass kernel handles (MA(1) per instrument with free or
fixed timescales, the global MA, terms added after the MA, acceleration,
parameterisations with phase and with time of periastron), several tiles
(752 epochs are three 256-epoch tiles), FP64 and the mixed path."""
import numpy as np
import pytest
import torch

from tests.conftest import make_sim
from tests.torch_reference import lnl_torch
from tests.test_gpu_grad import _draws

pytestmark = pytest.mark.gpu

CASES = {
    "plain": dict(folder="GJ876", k=2, parameterisation=1),
    "ma_ins": dict(folder="GJ876", k=1, parameterisation=1, moav={"order": 1, "global": False},
                   semantics="intended"),
    "ma_ins_fixed_tau": dict(folder="GJ876", k=1, parameterisation=1, moav={"order": 1, "global": False},
                             semantics="intended",
                             conditions=[(f"MATimescale {i} Order 1", "fixed", 5.0) for i in range(1, 9)]),
    "ma_global_acc_p6": dict(folder="GJ876", k=2, parameterisation=6, moav={"order": 1, "global": True},
                             acceleration=2),
    "post_ma_p3": dict(folder="GJ876", k=2, parameterisation=3, moav={"order": 1, "global": False},
                       semantics="intended", sinusoid=1, switch_SA=True),
    "magnetic_p4": dict(folder="GJ876", k=1, parameterisation=4, magnetic_cycle=1,
                        moav={"order": 1, "global": True}),
    "short_p0_acc": dict(folder="51Peg", k=2, parameterisation=0, acceleration=1),
}


def _model(case):
    kw = dict(CASES[case])
    folder, k = kw.pop("folder"), kw.pop("k")
    return make_sim(folder, k, **kw).model


@pytest.mark.parametrize("case", list(CASES))
def test_single_pass_matches_two_pass_and_autograd(case):
    m = _model(case)
    th = _draws(m, 12, np.random.default_rng(sum(map(ord, case))))
    out = {}
    for kern in ("two_pass", "single_pass"):
        m.set_grad_kernel(kern)
        out[kern] = {p: m.lnl_grad_batch(th, p) for p in ("fp64", "mixed")}
    m.set_grad_kernel("auto")
    for p in ("fp64", "mixed"):
        (l2, g2), (l1, g1) = out["two_pass"][p], out["single_pass"][p]
        assert np.array_equal(np.isfinite(l1), np.isfinite(l2))
        fin = np.isfinite(l2)
        dl = np.max(np.abs(l1[fin] - l2[fin]) / np.abs(l2[fin]))
        dg = np.max(np.abs(g1[fin] - g2[fin]) / np.maximum(np.abs(g2[fin]), 1.0))
        print(f"{case} {p}: single vs two-pass |dlnL|/|lnL| {dl:.1e}, gradient {dg:.1e}", flush=True)
        assert dl < 1e-12
        assert dg < 1e-9
    l1, g1 = out["single_pass"]["fp64"]
    worst = 0.0
    for b in range(3):
        t = torch.tensor(th[b], dtype=torch.float64, requires_grad=True)
        lnl_torch(m, t).backward()
        ga = t.grad.numpy()
        worst = max(worst, np.max(np.abs(g1[b] - ga) / np.maximum(np.abs(ga), 1.0)))
    print(f"{case}: single-pass vs autograd {worst:.1e}", flush=True)
    assert worst < 1e-8


def test_single_pass_refuses_higher_ma_orders():
    from astroemperor_amd import _abi
    m = make_sim("GJ876", 1, parameterisation=1, moav={"order": 2, "global": False}, semantics="intended").model
    with pytest.raises(_abi.EmpError):
        m.set_grad_kernel("single_pass")
    m.set_grad_kernel("two_pass")
    m.set_grad_kernel("auto")
