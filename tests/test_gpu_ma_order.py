"""Moving average of order q = 2..4 (moav00/moav01 with q > 1) on the GPU
against the oracle: the batched lnL at every launch shape (FP64 within 1e-10
relative, mixed within 1e-6 absolute of FP64), my_model, the PT engine's and
the batched-systems engine's stored lnL; the order-q adjoint against FP64
autograd and finite differences, the gradient move on such a model; orders
above 4 are refused (EMP_EUNSUP)."""
import numpy as np
import pytest

from tests.conftest import make_sim, oracle_for
from tests.test_gpu_lnl import _check_lnl, _draws

pytestmark = pytest.mark.gpu

CASES = {
    "q2_global": dict(moav={"order": 2, "global": True}),
    "q2_ins": dict(moav={"order": 2, "global": False}, semantics="intended"),
    "q3_global": dict(moav={"order": 3, "global": True}),
    "q4_ins": dict(moav={"order": 4, "global": False}, semantics="intended"),
    "q2_ins_post": dict(moav={"order": 2, "global": False}, semantics="intended", sinusoid=1, switch_SA=True),
}


def _model(case, k=1):
    return make_sim("GJ876", k, parameterisation=1, **CASES[case]).model


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("wpt", [1, 2, 4, 8])
def test_lnl_matches_oracle(case, wpt):
    m = _model(case)
    m.set_waves_per_theta(wpt)
    th = _draws(m, 64, np.random.default_rng(wpt))
    e, dm = _check_lnl(m, th, f"{case} wpt{wpt}")
    print(f"{case} wpt={wpt}: fp64 rel {e:.2e}, |mixed - fp64| {dm:.2e}", flush=True)
    assert dm < 1e-6


@pytest.mark.parametrize("case", ["q2_global", "q3_global", "q4_ins"])
def test_my_model_matches_oracle(case):
    m = _model(case)
    o = oracle_for(m)
    keep = [t for t in _draws(m, 60, np.random.default_rng(3)) if np.isfinite(o.my_likelihood(t))][:3]
    assert keep
    for th in keep:
        mu, e2 = m.my_model(th)
        mu_o, e2_o = o.my_model(th)
        assert np.allclose(mu, mu_o, rtol=1e-11, atol=1e-9)
        assert np.allclose(e2, e2_o, rtol=1e-14, atol=0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pt_and_systems_stored_lnl(precision):
    from astroemperor_amd import PTSampler, SystemsPT
    from astroemperor_amd.simulation import test_init
    m = _model("q2_ins", k=2)
    T, W = 4, 32
    rng = np.random.default_rng(5)
    p0 = test_init(m, T, W, rng)
    s = PTSampler(W, m.ndim__, m, ntemps=T, betas=np.array([1.0, 0.5, 0.2, 0.05]), seed=9, precision=precision)
    s.run_mcmc(p0, nsweeps=5)
    st = s.get_last_sample()
    o = oracle_for(m)
    tol = (1e-10, 0.0) if precision == "fp64" else (0.0, 1e-6)
    for t, w in zip(rng.integers(0, T, 12), rng.integers(0, W, 12)):
        assert np.isclose(st.log_like[t, w], o.my_likelihood(st.coords[t, w]), rtol=tol[0], atol=tol[1]), (t, w)
    models = [make_sim("51Peg", 1, parameterisation=1).model, _model("q3_global")]
    p0s = [test_init(mm, 4, 16, rng) for mm in models]
    sy = SystemsPT(models, 4, 16, p0s, seed=3, precision=precision)
    sy.run(4)
    for i, mm in enumerate(models):
        x, ll, _ = sy.state(i)
        oo = oracle_for(mm)
        for t, w in zip(rng.integers(0, 4, 8), rng.integers(0, 16, 8)):
            assert np.isclose(ll[t, w], oo.my_likelihood(x[t, w]), rtol=tol[0], atol=tol[1]), (i, t, w)


@pytest.mark.parametrize("case,wpt", [("q2_global", 0), ("q2_ins", 1), ("q3_global", 8), ("q4_ins", 2),
                                      ("q2_ins_post", 0)])
def test_adjoint(case, wpt):
    """The order-q adjoint (reverse recurrence over companion maps) against
    FP64 autograd of the torch reference and Richardson-extrapolated central
    differences; the mixed adjoint within 1e-6 of it."""
    import torch
    from tests.test_gpu_grad import _draws as _grad_draws
    from tests.test_gpu_grad import _fd, _fd_judged
    from tests.torch_reference import lnl_torch
    m = _model(case)
    m.set_waves_per_theta(wpt)
    th = _grad_draws(m, 3, np.random.default_rng(17))
    lnl, g = m.lnl_grad_batch(th)
    assert np.allclose(lnl, m.lnl_batch(th), rtol=1e-12, atol=0)
    worst_ad, worst_fd, judged = 0.0, 0.0, 0
    for b in range(len(th)):
        t = torch.tensor(th[b], dtype=torch.float64, requires_grad=True)
        lnl_torch(m, t).backward()
        ga = t.grad.numpy()
        worst_ad = max(worst_ad, np.max(np.abs(g[b] - ga) / np.maximum(np.abs(ga), 1.0)))
        _, rich = _fd(m, th[b])
        err, n = _fd_judged(g[b], rich, lnl[b], th[b])
        worst_fd, judged = max(worst_fd, err), judged + n
    print(f"{case} wpt={wpt}: adjoint vs autograd {worst_ad:.2e}, vs Richardson FD {worst_fd:.2e} "
          f"over {judged} of {g.size}", flush=True)
    assert worst_ad < 1e-8
    # order 4 per instrument has 64 MA parameters whose derivatives the
    # differences cannot resolve against |lnL| rounding; a quarter is judged
    assert judged >= 0.25 * g.size
    assert worst_fd < 1e-6
    lm, gm = m.lnl_grad_batch(th, "mixed")
    assert np.max(np.abs(lm - lnl)) < 1e-6
    assert np.max(np.abs(gm - g) / np.maximum(np.abs(g), 1.0)) < 1e-6


def test_gradient_move_runs():
    from astroemperor_amd import GradientMove, PTSampler, StretchMove
    from astroemperor_amd.simulation import test_init
    m = _model("q2_ins")
    T, W = 2, 32
    rng = np.random.default_rng(4)
    s = PTSampler(W, m.ndim__, m, ntemps=T, seed=1, moves=[(StretchMove(), 0.5), (GradientMove(), 0.5)])
    s.run_mcmc(test_init(m, T, W, rng), nsweeps=6)
    st = s.get_last_sample()
    o = oracle_for(m)
    for t, w in zip(rng.integers(0, T, 8), rng.integers(0, W, 8)):
        assert np.isclose(st.log_like[t, w], o.my_likelihood(st.coords[t, w]), rtol=1e-10), (t, w)


def test_order_above_four_is_refused():
    from astroemperor_amd import _abi
    m = make_sim("GJ876", 1, parameterisation=1, moav={"order": 5, "global": True}).model
    with pytest.raises(_abi.EmpError, match="EUNSUP"):
        m.lnl_batch(_draws(m, 2, np.random.default_rng(0)))
    # the reference's per-instrument MA is a no-op: any order runs
    m0 = make_sim("GJ876", 1, parameterisation=1, moav={"order": 5, "global": False}).model
    _check_lnl(m0, _draws(m0, 16, np.random.default_rng(1)), "q5 no-op")
