"""Moving average of order q = 2..4 (moav00/moav01 with q > 1) on the GPU
against the oracle: the batched lnL at every launch shape (FP64 within 1e-10
relative, mixed within 1e-6 absolute of FP64), my_model, the PT engine's and
the batched-systems engine's stored lnL; the adjoint and the gradient move
refuse these models (EMP_EUNSUP), as do orders above 4."""
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


def test_adjoint_and_gradient_move_refuse():
    from astroemperor_amd import GradientMove, PTSampler, StretchMove, _abi
    m = _model("q2_global")
    th = _draws(m, 4, np.random.default_rng(0))
    with pytest.raises(_abi.EmpError, match="EUNSUP"):
        m.lnl_grad_batch(th)
    from astroemperor_amd.simulation import test_init
    s = PTSampler(16, m.ndim__, m, ntemps=2, seed=1, moves=[(StretchMove(), 0.8), (GradientMove(), 0.2)])
    with pytest.raises(_abi.EmpError, match="EUNSUP"):
        s.run_mcmc(test_init(m, 2, 16, np.random.default_rng(2)), nsweeps=1)


def test_order_above_four_is_refused():
    from astroemperor_amd import _abi
    m = make_sim("GJ876", 1, parameterisation=1, moav={"order": 5, "global": True}).model
    with pytest.raises(_abi.EmpError, match="EUNSUP"):
        m.lnl_batch(_draws(m, 2, np.random.default_rng(0)))
    # the reference's per-instrument MA is a no-op: any order runs
    m0 = make_sim("GJ876", 1, parameterisation=1, moav={"order": 5, "global": False}).model
    _check_lnl(m0, _draws(m0, 16, np.random.default_rng(1)), "q5 no-op")
