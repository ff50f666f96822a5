"""Terms added after the moving average -- stellar-activity columns
(sai00.model), Sinusoid and MagneticCycle (sinusoid00.model,
magneticcycle00.model) -- on the GPU, against the oracle:

* lnL of the batch kernel, FP64 within 1e-10 relative, the mixed path within
  1e-6 absolute of FP64, at every launch shape, with and without the moving
  average (the terms are added after it);
* my_model per epoch;
* the adjoint gradient against FP64 autograd of the torch reference and
  Richardson-extrapolated finite differences, and the mixed adjoint against
  the FP64 one;
* the stored lnL of the parallel-tempering engine and of a batch of systems.
"""
import numpy as np
import pytest
import torch

from tests.conftest import make_sim, oracle_for
from tests.test_gpu_grad import _draws as _grad_draws
from tests.test_gpu_grad import _fd, _fd_judged
from tests.test_gpu_lnl import _check_lnl, _draws
from tests.torch_reference import lnl_torch

pytestmark = pytest.mark.gpu

CASES = {
    "sin": dict(folder="51Peg", k=1, sinusoid=1),
    "mag": dict(folder="51Peg", k=1, magnetic_cycle=1),
    "both_global_ma": dict(folder="GJ876", k=2, sinusoid=1, magnetic_cycle=1,
                           moav={"order": 1, "global": True}),
    "both_ins_ma": dict(folder="GJ876", k=1, sinusoid=1, magnetic_cycle=1,
                        moav={"order": 1, "global": False}, semantics="intended"),
    "sai": dict(folder="GJ876", k=1, switch_SA=True),
    "sai_cos_ins_ma": dict(folder="GJ876", k=2, switch_SA=True, sinusoid=1,
                           moav={"order": 1, "global": False}, semantics="intended"),
    "sai_global_ma": dict(folder="GJ876", k=1, switch_SA=True, moav={"order": 1, "global": True}),
}


def _sim(case, **kw):
    c = dict(CASES[case])
    folder, k = c.pop("folder"), c.pop("k")
    c.update(kw)
    return make_sim(folder, k, parameterisation=1, **c)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("wpt", [1, 2, 4, 8])
def test_lnl_matches_oracle(case, wpt):
    m = _sim(case).model
    m.set_waves_per_theta(wpt)
    th = _draws(m, 96, np.random.default_rng(wpt))
    e, dm = _check_lnl(m, th, f"{case} wpt{wpt}")
    print(f"{case} wpt={wpt}: fp64 rel {e:.2e}, |mixed - fp64| {dm:.2e}", flush=True)
    assert dm < 1e-6


@pytest.mark.parametrize("case", list(CASES))
def test_my_model_matches_oracle(case):
    m = _sim(case).model
    o = oracle_for(m)
    rng = np.random.default_rng(3)
    keep = [t for t in _draws(m, 60, rng) if np.isfinite(o.my_likelihood(t))][:4]
    assert keep
    for th in keep:
        mu, e2 = m.my_model(th)
        mu_o, e2_o = o.my_model(th)
        assert np.allclose(mu, mu_o, rtol=1e-11, atol=1e-9)
        assert np.allclose(e2, e2_o, rtol=1e-14, atol=0)


@pytest.mark.parametrize("case", list(CASES))
def test_adjoint(case):
    m = _sim(case).model
    th = _grad_draws(m, 4, np.random.default_rng(9))
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
    print(f"{case}: adjoint vs autograd {worst_ad:.2e}, vs Richardson FD {worst_fd:.2e} "
          f"over {judged} of {g.size}", flush=True)
    assert worst_ad < 1e-8
    assert judged >= 0.4 * g.size
    assert worst_fd < 1e-6
    lm, gm = m.lnl_grad_batch(th, "mixed")
    assert np.max(np.abs(lm - lnl)) < 1e-6
    assert np.max(np.abs(gm - g) / np.maximum(np.abs(g), 1.0)) < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pt_engine_stored_lnl(precision):
    from astroemperor_amd import PTSampler
    from astroemperor_amd.simulation import test_init
    m = _sim("sai_cos_ins_ma").model
    T, W = 4, 32
    p0 = test_init(m, T, W, np.random.default_rng(2))
    s = PTSampler(W, m.ndim__, m, ntemps=T, betas=np.array([1.0, 0.5, 0.2, 0.05]), seed=77,
                  precision=precision)
    s.run_mcmc(p0, nsweeps=6)
    st = s.get_last_sample()
    o = oracle_for(m)
    tol = (1e-10, 0.0) if precision == "fp64" else (0.0, 1e-6)
    rng = np.random.default_rng(1)
    for t, w in zip(rng.integers(0, T, 16), rng.integers(0, W, 16)):
        want = o.my_likelihood(st.coords[t, w])
        assert np.isclose(st.log_like[t, w], want, rtol=tol[0], atol=tol[1]), (t, w)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_systems_stored_lnl(precision):
    from astroemperor_amd import SystemsPT
    from astroemperor_amd.simulation import test_init
    models = [_sim("sin").model, _sim("mag").model, make_sim("51Peg", 1, parameterisation=1).model,
              _sim("sai").model]
    T, W = 4, 16
    rng = np.random.default_rng(4)
    p0 = [test_init(m, T, W, rng) for m in models]
    s = SystemsPT(models, T, W, p0, seed=6, precision=precision)
    s.run(4)
    tol = (1e-10, 0.0) if precision == "fp64" else (0.0, 1e-6)
    for i, m in enumerate(models):
        x, ll, _ = s.state(i)
        o = oracle_for(m)
        for t, w in zip(rng.integers(0, T, 8), rng.integers(0, W, 8)):
            want = o.my_likelihood(x[t, w])
            assert np.isclose(ll[t, w], want, rtol=tol[0], atol=tol[1]), (i, t, w)
