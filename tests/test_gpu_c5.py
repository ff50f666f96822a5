"""The large mixed workload on the GPU: 7 Keplerians (ln P, K, M0, e, w),
10000 epochs over 5 instruments (1000-3000 each), per-instrument offset,
jitter and MA(1) with a fixed 5 d timescale (intended semantics), 50 free
parameters, 16 temperatures x 1024 walkers, gradient (Langevin) sweeps mixed
1:4 with stretch sweeps.

* FP64 lnL against the oracle at 1e-10 relative on prior draws;
* mixed lnL against FP64 (< 1e-6 absolute) on 1024 draws and against the
  oracle on the same draws as the FP64 check;
* the MA(1) adjoint against the FP64 torch reference's autograd;
* three sweeps of the 1:4 mix at full size, in both precisions: the stored
  lnL of sampled walkers against the oracle at their stored positions.
"""
import numpy as np
import pytest
import torch

from tests.conftest import oracle_for
from tests.test_gpu_pt import M32, _philox, _u01
from tests.torch_reference import lnl_torch

pytestmark = pytest.mark.gpu

MOVE_TAG = 0x10000000          # the sweep-move draw's Philox tag


@pytest.fixture(scope="module")
def c5():
    from astroemperor_amd import synth
    w = synth.c5()
    return w["model"], oracle_for(w["model"]), w["setup"]


def _prior_draws(m, n, seed):
    from astroemperor_amd.simulation import test_init
    return test_init(m, 1, n, np.random.default_rng(seed))[0]


def test_shape(c5):
    m, _, (T, W, _, _) = c5
    assert (m.kplan__, m.nins__, m.data.ndat, m.ndim__, T, W) == (7, 5, 10000, 50, 16, 1024)


def test_fp64_and_mixed_vs_oracle(c5):
    m, o, _ = c5
    th = _prior_draws(m, 1024, 51)
    l64 = m.lnl_batch(th, "fp64")
    lmx = m.lnl_batch(th, "mixed")
    assert np.isfinite(l64).all() and np.isfinite(lmx).all()
    d = np.abs(lmx - l64)
    print(f"C5 mixed vs FP64 over 1024 draws: max {d.max():.2e}, p99 {np.percentile(d, 99):.2e}", flush=True)
    assert d.max() < 1e-6
    for b in np.linspace(0, 1023, 8).astype(int):
        want = o.my_likelihood(th[b])
        assert abs(l64[b] - want) <= 1e-10 * abs(want), (b, l64[b], want)
        assert abs(lmx[b] - want) < 1e-6, (b, lmx[b], want)


def test_ma1_adjoint_vs_torch_reference(c5):
    m, o, _ = c5
    th = _prior_draws(m, 64, 52)
    for prec, tol in (("fp64", 1e-8), ("mixed", 1e-6)):
        lnl, g = m.lnl_grad_batch(th, prec)
        assert np.isfinite(lnl).all() and np.isfinite(g).all()
        for b in (5, 40):
            t = torch.tensor(th[b], dtype=torch.float64, requires_grad=True)
            v = lnl_torch(m, t)
            v.backward()
            ga = t.grad.numpy()
            err = np.max(np.abs(g[b] - ga) / np.maximum(np.abs(ga), 1.0))
            print(f"C5 {prec} adjoint row {b}: max rel {err:.2e}", flush=True)
            assert err < tol, (prec, b, err)
            assert abs(lnl[b] - v.item()) <= (1e-10 * abs(v.item()) if prec == "fp64" else 1e-6)


def _moves_of(seed, n, ws=4.0, wg=1.0):
    """The move each sweep makes (1 = gradient), drawn as the engine draws it."""
    out = []
    for sweep in range(n):
        r = _philox([0, sweep & M32, sweep >> 32, MOVE_TAG], [seed & M32, seed >> 32])
        out.append(int(_u01(r[0], r[1]) * (ws + wg) < wg))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_size_mixed_moves_stored_lnl(c5, precision):
    from astroemperor_amd import GradientMove, PTSampler, StretchMove
    m, o, (T, W, _, _) = c5
    seed = next(s for s in range(1000, 2000) if _moves_of(s, 3) == [1, 0, 1])
    p0 = _prior_draws(m, T * W, 53).reshape(T, W, -1)
    s = PTSampler(W, m.ndim__, m, ntemps=T, seed=seed, precision=precision, record=False,
                  moves=[(StretchMove(), 4.0), (GradientMove(), 1.0)])
    s.run_mcmc(p0, nsweeps=3, capacity=3)
    st = s.get_last_sample()
    assert s.n_lnl_evaluated > 0
    moved = (st.coords != p0).any(-1)
    assert moved.any()
    rng = np.random.default_rng(7)
    picks = list(zip(rng.integers(0, T, 24), rng.integers(0, W, 24)))
    tw = np.argwhere(moved)
    picks += [tuple(tw[i]) for i in rng.choice(len(tw), 12, replace=False)]
    for t, w in picks:
        want = o.my_likelihood(st.coords[t, w])
        got = st.log_like[t, w]
        if precision == "fp64":
            assert abs(got - want) <= 1e-10 * abs(want), (t, w, got, want)
        else:
            assert abs(got - want) < 1e-6, (t, w, got, want)
        assert st.log_prior[t, w] == pytest.approx(o.my_prior(st.coords[t, w]), rel=1e-14, abs=1e-13)
    s.close()
