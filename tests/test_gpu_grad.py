"""Adjoint-kernel gradients against an FP64 autograd reference, FP64 central
finite differences (h = 1e-6 max(|theta|, 1), Richardson-extrapolated), the
torch.autograd.Function wrapper and the batched L-BFGS MAP search."""
import numpy as np
import pytest
import torch

from tests.conftest import make_sim, oracle_for
from tests.torch_reference import lnl_torch

pytestmark = pytest.mark.gpu


def _draws(model, n, rng, shrink=0.6):
    lim = model.limits_free()
    mid, half = lim.mean(1), (lim[:, 1] - lim[:, 0]) / 2 * shrink
    th = mid + half * rng.uniform(-1, 1, size=(n, len(mid)))
    if model.parameterisation in (1, 2, 4, 7):        # keep e = S^2 + C^2 < 0.8
        for k in range(model.kplan__):
            s = th[:, 5 * k + 3] ** 2 + th[:, 5 * k + 4] ** 2
            th[:, 5 * k + 3:5 * k + 5] *= np.minimum(1.0, np.sqrt(0.8 / s))[:, None]
    return th


def _fd(model, th, h_rel=1e-6):
    """Central differences of the GPU FP64 lnL, plain and Richardson."""
    D = len(th)
    h = h_rel * np.maximum(np.abs(th), 1.0)
    rows = []
    for j in range(D):
        for s in (h[j], -h[j], h[j] / 2, -h[j] / 2):
            t = th.copy()
            t[j] += s
            rows.append(t)
    v = model.lnl_batch(np.array(rows)).reshape(D, 4)
    plain = (v[:, 0] - v[:, 1]) / (2 * h)
    half = (v[:, 2] - v[:, 3]) / h
    return plain, (4 * half - plain) / 3


def _fd_judged(g, rich, lnl, th, h_rel=1e-6):
    """Relative FD mismatch over the entries the differences can judge.

    Each lnL carries a rounding error of a few 2^-52 |lnL|, which the central
    difference divides by h: an entry is judged where 4 * 2^-52 |lnL| / h is at
    most 1e-7 max(|fd|, 1).  Returns (largest relative mismatch over judged
    entries, their number)."""
    h = h_rel * np.maximum(np.abs(th), 1.0)
    scale = np.maximum(np.abs(rich), 1.0)
    judged = 4 * 2.0 ** -52 * abs(lnl) / h <= 1e-7 * scale
    err = np.abs(g - rich) / scale
    return (float(np.max(err[judged])) if judged.any() else 0.0), int(judged.sum())


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 6, 7])
def test_adjoint_vs_autograd_and_fd(p):
    rng = np.random.default_rng(50 + p)
    sim = make_sim("51Peg", 2, parameterisation=p, acceleration=1 if p == 0 else 0)
    m = sim.model
    th = _draws(m, 6, rng)
    lnl, g = m.lnl_grad_batch(th)
    ref_l = m.lnl_batch(th)
    assert np.allclose(lnl, ref_l, rtol=1e-12, atol=0)
    worst_ad, worst_fd = 0.0, 0.0
    for b in range(len(th)):
        t = torch.tensor(th[b], dtype=torch.float64, requires_grad=True)
        v = lnl_torch(m, t)
        v.backward()
        ga = t.grad.numpy()
        scale = np.maximum(np.abs(ga), 1.0)
        worst_ad = max(worst_ad, np.max(np.abs(g[b] - ga) / scale))
        plain, rich = _fd(m, th[b])
        worst_fd = max(worst_fd, np.max(np.abs(g[b] - rich) / np.maximum(np.abs(rich), 1.0)))
    print(f"p={p}: adjoint vs autograd {worst_ad:.2e}, vs Richardson FD {worst_fd:.2e}", flush=True)
    assert worst_ad < 1e-8
    assert worst_fd < 1e-6


def test_adjoint_c4_shape():
    """5 planets, 2 instruments, 2000 epochs: batch gradient vs autograd on a few rows."""
    from astroemperor_amd import synth
    w = synth.c4()
    m = w["model"]
    rng = np.random.default_rng(8)
    th = _draws(m, 1024, rng)
    lnl, g = m.lnl_grad_batch(th)
    assert np.isfinite(lnl).all() and np.isfinite(g).all()
    for b in rng.choice(1024, 3, replace=False):
        t = torch.tensor(th[b], dtype=torch.float64, requires_grad=True)
        v = lnl_torch(m, t)
        v.backward()
        assert np.max(np.abs(g[b] - t.grad.numpy()) / np.maximum(np.abs(t.grad.numpy()), 1.0)) < 1e-8
        assert abs(v.item() - lnl[b]) < 1e-10 * abs(lnl[b])


def test_autograd_function_gradcheck():
    from astroemperor_amd.autograd import LogLikelihood
    sim = make_sim("51Peg", 1, parameterisation=1)
    m = sim.model
    th = torch.tensor(_draws(m, 3, np.random.default_rng(1)), dtype=torch.float64, device="cuda",
                      requires_grad=True)
    out = LogLikelihood.apply(th, m)
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, device="cuda")
    (out * w).sum().backward()
    _, g = m.lnl_grad_batch(th.detach().cpu().numpy())
    assert np.allclose(th.grad.cpu().numpy(), w.cpu().numpy()[:, None] * g, rtol=1e-14, atol=0)


def test_map_search_51peg_reaches_kat():
    from astroemperor_amd.autograd import map_search
    sim = make_sim("51Peg", 1, parameterisation=1,
                   conditions=[("Period 1", "limits", [4.225, 4.236]), ("Amplitude 1", "limits", [45, 65]),
                               ("Offset 1", "limits", [-10.0, 10.0])])
    m = sim.model
    rng = np.random.default_rng(0)
    starts = _draws(m, 256, rng, shrink=0.9)
    th, post = map_search(m, starts, iters=150)
    best = th[np.nanargmax(post)]
    o = oracle_for(m)
    # the max-posterior lnL sits at the recorded KAT (-869.480) within the sampler's spread
    assert o.my_likelihood(best) > -869.55, (o.my_likelihood(best), best)
    assert np.nanmax(post) >= np.nanmax([o.my_prior(s) + o.my_likelihood(s) for s in starts])


@pytest.mark.parametrize("case", ["per_instrument", "per_instrument_fixed_tau", "global"])
def test_adjoint_ma1(case):
    """MA(1) adjoint (reverse recurrence of the residual filter): against the
    torch reference's autograd and Richardson-extrapolated central differences,
    with free timescales, fixed ones (precomputed weights) and the global series."""
    rng = np.random.default_rng({"per_instrument": 1, "per_instrument_fixed_tau": 2, "global": 3}[case])
    conds = []
    if case == "per_instrument_fixed_tau":
        conds = [(f"MATimescale {i} Order 1", "fixed", 5.0) for i in range(1, 9)]
    sim = make_sim("GJ876", 1, parameterisation=1, moav={"order": 1, "global": case == "global"},
                   semantics="intended", conditions=conds)
    m = sim.model
    # waves per vector: 1 (no cross-wave carry), 8, and the default (2 at 752 epochs)
    if case == "global":
        m.set_waves_per_theta(1)
    elif case == "per_instrument_fixed_tau":
        m.set_waves_per_theta(8)
    th = _draws(m, 4, rng)
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
          f"over {judged} of {g.size} entries", flush=True)
    assert worst_ad < 1e-8
    assert judged >= 0.4 * g.size
    assert worst_fd < 1e-6


@pytest.mark.parametrize("ma", [False, True])
def test_mixed_adjoint_matches_fp64(ma):
    """The mixed path's adjoint (FP32 Kepler start, one FP64 step; table and
    Markley starts both hit, e up to 0.8 at 752 epochs) against the FP64
    adjoint: lnL within 1e-6 absolute, gradient within 1e-6 of max(|g|, 1)."""
    kw = {"moav": {"order": 1, "global": False}, "semantics": "intended"} if ma else {}
    sim = make_sim("GJ876", 2, parameterisation=1, **kw)
    m = sim.model
    th = _draws(m, 96, np.random.default_rng(11 + ma))
    l64, g64 = m.lnl_grad_batch(th)
    lm, gm = m.lnl_grad_batch(th, "mixed")
    assert np.allclose(lm, m.lnl_batch(th, "mixed"), rtol=1e-12, atol=0)
    dl = np.max(np.abs(lm - l64))
    dg = np.max(np.abs(gm - g64) / np.maximum(np.abs(g64), 1.0))
    print(f"ma={ma}: |dlnL| {dl:.2e}, gradient {dg:.2e}", flush=True)
    assert dl < 1e-6
    assert dg < 1e-6


def test_map_init_seeds_walkers_at_the_mode():
    """Walkers seeded around the batched L-BFGS optimum (autograd.map_init):
    the mode's lnL reaches the recorded known answer (51 Peg, one planet,
    parameterisation 1: lnL -869.480 at the max-posterior sample,
    tests/01_51peg_basic.ipynb and 00_mini_test.ipynb); with a small spread
    every cold walker starts above it; with the Laplace spread the cold
    walkers' lnL sit about D/2 below the mode, as a Gaussian posterior puts
    them; every walker is inside the prior; and the sampler runs from there."""
    from astroemperor_amd import PTSampler
    from astroemperor_amd.autograd import map_init
    sim = make_sim("51Peg", 1, parameterisation=1,
                   conditions=[("Period 1", "limits", [4.225, 4.236]), ("Amplitude 1", "limits", [45, 65]),
                               ("Offset 1", "limits", [-10.0, 10.0])])
    m = sim.model
    o = oracle_for(m)
    T, W, D = 4, 64, m.ndim__
    betas = np.array([1.0, 0.5, 0.2, 0.05])
    tight, mode, lpost = map_init(m, T, W, np.random.default_rng(2), betas=betas, starts=256, iters=150,
                                  spread=0.01)
    l_mode = o.my_likelihood(mode)
    assert l_mode >= -869.48, l_mode
    assert abs(lpost - (l_mode + o.my_prior(mode))) < 1e-8 * abs(lpost)
    cold = np.array([o.my_likelihood(x) for x in tight[0]])
    assert np.all(cold >= -869.48), cold.min()
    pos, _, _ = map_init(m, T, W, np.random.default_rng(3), betas=betas, starts=256, iters=150)
    assert np.isfinite(m.lnprior_batch(pos.reshape(-1, D))).all()
    drop = l_mode - np.array([o.my_likelihood(x) for x in pos[0]])
    print(f"mode lnL {l_mode:.4f}, Laplace-spread cold walkers: mean drop {drop.mean():.2f} (D/2 = {D / 2})",
          flush=True)
    assert 0.4 * D / 2 < drop.mean() < 2.5 * D / 2
    hot = l_mode - np.array([o.my_likelihood(x) for x in pos[-1]])
    assert hot.mean() > drop.mean()            # hotter temperatures spread wider
    s = PTSampler(W, D, m, ntemps=T, betas=betas, adapt_mode=-1, seed=4)
    s.run_mcmc(pos, nsweeps=50)
    assert np.median(s.get_log_like()[0, -1]) > l_mode - 3 * D
