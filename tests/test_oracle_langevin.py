"""The oracle's gradient move (ensemble-preconditioned Metropolis-adjusted
Langevin) against what the mathematics fixes:

* the factor L L^T is the complementary walkers' population covariance with
  the diagonal times 1 + 1e-6;
* log q(y | x) - log q(x | y) equals the difference of scipy's multivariate
  normal densities N(y; x + (eps^2/2) beta Sigma g(x), eps^2 Sigma) and back;
* on correlated Gaussian targets the move leaves N(m, Sigma / beta) invariant.
  The step is large (eps = 1.5), where the unadjusted Langevin chain's variance
  is more than twice the target's, so a dropped or inverted proposal ratio
  fails the moment check.
"""
import numpy as np
from scipy import stats

from oracle import emp_oracle as orc


def _spd(rng, D):
    A = rng.normal(size=(D, D))
    return A @ A.T + D * np.eye(D) * 0.3


def test_factor_is_ridged_population_covariance():
    rng = np.random.default_rng(0)
    C = rng.multivariate_normal(np.zeros(4), _spd(rng, 4), size=50)
    L, act = orc.langevin_factor(C)
    S = np.cov(C.T, bias=True)
    S[np.diag_indices(4)] *= 1 + 1e-6
    assert act.all()
    assert np.allclose(L @ L.T, S, rtol=1e-12, atol=1e-12)
    assert np.allclose(np.triu(L, 1), 0.0)


def test_factor_parameter_without_spread_is_inactive():
    rng = np.random.default_rng(1)
    C = rng.normal(size=(20, 3))
    C[:, 1] = 7.0
    L, act = orc.langevin_factor(C)
    assert list(act) == [True, False, True]
    assert L[1, 1] == 1.0 and L[1, 0] == 0.0 and L[2, 1] == 0.0
    # the inactive parameter never moves (all proposals accepted: u ~ 0)
    X = rng.normal(size=(4, 3))
    X[:, 1] = 7.0

    class P:
        prior = staticmethod(lambda th: 0.0)
        like = staticmethod(lambda th: -0.5 * float(th @ th))
        grad = staticmethod(lambda th: -th)

    Xc = np.vstack([X, C])
    lnl = np.array([P.like(x) for x in Xc])
    g = -Xc.copy()
    before = Xc[:4, 1].copy()
    orc.langevin_halfstep(Xc, lnl, np.zeros(len(Xc)), g, 1.0, [0, 1, 2, 3], list(range(4, len(Xc))),
                          rng.normal(size=(4, 3)), np.full(4, 1e-300), P, 0.5)
    assert np.array_equal(Xc[:4, 1], before)


def test_proposal_ratio_matches_multivariate_normal():
    rng = np.random.default_rng(2)
    D, eps, beta = 5, 0.7, 0.4
    C = rng.multivariate_normal(np.zeros(D), _spd(rng, D), size=40)
    L, act = orc.langevin_factor(C)
    Sig = L @ L.T
    for _ in range(20):
        x, y = rng.normal(size=D) * 2, rng.normal(size=D) * 2
        gx, gy = rng.normal(size=D) * 3, rng.normal(size=D) * 3
        mine = orc.langevin_logq(x, y, gx, L, act, beta, eps) - orc.langevin_logq(y, x, gy, L, act, beta, eps)
        fwd = stats.multivariate_normal(x + 0.5 * eps ** 2 * beta * Sig @ gx, eps ** 2 * Sig).logpdf(y)
        rev = stats.multivariate_normal(y + 0.5 * eps ** 2 * beta * Sig @ gy, eps ** 2 * Sig).logpdf(x)
        assert abs(mine - (fwd - rev)) < 1e-9 * max(1.0, abs(fwd - rev))


def _run_gaussian(beta, eps, sweeps, seed):
    rng = np.random.default_rng(seed)
    D, W = 3, 40
    m = np.array([1.0, -2.0, 0.5])
    Sig = np.array([[2.0, 1.2, 0.3], [1.2, 1.5, -0.4], [0.3, -0.4, 0.8]])
    P_ = np.linalg.inv(Sig)

    class Post:
        prior = staticmethod(lambda th: 0.0 if np.all(np.abs(th) < 1e3) else -np.inf)
        like = staticmethod(lambda th: -0.5 * float((th - m) @ P_ @ (th - m)))
        grad = staticmethod(lambda th: -P_ @ (th - m))

    # started at stationarity: from a spread several times wider, the ensemble
    # covariance makes every drift overshoot and nothing is accepted (the
    # stretch moves contract such an ensemble when the two are mixed)
    X = rng.multivariate_normal(m, Sig / beta, size=W)
    lnl = np.array([Post.like(x) for x in X])
    lnp = np.zeros(W)
    g = np.array([Post.grad(x) for x in X])
    halves = [list(range(W // 2)), list(range(W // 2, W))]
    out, nacc = [], 0
    for s in range(sweeps):
        for h in (0, 1):
            acc = orc.langevin_halfstep(X, lnl, lnp, g, beta, halves[h], halves[1 - h],
                                        rng.normal(size=(W // 2, D)), rng.uniform(size=W // 2), Post, eps)
            nacc += acc.sum()
        if s >= 200:
            out.append(X.copy())
    x = np.concatenate(out)
    return x, m, Sig / beta, nacc / (sweeps * W)


def test_gaussian_stationary_beta_1():
    x, m, S, af = _run_gaussian(1.0, 1.5, 1400, 3)
    assert 0.15 < af < 0.95
    assert np.all(np.abs(x.mean(0) - m) < 6 * np.sqrt(np.diag(S) / 4000))
    assert np.allclose(np.cov(x.T), S, rtol=0.1, atol=0.08 * np.sqrt(np.outer(np.diag(S), np.diag(S))).max())


def test_gaussian_stationary_tempered():
    x, m, S, af = _run_gaussian(0.3, 1.5, 1400, 4)
    assert 0.15 < af < 0.95
    assert np.all(np.abs(np.diag(np.cov(x.T)) / np.diag(S) - 1) < 0.1)
