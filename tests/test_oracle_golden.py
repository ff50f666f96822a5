"""Pins the oracle against 50-digit mpmath vectors and closed forms (CPU)."""
import json
import os

import numpy as np
import pytest

from oracle import emp_oracle as orc


@pytest.fixture(scope="module")
def kep(golden_dir):
    return np.load(os.path.join(golden_dir, "kepler_mpmath.npz"))


@pytest.fixture(scope="module")
def lnl_gold(golden_dir):
    with open(os.path.join(golden_dir, "lnl_mpmath.json")) as f:
        return json.load(f)


def _tol(kep):
    """2 ulp of the larger of |M| and E (E in [0, 2 pi)), amplified by the
    conditioning 1/(1 - e) of the root; |M| up to 12566 rad carries its own
    representation error into the reduction."""
    scale = np.maximum(np.maximum(np.abs(kep["M"]), np.abs(kep["E"])), 1.0)
    return 2 * np.spacing(scale) / (1 - kep["e"])


def _angle_err(a, b):
    d = np.mod(a - b + np.pi, 2 * np.pi) - np.pi
    return np.abs(d)


def test_kepler_c_matches_mpmath(kep):
    E = orc.kepler_solve(kep["M"], kep["e"])
    err = _angle_err(E, kep["E"])
    assert np.all(err <= _tol(kep)), (err / _tol(kep)).max()


def test_kepler_numpy_fallback_matches_mpmath(kep):
    E = orc._kepler_solve_numpy(kep["M"], kep["e"])
    err = _angle_err(E, kep["E"])
    assert np.all(err <= _tol(kep))


def test_kepler_identities():
    e = np.array([0.0, 0.1, 0.5, 0.9, 0.99])
    assert np.all(orc.kepler_solve(np.zeros(5), e) == 0.0)
    assert np.allclose(orc.kepler_solve(np.full(5, np.pi), e), np.pi, atol=1e-15)
    M = np.linspace(0.01, 3.1, 50)
    for ee in e:
        Ep = orc.kepler_solve(M, ee)
        Em = orc.kepler_solve(-M, ee)
        assert np.allclose(np.mod(-Em, 2 * np.pi), Ep, atol=2e-15)
        # residual of Kepler's equation
        assert np.max(np.abs(Ep - ee * np.sin(Ep) - M)) < 4e-16 * 8
    # e = 0: E = M
    M = np.linspace(-10, 10, 101)
    assert np.allclose(orc.kepler_solve(M, 0.0), np.mod(M, 2 * np.pi), atol=1e-14)


def _spec_for(case, nins):
    """Minimal spec: k Keplerians, Offset, Jitter blocks (no priors needed)."""
    k = len(case["planets"])
    blocks = [{"type": "Keplerian", "start": 5 * i, "len": 5, "parameterisation": case["parameterisation"]}
              for i in range(k)]
    blocks += [{"type": "Offset", "start": 5 * k, "len": nins},
               {"type": "Jitter", "start": 5 * k + nins, "len": nins}]
    params = [{"name": f"p{i}", "prior": "Uniform", "limits": [-1e9, 1e9], "prargs": 0.0, "fixed": None}
              for i in range(5 * k + 2 * nins)]
    return {"blocks": blocks, "params": params, "nins": nins, "semantics": "reference"}


def test_oracle_lnl_matches_mpmath(lnl_gold):
    for case in lnl_gold["cases"]:
        ds = lnl_gold["datasets"][case["data"]]
        spec = _spec_for(case, ds["nins"])
        o = orc.OracleModel(spec, ds["X"], ds["Y"], ds["YERR"], ds["FLAG"])
        theta = sum(case["planets"], []) + case["offsets"] + case["jitters"]
        mu, _ = o.my_model(theta)
        assert np.allclose(mu, case["mu"], rtol=0, atol=1e-9), case["parameterisation"]
        ll = o.my_likelihood(theta)
        assert abs(ll - case["lnl"]) <= 1e-10 * abs(case["lnl"]), (case["parameterisation"], ll, case["lnl"])


def test_keplerian_closed_forms():
    X = np.linspace(0, 100, 33)
    # below the omega threshold (p1): w = 0 so mu = K cos(f) + K e, and with e
    # tiny f ~ M: mu ~ K cos(M)
    th = [7.0, 10.0, 0.3, 1e-4, 1e-4]   # e = 2e-8 < 1e-6
    mu = orc.keplerian_term(th, X, 1)
    M = 2 * np.pi / 7.0 * X + 0.3
    assert np.allclose(mu, 10.0 * np.cos(M), atol=1e-6)
    # at M = 0 (periastron) mu = K (1 + e) cos w; at M = pi mu = K (e - 1) cos w
    for e in [0.1, 0.5, 0.8]:
        w = 0.7
        mu0 = orc.keplerian_term([5.0, 3.0, 0.0, e, w], np.array([0.0]), 0)[0]
        assert abs(mu0 - 3.0 * (1 + e) * np.cos(w)) < 1e-12
        mupi = orc.keplerian_term([5.0, 3.0, np.pi, e, w], np.array([0.0]), 0)[0]
        assert abs(mupi - 3.0 * (e - 1) * np.cos(w)) < 1e-12
    # use_c variant has no small-e guard: NaN at e = 0 (clib/ckep01.model:3)
    assert np.all(np.isnan(orc.keplerian_term([5.0, 3.0, 0.1, 0.0, 0.0], X, 1, use_c=True)))


def test_ma_literal_loop():
    """moav01 recurrence against a hand-unrolled 4-epoch example."""
    t = np.array([0.0, 1.0, 3.0, 3.5])
    r = np.array([1.0, -2.0, 0.5, 4.0])
    c, tau = 0.3, 2.0
    e0 = r[0]
    e1 = r[1] - c * np.exp(-1.0 / tau) * e0
    e2 = r[2] - c * np.exp(-2.0 / tau) * e1
    e3 = r[3] - c * np.exp(-0.5 / tau) * e2
    out, added = orc.ma_residuals(t, r, [c], [tau])
    assert np.allclose(out, [e0, e1, e2, e3], rtol=0, atol=1e-15)
    assert np.allclose(r - out, added, atol=1e-15)


def test_priors_closed_forms():
    from scipy.stats import beta, truncnorm
    lim = [1e-5, 40.0]
    mu, s = 5.0, 5.0
    args = orc.normalise_prargs("Normal", lim, [mu, s])
    for x in [1e-5, 0.3, 5.0, 17.0, 40.0]:
        a, b = (lim[0] - mu) / s, (lim[1] - mu) / s
        assert abs(orc.prior_normal(x, lim, args) - truncnorm.logpdf(x, a, b, loc=mu, scale=s)) < 1e-12
    assert orc.prior_normal(41.0, lim, args) == -np.inf
    assert orc.prior_uniform(0.5, [0, 2], np.log(0.5)) == np.log(0.5)
    assert orc.prior_uniform(2.5, [0, 2], np.log(0.5)) == -np.inf
    # the reference's Jeffreys prior is uniform in x (Jeffreys.prior:5)
    assert orc.prior_jeffreys(3.0, [1.0, 5.0]) == np.log(1 / 4.0)
    for x in [0.01, 0.3, 0.9]:
        assert abs(orc.prior_beta(x, [0, 1], [0.867, 3.03]) - beta.logpdf(x, 0.867, 3.03)) < 1e-12


def test_ladder_update_fixed_points():
    betas = np.array([1.0, 0.5, 0.25, 0.0])
    # equal swap rates: the ladder does not move
    assert np.allclose(orc.ladder_update(betas, [0.3, 0.3, 0.3], 10, 1000, 1), betas)
    # a colder pair accepting more than the next pushes T_1 up (beta_1 down)
    nb = orc.ladder_update(betas, [0.6, 0.2, 0.2], 0, 1000, 1)
    assert nb[1] < 0.5 and nb[0] == 1.0 and nb[-1] == 0.0


def test_ladder_modes_swap_distance():
    betas = np.array([1.0, 0.5, 0.25, 0.0])
    # equal mean |log swap ratio|: no motion; zero distances: no motion, no NaN
    for d in ([0.4, 0.4, 0.4], [0.0, 0.0, 0.0]):
        assert np.array_equal(orc.ladder_update(betas, None, 3, 1000, 1, mode=1, smd=d), betas)
    # pair (0, 1) closer than pair (1, 2): its spacing widens, beta_1 falls
    nb = orc.ladder_update(betas, None, 0, 1000, 1, mode=1, smd=[0.1, 0.3, 0.3])
    assert nb[1] < 0.5 and nb[0] == 1.0 and nb[-1] == 0.0
    # the balance term is (d_1 - d_0) / (d_1 + d_0) = 0.5 at kappa = 1
    assert np.isclose(1.0 / nb[1], 1.0 + np.exp(0.5), rtol=1e-15)
    assert np.allclose(orc.ladder_update(betas, None, 0, 1000, 1, mode=-1, smd=[0.1, 0.3, 0.3]), betas)


def test_ladder_mode_specific_heat_fixed_point():
    # A Gaussian likelihood in d dimensions has sd(lnL) = sqrt(d / 2) / beta at
    # every temperature; then every pair of a geometric ladder has the same
    # thermodynamic length and the ladder stays.
    geo = np.geomspace(1.0, 1e-4, 6)
    sig = np.sqrt(1.5) / geo
    nb = orc.ladder_update(geo, None, 0, 1000, 1, mode=2, spread=sig)
    assert np.allclose(nb, geo, rtol=1e-13, atol=0)
    # A perturbed ladder moves back to the geometric one under the same law
    # (a flipped balance sign would run away from it).
    b = np.array([1.0, 0.9, 0.5, 0.05, 0.002, 1e-4])
    for n in range(400):
        b = orc.ladder_update(b, None, n, 50, 1, mode=2, spread=np.sqrt(1.5) / b)
    assert np.allclose(np.log(b), np.log(geo), rtol=0, atol=1e-3)
    # known answer at T = 3: lengths 0.5 * (1 + 2) / 2 = 0.75 and
    # 0.4 * (2 + 10) / 2 = 2.4, balance (2.4 - 0.75) / 3.15, kappa = 1
    nb = orc.ladder_update([1.0, 0.5, 0.1], None, 0, 1000, 1, mode=2, spread=[1.0, 2.0, 10.0])
    assert np.isclose(1.0 / nb[1], 1.0 + np.exp(1.65 / 3.15), rtol=1e-15)


def test_lnl_spread():
    rng = np.random.default_rng(0)
    L = rng.normal(-100, 7, size=(4, 33))
    L[1, 5] = -np.inf
    L[2, :] = -np.inf
    L[2, 3] = -5.0
    got = orc.lnl_spread(L)
    assert np.isclose(got[0], np.std(L[0], ddof=1), rtol=1e-14)
    assert np.isclose(got[1], np.std(np.delete(L[1], 5), ddof=1), rtol=1e-14)
    assert got[2] == 0.0
    assert np.isclose(got[3], np.std(L[3], ddof=1), rtol=1e-14)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_oracle_acceleration_first_coefficient_on_highest_power(order):
    """acc.model:2, np.polyval(concat(theta, [0]), X_ - X_[0]): the oracle's
    model is a1 x^q + ... + aq x + gamma, the sum written out term by term."""
    from tests.conftest import make_sim, oracle_for
    m = make_sim("51Peg", 0, acceleration=order).model
    o = oracle_for(m)
    x = m.data.X - m.data.X[0]
    rng = np.random.default_rng(order)
    a = rng.uniform(-10, 10, order) / np.array([x[-1] ** (order - i) for i in range(order)])
    mu, _ = o.my_model(np.concatenate([a, [0.5, 2.0]]))
    want = 0.5 + sum(a[i] * x ** (order - i) for i in range(order))
    assert np.max(np.abs(mu - want)) <= 1e-13 * np.max(np.abs(want))
    if order > 1:
        rev = 0.5 + sum(a[i] * x ** (i + 1) for i in range(order))
        assert np.max(np.abs(mu - rev)) > 1.0
