"""Host-side evidence estimators against closed forms (no device)."""
import numpy as np
import pytest

from astroemperor_amd.evidence import ss_log_evidence, ti_log_evidence


def test_ti_exact_for_linear_mean():
    betas = np.array([1.0, 0.6, 0.25, 0.1, 0.0])
    mean = -50.0 + 12.0 * betas                      # int_0^1 = -50 + 6
    for pchip in (False, True):
        z, err = ti_log_evidence(betas, mean, pchip=pchip)
        assert abs(z - (-44.0)) < 1e-12
        assert err < 1e-12


def test_ti_hottest_mean_stands_in_at_zero():
    betas = np.array([1.0, 0.5, 0.2])
    mean = np.array([-10.0, -12.0, -20.0])
    z, _ = ti_log_evidence(betas, mean, pchip=False)
    want = 0.2 * -20.0 + 0.5 * (-20.0 - 12.0) * 0.3 + 0.5 * (-12.0 - 10.0) * 0.5
    assert abs(z - want) < 1e-12


def test_ti_pchip_converges_on_smooth_curve():
    # <lnL>(beta) = -a / (beta + c): int_0^1 = -a ln((1 + c) / c)
    a, c = 3.0, 0.05
    want = -a * np.log((1 + c) / c)
    errs = []
    for T in (12, 24, 48):
        betas = np.concatenate([np.geomspace(1.0, 1e-4, T - 1), [0.0]])
        z, est = ti_log_evidence(betas, -a / (betas + c), pchip=True)
        errs.append(abs(z - want))
        assert abs(z - want) < max(est, 1e-3)
    assert errs[2] < errs[0]


def test_stepping_stones_gaussian_samples():
    # lnL at beta_t ~ N(mu_t, s^2): log E[exp(d X)] = d mu + d^2 s^2 / 2
    rng = np.random.default_rng(0)
    betas = np.array([1.0, 0.5, 0.2, 0.05, 0.0])
    mu = np.array([-10.0, -12.0, -16.0, -30.0, -60.0])
    s = 1.5
    lse, var = np.zeros(5), np.zeros(5)
    want = 0.0
    for t in range(1, 5):
        dlt = betas[t - 1] - betas[t]
        x = np.exp(dlt * rng.normal(mu[t], s, 400_000))
        lse[t] = np.log(x.mean())
        var[t] = (np.mean(x * x) / x.mean() ** 2 - 1) / x.size
        want += dlt * mu[t] + 0.5 * dlt * dlt * s * s
    z, err = ss_log_evidence(betas, mu, lse, var)
    assert abs(z - want) < 5 * err + 1e-9
    assert 0 < err < 0.01


def test_stepping_stones_needs_ordered_ladder():
    with pytest.raises(ValueError):
        ss_log_evidence(np.array([0.5, 1.0]), np.zeros(2), np.zeros(2), np.zeros(2))


def test_stepping_stones_adds_trapezoid_below_hottest():
    betas = np.array([1.0, 0.1])
    z, _ = ss_log_evidence(betas, np.array([-5.0, -7.0]), np.array([0.0, -3.0]), np.zeros(2))
    assert abs(z - (-3.0 + 0.1 * -7.0)) < 1e-12
