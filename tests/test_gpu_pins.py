"""Closed-form pins of two model variants on the GPU.

* ``use_c`` (support/models/clib/ckep00-04.model): the same RV term without
  the small-eccentricity guard of kep01/02/04.  Away from e = 0 it equals the
  guarded model and the oracle; for 0 < e below the guard's threshold it uses
  w = arccos(C / sqrt(e)) where the guarded model uses w = 0, which the
  closed form mu = K (cos(M + w) + e cos w) + O(e) separates; at e = 0 the
  reference's arccos(C / 0) is NaN, mapped to -inf here and counted.
* Acceleration (support/models/acc.model:2): ``np.polyval(concat(theta,
  [0]), X_ - X_[0])`` puts the first coefficient on the highest power,
  mu = a1 x^q + ... + aq x, checked against that sum written out term by
  term (not through np.polyval) for orders 1-3.
"""
import numpy as np
import pytest

from tests.conftest import make_sim, oracle_for

pytestmark = pytest.mark.gpu


def _draws(model, n, rng):
    lim = model.limits_free()
    th = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(size=(n, len(lim)))
    if model.parameterisation in (1, 2, 4):
        for k in range(model.kplan__):
            s = th[:, 5 * k + 3] ** 2 + th[:, 5 * k + 4] ** 2
            th[:, 5 * k + 3:5 * k + 5] *= np.minimum(1.0, np.sqrt(0.9 / s))[:, None]
    return th


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4])
def test_use_c_matches_oracle_and_guarded_model(p):
    rng = np.random.default_rng(300 + p)
    mc = make_sim("51Peg", 2, parameterisation=p, use_c=True).model
    mg = make_sim("51Peg", 2, parameterisation=p, use_c=False).model
    assert mc.use_c and not mg.use_c
    th = _draws(mc, 200, rng)
    o = oracle_for(mc)
    ref = np.array([o.my_likelihood(t) for t in th])
    got = mc.lnl_batch(th, "fp64")
    assert np.isfinite(ref).all()
    assert np.all(np.abs(got - ref) <= 1e-10 * np.abs(ref)), np.max(np.abs(got - ref) / np.abs(ref))
    assert np.all(np.abs(mc.lnl_batch(th, "mixed") - got) < 1e-6)
    # e well above the guard's threshold: the variants coincide
    assert np.all(np.abs(mg.lnl_batch(th, "fp64") - got) <= 1e-12 * np.abs(got))
    if p in (1, 2, 4):
        thr = 1e-6 if p == 1 else 1e-5
        x = th[:3].copy()
        # planet 1 at e = thr / 4 with S < 0, C < 0: w = 2 pi - arccos(C / sqrt(e))
        S, C = -np.sqrt(thr / 8), -np.sqrt(thr / 8)
        x[:, 3], x[:, 4] = S, C
        e = S * S + C * C
        w = 2 * np.pi - np.arccos(C / np.sqrt(e))
        mu_c, _ = mc.my_model(x[0])
        mu_g, _ = mg.my_model(x[0])
        K = x[0, 1] if p != 2 else x[0, 1] ** 2 + x[0, 2] ** 2
        # the planet's two terms differ by K (cos(M + w) - cos M) + O(K e)
        X = mc.data.X
        if p == 1:
            M = 2 * np.pi / x[0, 0] * X + x[0, 2]
        elif p == 2:
            A = K
            ph = np.arccos(x[0, 2] / np.sqrt(A))
            ph = 2 * np.pi - ph if x[0, 1] < 0 else ph
            M = 2 * np.pi / np.exp(x[0, 0]) * X + ph
        else:
            M = 2 * np.pi / x[0, 0] * (X - x[0, 2])
        want = K * (np.cos(M + w) - np.cos(M))
        assert np.max(np.abs((mu_c - mu_g) - want)) < 10 * K * e + 1e-9
        # and the oracle agrees on these rows
        ref = np.array([o.my_likelihood(t) for t in x])
        assert np.all(np.abs(mc.lnl_batch(x) - ref) <= 1e-10 * np.abs(ref))
        # e = 0 exactly: NaN in the reference, -inf here, counted
        z = x[:1].copy()
        z[:, 3] = z[:, 4] = 0.0
        assert np.isnan(o.my_likelihood(z[0]))
        n0 = mc.nonfinite_count()
        assert mc.lnl_batch(z)[0] == -np.inf
        assert mc.nonfinite_count() == n0 + 1
        assert np.isfinite(mg.lnl_batch(z)[0])


@pytest.mark.parametrize("order", [1, 2, 3])
def test_acceleration_coefficient_order(order):
    m = make_sim("51Peg", 0, acceleration=order).model
    names = m.names_free()
    assert all(n.startswith("Acceleration") for n in names[:order]) and names[order] == "Offset 1", names
    rng = np.random.default_rng(order)
    X = m.data.X
    x = X - X[0]
    o = oracle_for(m)
    for _ in range(4):
        span = x[-1]
        # each term up to 10 m/s over the span: a[i] multiplies x^(order - i)
        a = rng.uniform(-10, 10, order) / np.array([span ** (order - i) for i in range(order)])
        gam, jit = rng.uniform(-5, 5), rng.uniform(1, 5)
        th = np.concatenate([a, [gam, jit]])
        mu, err2 = m.my_model(th)
        # acc.model:2 with the first coefficient on the highest power, written out
        want = np.full_like(X, gam)
        for i in range(order):
            want += a[i] * x ** (order - i)
        assert np.max(np.abs(mu - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
        e2 = m.data.YERR ** 2 + jit ** 2
        lnl = -0.5 * np.sum((m.data.Y - want) ** 2 / e2 + np.log(e2)) - 0.5 * len(X) * np.log(2 * np.pi)
        got = m.lnl_batch(th[None])[0]
        assert abs(got - lnl) <= 1e-12 * abs(lnl)
        assert abs(o.my_likelihood(th) - lnl) <= 1e-12 * abs(lnl)
        if order >= 2:
            # the reversed order (first coefficient on x) is a different model
            rev = np.full_like(X, gam)
            for i in range(order):
                rev += a[i] * x ** (i + 1)
            assert np.max(np.abs(rev - want)) > 1e3 * np.max(np.abs(mu - want)) + 1e-9
