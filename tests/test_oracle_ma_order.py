"""The oracle's MA(q) recurrence for q >= 2 (moav01.model:7-15, the literal
loop in ``ma_residuals``) against an independent statement of it: the
in-place update makes eps the solution of the unit lower-triangular banded
system L eps = r with L[i, i-1-c] = phi_c exp(-|t_i - t_{i-1-c}| / tau_c).
Using the original residuals instead of the updated ones (a plausible slip)
would be the product (2I - L) r instead, and fails this test."""
import numpy as np
import pytest
from scipy.linalg import solve_triangular

from oracle.emp_oracle import ma_residuals


@pytest.mark.parametrize("q", [1, 2, 3, 4])
def test_recurrence_is_the_triangular_solve(q):
    rng = np.random.default_rng(q)
    n = 200
    t = np.sort(rng.uniform(0, 300, n))
    r = rng.normal(0, 5, n)
    coefs = rng.uniform(0.05, 0.3, q)
    taus = rng.uniform(5, 25, q)
    eps, added = ma_residuals(t, r, list(coefs), list(taus))
    L = np.eye(n)
    for i in range(n):
        for c in range(q):
            if i > c:
                L[i, i - 1 - c] = coefs[c] * np.exp(-abs(t[i] - t[i - 1 - c]) / taus[c])
    want = solve_triangular(L, r, lower=True, unit_diagonal=True)
    assert np.allclose(eps, want, rtol=1e-12, atol=1e-12)
    assert np.allclose(added, r - want, rtol=1e-12, atol=1e-12)
    if q > 1:
        wrong = (2 * np.eye(n) - L) @ r
        assert not np.allclose(eps, wrong, rtol=1e-6)
