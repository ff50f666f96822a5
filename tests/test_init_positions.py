"""Initial positions (emp/emp.py:617-682), on the CPU with the oracle's prior.

The generated ``set_init`` draws, per temperature and free parameter,
``r (2 sort(U(0, 1, W)) - 1) + m`` and shuffles it: W iid uniforms on
[m - r, m + r], m and r the centre and half-width of ``init_pos`` (rounded to
8 decimals) or of the limits, r shrunk by 0.707 for the ``is_hou``
parameters (Amp_sin/Amp_cos/Ecc_sin/Ecc_cos, param_repo.py:86-111), which
keeps S^2 + C^2 <= 2 (0.707)^2 < 1.  ``test_init`` redraws every walker whose
prior is -inf, up to 100 rounds, then prints 'COULDNT FIND VALID INITIAL
POSITION' and returns what it has.
"""
import numpy as np
import pytest
from scipy import stats

from tests.conftest import make_sim, oracle_for


def _with_oracle_prior(model):
    o = oracle_for(model)
    model.lnprior_batch = lambda th: np.array([o.my_prior(t) for t in np.atleast_2d(th)])
    return model, o


def test_set_init_spread_and_hou_shrink():
    from astroemperor_amd.simulation import set_init
    m = make_sim("51Peg", 1, parameterisation=1).model
    free = [p for p in m.params() if p.fixed is None]
    assert [p.name for p in free if p.is_hou] == ["Ecc_sin 1", "Ecc_cos 1"]
    T, W = 3, 3000
    pos = set_init(m, T, W, np.random.default_rng(0))
    assert pos.shape == (T, W, len(free))
    for j, p in enumerate(free):
        lo, hi = p.limits
        mid, r = (lo + hi) / 2, (hi - lo) / 2 * (0.707 if p.is_hou else 1.0)
        v = pos[:, :, j]
        assert v.min() >= mid - r and v.max() <= mid + r, p.name
        # the extremes of 3000 uniforms sit within 0.5% of the interval's edges
        assert v.min() < mid - 0.99 * r and v.max() > mid + 0.99 * r, p.name
        for t in range(T):
            u = (v[t] - (mid - r)) / (2 * r)
            assert stats.kstest(u, "uniform").pvalue > 1e-4, (p.name, t)
        # shuffled independently per temperature (not the same permutation)
        assert not np.array_equal(np.argsort(v[0]), np.argsort(v[1]))
    # the shrink keeps every eccentricity inside the unit disk
    e = pos[..., 3] ** 2 + pos[..., 4] ** 2
    assert e.max() < 1.0 and e.max() > 0.9


def test_set_init_uses_rounded_init_pos():
    from astroemperor_amd.simulation import set_init
    m = make_sim("51Peg", 1, parameterisation=1).model
    free = [p for p in m.params() if p.fixed is None]
    free[0].init_pos = [4.230000004, 4.240000006]          # rounded to 8 decimals: 4.23, 4.24000001
    pos = set_init(m, 2, 2000, np.random.default_rng(1))
    v = pos[:, :, 0]
    assert v.min() >= 4.23 and v.max() <= 4.24000001
    assert v.min() < 4.2301 and v.max() > 4.2399


def test_test_init_redraws_invalid_walkers():
    """init_pos of the period half outside the prior box: about half the
    walkers start with a -inf prior and are redrawn until every walker is
    valid; the redrawn values stay uniform on the valid part."""
    from astroemperor_amd.simulation import test_init
    m, o = _with_oracle_prior(make_sim("51Peg", 1, parameterisation=1).model)
    free = [p for p in m.params() if p.fixed is None]
    lo, hi = free[0].limits
    free[0].init_pos = [lo - (hi - lo), hi]                # the lower half lies below the limit
    T, W = 2, 400
    p0 = test_init(m, T, W, np.random.default_rng(2))
    lp = np.array([[o.my_prior(x) for x in p0[t]] for t in range(T)])
    assert np.isfinite(lp).all()
    v = p0[:, :, 0].ravel()
    assert v.min() >= lo and v.max() <= hi
    assert stats.kstest((v - lo) / (hi - lo), "uniform").pvalue > 1e-4


def test_test_init_gives_up_after_max_repeats(capsys):
    from astroemperor_amd.simulation import test_init
    m, o = _with_oracle_prior(make_sim("51Peg", 1, parameterisation=1).model)
    free = [p for p in m.params() if p.fixed is None]
    lo, hi = free[0].limits
    free[0].init_pos = [hi + 1.0, hi + 2.0]               # never inside the prior box
    calls = []
    inner = m.lnprior_batch
    m.lnprior_batch = lambda th: (calls.append(1), inner(th))[1]
    p0 = test_init(m, 1, 8, np.random.default_rng(3), max_repeats=100)
    assert "COULDNT FIND VALID INITIAL POSITION" in capsys.readouterr().out
    assert len(calls) == 100
    assert p0.shape == (1, 8, len(free)) and np.all(p0[0, :, 0] > hi)
