"""Stellar-activity (SAI) columns: the loader's scaling (qol_utils.py:73-98),
the block (block_repo.py:243-272, 742-749), the switch (emp/emp.py:2305-2312)
and the oracle's term (sai00.model, emp_model.py:425-432, 736-745).

Pins:

* each activity column of a file is an increasing affine image of the raw
  column whose minimum and maximum land on the file's RV minimum and maximum
  (an affine map is fixed by two points, so this is the whole rule); epochs of
  other files hold zeros (fillna(0));
* a unit coefficient reproduces its column in the model, and the term is
  linear in the coefficients;
* the term follows MOAV in block order: with the MA active, removing it
  changes the model by exactly sum_j zeta_j a_ij.
"""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN, make_sim, oracle_for


def test_loader_scaling_and_zero_fill():
    from astroemperor_amd.data import DataWrapper
    w = DataWrapper("GJ876", read_loc=GOLDEN + "/")
    d = w.get_data__()
    assert d.nsai == [0, 0, 0, 0, 6, 0, 0, 4]
    assert d.SA.shape == (d.ndat, 10)
    col = 0
    for k, fn in enumerate(w.filenames):
        raw = np.loadtxt(os.path.join(w.path, fn))
        m = d.FLAG == k + 1
        own = slice(col, col + d.nsai[k])
        other = np.ones(d.SA.shape[1], dtype=bool)
        other[own] = False
        assert np.all(d.SA[m][:, other] == 0.0)
        rv = d.Y[m]
        # epochs of a file keep their file order under the stable-enough sort here
        order = np.argsort(raw[:, 0], kind="quicksort")
        for j in range(d.nsai[k]):
            a = d.SA[m][:, col + j]
            x = raw[order, 3 + j]
            assert abs(a.min() - rv.min()) <= 1e-12 * abs(rv.min())
            assert abs(a.max() - rv.max()) <= 1e-12 * abs(rv.max())
            slope, icpt = np.polyfit(x, a, 1)
            assert slope > 0
            assert np.max(np.abs(a - (slope * x + icpt))) <= 1e-9 * np.ptp(rv)
        col += d.nsai[k]


def test_switch_and_block():
    off = make_sim("GJ876", 1)
    assert off.model.block("StellarActivity") is None
    assert off.my_data.SA.shape[1] == 0 and off.cornums == [0] * 8
    on = make_sim("GJ876", 1, switch_SA=True)
    b = on.model.block("StellarActivity")
    assert [bb.type_ for bb in on.model.blocks] == ["Keplerian", "Offset", "Jitter", "StellarActivity"]
    names = [p.name for p in b]
    assert names == [f"Staract 5 {j}" for j in range(1, 7)] + [f"Staract 8 {j}" for j in range(1, 5)]
    assert np.allclose(b.get_attr("limits"), [[-1.0, 1.0]] * 10)
    assert all(p.prior == "Uniform" for p in b)


def _with_zeta(m, th, zeta):
    full = np.array(th, dtype=float)
    s = m.block("StellarActivity").slice.start
    pos = {int(c): j for j, c in enumerate(m.C_)}
    for i, z in enumerate(zeta):
        full[pos[s + i]] = z
    return full


@pytest.mark.parametrize("moav", [None, {"order": 1, "global": True}, {"order": 1, "global": False}])
def test_oracle_term(moav):
    sim = make_sim("GJ876", 1, switch_SA=True, moav=moav, semantics="intended")
    m = sim.model
    o = oracle_for(m)
    rng = np.random.default_rng(0)
    lim = m.limits_free()
    th = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(0.2, 0.8, size=len(lim))
    base = o.my_model(_with_zeta(m, th, np.zeros(10)))[0]
    SA = m.data.SA
    for j in (0, 5, 6, 9):
        e = np.zeros(10)
        e[j] = 1.0
        got = o.my_model(_with_zeta(m, th, e))[0] - base
        assert np.allclose(got, SA[:, j], rtol=0, atol=1e-9 * np.abs(SA[:, j]).max())
    zeta = rng.uniform(-1, 1, 10)
    got = o.my_model(_with_zeta(m, th, zeta))[0] - base
    assert np.allclose(got, SA @ zeta, rtol=0, atol=1e-9 * np.abs(SA).max())


def test_constant_activity_column_is_zero_filled():
    """A constant column maps 0/0 to NaN, which the loader's fillna(0) zeroes."""
    from astroemperor_amd.data import RVData
    rng = np.random.default_rng(0)
    t = np.sort(rng.uniform(0, 100, 20))
    sa = np.stack([np.full(20, 3.0), rng.normal(size=20)], axis=1)
    d = RVData.from_arrays([t], [rng.normal(size=20)], [np.ones(20)], sa_list=[sa])
    assert np.all(d.SA[:, 0] == 0.0)
    assert np.all(np.isfinite(d.SA))
