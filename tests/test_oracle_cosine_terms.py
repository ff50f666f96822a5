"""The oracle's Sinusoid and MagneticCycle terms (sinusoid00.model,
magneticcycle00.model), the blocks' names and default limits
(block_repo.py:75-100, 275-314, 805-830) and where the terms enter relative to
the moving average.

Pins, each chosen so that a plausible slip fails:

* a sinusoid is a circular Keplerian with omega = 0 (kep00.model with e = 0,
  whose E = M; the Keplerian oracle is pinned to 50-digit mpmath vectors):
  catches a sign or factor on the phase, the period used as a frequency;
* a magnetic cycle equals a sinusoid at twice the period (pi / P) plus one at
  the period (2 pi / P), with the amplitudes and phases crossed over as the
  block orders them (P, A1, A2, phase1, phase2): catches swapped harmonics
  or parameter slots;
* values at t = 0 (A cos phase) and a quarter period later;
* the terms follow MOAV in block order, so with a global MA(1) removing the
  sinusoid changes the model by exactly the unfiltered sinusoid.
"""
import numpy as np
import pytest

from tests.conftest import make_sim, oracle_for


def _model_with(sim, values, rng):
    """A free vector inside the limits with the block of type values[0]
    set to values[1] (by position: the reference names a sinusoid's and the
    first Keplerian's period alike, 'Period 1')."""
    m = sim.model
    lim = m.limits_free()
    th = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(0.2, 0.8, size=len(lim))
    typ, vals = values
    start = m.block(typ).slice.start
    pos = {int(c): j for j, c in enumerate(m.C_)}
    for i, v in enumerate(vals):
        th[pos[start + i]] = v
    names = [m.params()[int(c)].name for c in m.C_]
    return th, names


def test_block_names_and_defaults():
    sim = make_sim("51Peg", 1, sinusoid=1, magnetic_cycle=1)
    m = sim.model
    types = [b.type_ for b in m.blocks]
    assert types == ["Keplerian", "Offset", "Jitter", "Sinusoid", "MagneticCycle"]
    sn, mc = m.block("Sinusoid"), m.block("MagneticCycle")
    assert [p.name for p in sn] == ["Period 1", "Amplitude 1", "Phase 1"]
    assert [p.name for p in mc] == ["Period S1", "Amplitude S1", "Amplitude S2", "Phase S1", "Phase S2"]
    d = m.data
    span = float(d.X.max() - d.X.min())
    amp = float(np.std(d.Y)) * 2.0
    assert np.allclose(sn.get_attr("limits"), [[1.5, span], [1e-6, amp], [0, 2 * np.pi]])
    assert np.allclose(mc.get_attr("limits"), [[1.5, span], [1e-6, amp], [1e-6, amp], [0, 2 * np.pi],
                                               [0, 2 * np.pi]])
    assert all(p.prior == "Uniform" for p in list(sn) + list(mc))


def _term(o, th):
    """Model contribution of the cosine blocks: my_model minus my_model with
    every cosine amplitude at zero."""
    full = o.full_theta(th)
    base = full.copy()
    for b in o.spec["blocks"]:
        if b["type"] == "Sinusoid":
            base[b["start"] + 1] = 0.0
        elif b["type"] == "MagneticCycle":
            base[b["start"] + 1] = base[b["start"] + 2] = 0.0
    free = o.C_
    m1, _ = o.my_model(full[free])
    m0, _ = o.my_model(base[free])
    return m1 - m0


def test_sinusoid_is_circular_keplerian():
    from oracle.emp_oracle import keplerian_term
    rng = np.random.default_rng(3)
    sim = make_sim("51Peg", 0, sinusoid=1)
    o = oracle_for(sim.model)
    X = o.X
    for _ in range(20):
        P = rng.uniform(1.5, 400.0)
        A = rng.uniform(1.0, 80.0)
        ph = rng.uniform(0, 2 * np.pi)
        th, names = _model_with(sim, ("Sinusoid", (P, A, ph)), rng)
        got = _term(o, th)
        kep = keplerian_term([P, A, ph, 0.0, 0.0], X, 0)
        assert np.allclose(got, kep, rtol=0, atol=1e-11 * A)


def test_sinusoid_spot_values():
    sim = make_sim("51Peg", 0, sinusoid=1)
    o = oracle_for(sim.model)
    P, A, ph = 10.0, 3.0, 0.7
    th, _ = _model_with(sim, ("Sinusoid", (P, A, ph)), np.random.default_rng(0))
    got = _term(o, th)
    # X[0] = 0 (the time origin is the first epoch)
    assert o.X[0] == 0.0
    assert abs(got[0] - A * np.cos(ph)) < 1e-13
    # a quarter period on: cos(ph + pi/2) = -sin(ph)
    o2 = oracle_for(sim.model)
    o2.X = np.array([0.0, P / 4, P / 2, P])
    o2.Y = np.zeros(4)
    o2.YERR = np.ones(4)
    o2.FLAG = np.ones(4, dtype=int)
    o2.ndat = 4
    o2.masks = [np.ones(4, dtype=bool)]
    got2 = _term(o2, th)
    assert np.allclose(got2, A * np.array([np.cos(ph), -np.sin(ph), -np.cos(ph), np.cos(ph)]), atol=1e-13)


def test_magnetic_cycle_is_two_sinusoids():
    rng = np.random.default_rng(4)
    sim_m = make_sim("51Peg", 0, magnetic_cycle=1)
    sim_s = make_sim("51Peg", 0, sinusoid=1)
    om, os_ = oracle_for(sim_m.model), oracle_for(sim_s.model)
    for _ in range(20):
        P = rng.uniform(1.5, 300.0)
        A1, A2 = rng.uniform(1.0, 50.0, 2)
        p1, p2 = rng.uniform(0, 2 * np.pi, 2)
        th, _ = _model_with(sim_m, ("MagneticCycle", (P, A1, A2, p1, p2)), rng)
        got = _term(om, th)
        ts, _ = _model_with(sim_s, ("Sinusoid", (2 * P, A1, p1)), rng)
        slow = _term(os_, ts)
        ts, _ = _model_with(sim_s, ("Sinusoid", (P, A2, p2)), rng)
        fast = _term(os_, ts)
        assert np.allclose(got, slow + fast, rtol=0, atol=1e-11 * (A1 + A2))


@pytest.mark.parametrize("moav,semantics", [({"order": 1, "global": True}, "reference"),
                                            ({"order": 1, "global": False}, "intended")])
def test_cosine_terms_follow_the_moving_average(moav, semantics):
    """With the MA active, dropping the sinusoid changes the model by exactly
    A cos(2 pi X / P + phase): the MA residuals never contain it."""
    rng = np.random.default_rng(5)
    sim = make_sim("GJ876", 1, moav=moav, semantics=semantics, sinusoid=1)
    o = oracle_for(sim.model)
    P, A, ph = 37.0, 20.0, 1.1
    th, names = _model_with(sim, ("Sinusoid", (P, A, ph)), rng)
    for n in names:
        if n.startswith("MACoefficient"):
            th[names.index(n)] = 0.25
    got = _term(o, th)
    assert np.allclose(got, A * np.cos(2 * np.pi / P * o.X + ph), rtol=0, atol=1e-11 * A)
    # and the MA does act on this model: a different coefficient moves it
    th2 = th.copy()
    for n in names:
        if n.startswith("MACoefficient"):
            th2[names.index(n)] = 0.05
    assert np.max(np.abs(o.my_model(th)[0] - o.my_model(th2)[0])) > 1e-3
