"""End-to-end ``Simulation.autorun`` on the device engine (emp/emp.py:2414-2456):
51 Peg from 0 Keplerians, BIC comparison, constrain-next-run by range.

Checks: the loop adds the planet and goes on exactly while BIC improves by
more than 5 (up to k_end); the statistics of each run agree with the oracle evaluated at
the run's own point estimates (like_max = lnL at the max-posterior sample,
chi2 and RMSE from my_model there, BIC/AIC/HQIC from like_max, DIC from lnL
at the posterior mean); the planet's next-run limits are its 95% HDI where
that is wider than the tolerance.
"""
import numpy as np
import pytest

from tests.conftest import GOLDEN, oracle_for

pytestmark = pytest.mark.gpu


def _sim():
    from astroemperor_amd import Simulation
    sim = Simulation()
    sim.read_loc = GOLDEN + "/"
    sim.load_data("51Peg")
    sim.set_engine("reddemcee")
    sim.keplerian_parameterisation = 1
    sim.engine_config["setup"] = np.array([5, 64, 600, 1])
    sim.engine_config["seed"] = 1234
    sim.run_config["burnin"] = 0.5
    sim.add_condition(["Period 1", "limits", [3.5, 5.0]])
    return sim


def test_autorun_51peg():
    sim = _sim()
    res = sim.autorun(0, 2)
    ks = [r["k"] for r in res]
    print("autorun:", [(r["k"], round(r["BIC"], 2), round(r["like_max"], 3)) for r in res], flush=True)
    assert ks == list(range(len(res))) and len(res) >= 2
    assert res[0]["BIC"] - res[1]["BIC"] > 5             # the planet is detected
    # the loop went on exactly while BIC improved by more than 5
    for a, b in zip(res[:-2], res[1:-1]):
        assert a["BIC"] - b["BIC"] > 5
    if len(res) < 3:
        assert res[-2]["BIC"] - res[-1]["BIC"] <= 5
    assert res[1]["like_max"] > -875.0                    # near the recorded KAT (-869.48)
    # for the run after k = 1 the planet's amplitude is limited to its 95% HDI
    # from that run; the period's HDI is narrower than tol = 1e-4 of its
    # value, so its limits stay (emp.py:2765-2767); both start from the
    # 1-sigma ranges
    per, amp = sim.blocks__[0][0], sim.blocks__[0][1]
    assert (per.name, amp.name) == ("Period 1", "Amplitude 1")
    if len(res) == 3:
        assert per.limits == [3.5, 5.0]
        assert 45.0 < amp.limits[0] < 55.9 < amp.limits[1] < 65.0
        assert per.init_pos[0] <= res[1]["fit"][0] <= per.init_pos[1]


def test_run_statistics_match_oracle():
    sim = _sim()
    sim._autorun_add_RV_ins()
    sim.add_keplerian()
    sim.run()
    m = sim.model
    o = oracle_for(m)
    assert abs(o.my_likelihood(sim.fit_max) - sim.like_max) <= 1e-10 * abs(sim.like_max)
    assert abs(o.my_prior(sim.fit_max) + sim.like_max - sim.post_max) <= 1e-9 * abs(sim.post_max)
    ymod, err2 = o.my_model(sim.ajuste)
    r = m.data.Y - ymod
    assert np.isclose(sim.chi2, np.sum(r ** 2 / err2), rtol=1e-10)
    assert np.isclose(sim.RMSE, np.sqrt(np.mean(r ** 2)), rtol=1e-10)
    n, k = m.data.ndat, m.ndim__
    assert np.isclose(sim.BIC, np.log(n) * k - 2 * sim.like_max, rtol=1e-14)
    assert np.isclose(sim.AIC, 2 * k - 2 * sim.like_max, rtol=1e-14)
    assert np.isclose(sim.HQIC, 2 * k * np.log(np.log(n)) - 2 * sim.like_max, rtol=1e-14)
    dic = -2 * o.my_likelihood(sim.fit_mean) + np.var(-2 * sim.likes[0])
    assert np.isclose(sim.DIC, dic, rtol=1e-10)
    z, dz = sim.evidence
    assert np.isfinite(z) and z < sim.like_max
    # EMPEROR's default 'hybrid' is not available: the flow falls back to
    # classic TI (emp.py:1443-1446)
    assert sim.evidence_method == "hybrid"
    classic = sim.sampler.get_evidence_ti(discard=sim.reddemcee_discard, pchip=False)[:2]
    assert (z, dz) == (float(classic[0]), float(classic[1]))
    for j, p in enumerate(q for b in sim.blocks__ for q in b if q.fixed is None):
        assert p.value == sim.fit_max[j]
        assert p.limits[0] <= p.value_range[0] <= p.value_range[1] <= p.limits[1]


def test_run_with_two_steps_per_sweep():
    """The reference's default setup has nsteps = 2: the burn-in and the
    chain are counted in steps (niter = nsteps x nsweeps, emp.py:2512-2527)."""
    sim = _sim()
    sim.engine_config["setup"] = np.array([3, 32, 40, 2])
    sim._autorun_add_RV_ins()
    sim.add_keplerian()
    sim.run()
    niter = 2 * 40
    assert sim.reddemcee_discard == niter // 2
    assert sim.chain.shape == (3, (niter - niter // 2) * 32, sim.model.ndim__)
    assert sim.sampler.iteration == niter
    assert np.isfinite(sim.like_max)


def test_run_with_gradient_moves():
    """engine_config['moves'] reaches the sampler: a run mixing gradient
    (Langevin) sweeps 1:3 with stretch sweeps."""
    from astroemperor_amd import GradientMove, StretchMove
    sim = _sim()
    sim.engine_config["setup"] = np.array([3, 32, 30, 1])
    sim.engine_config["moves"] = [(StretchMove(), 0.75), (GradientMove(), 0.25)]
    sim._autorun_add_RV_ins()
    sim.add_keplerian()
    sim.run()
    assert [type(m).__name__ for m, _ in sim.sampler.moves] == ["StretchMove", "GradientMove"]
    assert sim.chain.shape[1] == 15 * 32 and np.isfinite(sim.like_max)


@pytest.mark.parametrize("mode", [0, 2])
def test_adaptation_then_frozen_ladder(mode):
    """run_config['adaptation_batches'/'adaptation_nsweeps'] selects the
    freeze script (emp.py:690-704, endit_freeze1.scr:1-16): the ladder adapts
    during the burn-in sweeps, select_adjustment('00') freezes it, and the
    remaining sweeps keep it."""
    sim = _sim()
    sim.engine_config["setup"] = np.array([4, 32, 60, 1])
    sim.engine_config["adapt_mode"] = mode
    sim.engine_config["adapt_tau"] = 20
    sim.run_config["adaptation_batches"] = 2
    sim.run_config["adaptation_nsweeps"] = 100
    sim._autorun_add_RV_ins()
    sim.add_keplerian()
    sim.run()
    d = sim.reddemcee_discard
    bh = sim.sampler.get_betas()                        # ladder after each sweep
    assert d == 30 and bh.shape == (60, 4)
    assert np.max(np.abs(np.diff(bh[:d], axis=0))) > 0  # adapting
    assert np.all(bh[d:] == bh[d - 1])                  # frozen
    assert sim.sampler.adapt_mode == -1
    assert sim.chain.shape == (4, 30 * 32, sim.model.ndim__) and np.isfinite(sim.like_max)
