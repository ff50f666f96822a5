"""Host post-processing and the autorun loop (CPU): highest-density
intervals, comparison criteria, constrain-next-run and the loop's control
flow (emp/emp.py:1112-1212, 1720-1767, 2414-2456, 2654-2778,
emp/math_utils.py:72-115)."""
import numpy as np
import pytest

from astroemperor_amd.postprocess import hdi_of_chain, hdi_of_samples


def test_hdi_known_distributions():
    rng = np.random.default_rng(0)
    x = rng.normal(size=200_000)
    lo, hi = hdi_of_samples(x, 0.90)
    assert abs(lo + 1.6449) < 0.03 and abs(hi - 1.6449) < 0.03
    e = rng.exponential(size=200_000)
    lo, hi = hdi_of_samples(e, 0.90)
    assert lo == e.min()                              # a decreasing density starts at its mode
    assert abs(hi - np.log(10.0)) < 0.03


def test_hdi_is_the_shortest_window():
    rng = np.random.default_rng(1)
    ch = rng.gamma(2.0, size=(301, 4))
    for c in (0.36, 0.9, 0.95):
        got = hdi_of_chain(ch, c)
        n = len(ch)
        k = int(np.floor(c * n))
        for d in range(4):
            s = np.sort(ch[:, d])
            inside = np.sum((ch[:, d] >= got[d, 0]) & (ch[:, d] <= got[d, 1]))
            assert inside >= k + 1
            best = min(s[i + k] - s[i] for i in range(n - k))
            assert got[d, 1] - got[d, 0] == best


def _sim_with_script(bics, k_start=0, k_end=5):
    from astroemperor_amd import Simulation
    sim = Simulation()
    calls = []
    state = {"i": 0}

    def fake_run():
        calls.append(sim.kplanets__)
        sim.BIC = bics[state["i"]]
        state["i"] += 1

    sim.run = fake_run
    sim._autorun_add_RV_ins = lambda: None
    sim.add_keplerian = lambda: setattr(sim, "kplanets__", sim.kplanets__ + 1)
    sim.update_model = lambda: None
    sim._set_init_pos = lambda: None
    sim._constrain_next_run = lambda: None
    sim.autorun(k_start, k_end)
    return calls


def test_autorun_stops_when_bic_stops_improving():
    # k=0: 1000, k=1: 900 (better by 100), k=2: 897 (better by 3 < 5): stop after k=2
    assert _sim_with_script([1000.0, 900.0, 897.0, 800.0]) == [0, 1, 2]
    # always improving: runs up to k_end
    assert _sim_with_script([100.0, 90.0, 80.0, 70.0], k_start=1, k_end=3) == [1, 2, 3]


def test_criteria_direction_and_pass():
    from astroemperor_amd import Simulation
    sim = Simulation()
    assert sim.criteria_compare(90.0, 100.0)            # BIC: lower by more than 5
    assert not sim.criteria_compare(96.0, 100.0)
    sim.set_comparison_criteria("Evidence")
    assert sim.criteria_compare(-90.0, -100.0)          # evidence: higher by more than 5
    assert not sim.criteria_compare(-100.0, -90.0)
    sim.set_tolerance(20)
    assert not sim.criteria_compare(-90.0, -100.0)
    sim.set_comparison_criteria("Pass")
    assert sim.criteria_compare(0.0, 0.0)


def test_constrain_range_and_sigma():
    from astroemperor_amd.blocks import JitterBlock, KeplerianBlock
    from astroemperor_amd import Simulation
    sim = Simulation()
    kb = KeplerianBlock(1, 1)
    jb = JitterBlock(1)
    for p, lim in zip(list(kb) + list(jb), [[1.5, 100], [0, 50], [0, 6.3], [-1, 1], [-1, 1], [0, 20]]):
        p.limits = lim
        p.value, p.sigma = 10.0, 1.0
        p.value_low3, p.value_high3 = 8.0, 12.5
    kb[3].value, kb[3].value_low3, kb[3].value_high3 = 0.5, 0.49999, 0.50001    # too narrow: kept
    kb[2].value = 5.5
    kb[1].value_low3 = -5.0                             # below the limit: clipped to it
    sim.blocks__ = [kb, jb]
    sim._apply_constrain_range()
    got = {c[0]: c[2] for c in sim.conds}
    assert got["Period 1"] == [8.0, 12.5]
    assert got["Amplitude 1"] == [0, 12.5]
    assert "Ecc_sin 1" not in got
    assert got["Jitter 1"] == [0, 12.5]                 # jitter keeps its lower limit
    sim.conds = []
    sim._apply_constrain_sigma()
    got = {c[0]: c[2] for c in sim.conds}
    assert got["Period 1"] == [7.0, 13.0]
    assert got["Phase 1"] == [2.5, 6.3]                 # 5.5 -/+ 3 sigma, upper end clipped
    sim.constrain["method"] = "GM"
    with pytest.raises(NotImplementedError):
        sim._constrain_next_run()


class _FakeEvidenceSampler:
    """Records which estimator the evidence flow asked for."""

    def __init__(self, ti_value=-10.0, pchip_value=-11.0, hybrid=None):
        self.calls = []
        self.ti_value, self.pchip_value, self.hybrid = ti_value, pchip_value, hybrid

    def get_evidence_hybrid(self, discard=0):
        self.calls.append(("hybrid", discard))
        if self.hybrid is None:
            raise NotImplementedError
        return self.hybrid

    def get_evidence_ti(self, discard=0, pchip=True):
        self.calls.append(("ti", discard, pchip))
        return (self.pchip_value, 0.5) if pchip else (self.ti_value, 0.25)


def test_evidence_flow_follows_emperor():
    """emp.py:1432-1447: the configured estimator first; -inf falls back to TI
    with PCHIP, then classic TI; an exception goes straight to classic TI."""
    from astroemperor_amd import Simulation
    sim = Simulation()
    assert sim.evidence_method == "hybrid"
    sim.reddemcee_discard = 7
    sim.sampler = _FakeEvidenceSampler()
    assert sim._get_evidence() == (-10.0, 0.25)
    assert sim.sampler.calls == [("hybrid", 7), ("ti", 7, False)]
    sim.sampler = _FakeEvidenceSampler(hybrid=(-np.inf, 0.0))
    assert sim._get_evidence() == (-11.0, 0.5)
    assert sim.sampler.calls == [("hybrid", 7), ("ti", 7, True)]
    sim.sampler = _FakeEvidenceSampler(hybrid=(-np.inf, 0.0), pchip_value=-np.inf)
    assert sim._get_evidence() == (-10.0, 0.25)
    assert sim.sampler.calls == [("hybrid", 7), ("ti", 7, True), ("ti", 7, False)]
