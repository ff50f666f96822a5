"""GPU likelihood / prior / my_model parity against the oracle and the golden
vectors.  Bar: |lnL_gpu - lnL_oracle| <= 1e-10 |lnL| for FP64; the mixed path
within 1e-6 absolute of FP64; priors bit-identical (Beta: 1e-13)."""
import json
import os

import numpy as np
import pytest

from tests.conftest import make_sim, oracle_for

pytestmark = pytest.mark.gpu

REL = 1e-10


def _draws(model, n, rng):
    lim = model.limits_free()
    lo, hi = lim[:, 0], lim[:, 1]
    return lo + (hi - lo) * rng.uniform(size=(n, len(lo)))


def _check_lnl(model, theta, label):
    o = oracle_for(model)
    ref = np.array([o.my_likelihood(t) for t in theta])
    got = model.lnl_batch(theta, "fp64")
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(got), fin), label
    err = np.abs(got[fin] - ref[fin]) / np.abs(ref[fin])
    assert err.max() <= REL, (label, err.max())
    mixed = model.lnl_batch(theta, "mixed")
    assert np.array_equal(np.isfinite(mixed), fin), label
    return err.max(), np.abs(mixed[fin] - got[fin]).max()


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 6, 7])
def test_lnl_51peg_parameterisations(p):
    rng = np.random.default_rng(100 + p)
    for k in (1, 2):
        sim = make_sim("51Peg", k, parameterisation=p)
        th = _draws(sim.model, 300, rng)
        e, dm = _check_lnl(sim.model, th, f"p{p} k{k}")
        print(f"p={p} k={k}: max rel err fp64 {e:.2e}, max |mixed - fp64| {dm:.2e}", flush=True)
        assert dm < 1e-6


def test_lnl_gj876_eight_instruments():
    rng = np.random.default_rng(7)
    sim = make_sim("GJ876", 2, parameterisation=1)
    th = _draws(sim.model, 200, rng)
    e, dm = _check_lnl(sim.model, th, "gj876")
    assert dm < 1e-6


@pytest.mark.parametrize("wpt", [1, 2, 4, 8])
@pytest.mark.parametrize("moav", [{"order": 1, "global": False}, {"order": 1, "global": True}])
@pytest.mark.parametrize("semantics", ["reference", "intended"])
def test_lnl_moving_average(wpt, moav, semantics):
    rng = np.random.default_rng(11)
    sim = make_sim("GJ876", 1, parameterisation=1, moav=moav, semantics=semantics)
    sim.model.set_waves_per_theta(wpt)
    th = _draws(sim.model, 64, rng)
    e, dm = _check_lnl(sim.model, th, f"ma {moav} {semantics} wpt{wpt}")
    assert dm < 1e-6


def test_lnl_acceleration_and_fixed():
    rng = np.random.default_rng(12)
    sim = make_sim("GJ876", 2, parameterisation=0, acceleration=2,
                   conditions=[("Eccentricity 2", "fixed", 0.1), ("Jitter 3", "fixed", 2.0)])
    th = _draws(sim.model, 100, rng)
    _check_lnl(sim.model, th, "acc+fixed")


def test_lnl_golden_mpmath(golden_dir):
    """GPU lnL and my_model against the 50-digit vectors directly."""
    from astroemperor_amd import JitterBlock, KeplerianBlock, OffsetBlock, RVData, ReddModel
    with open(os.path.join(golden_dir, "lnl_mpmath.json")) as f:
        gold = json.load(f)
    for case in gold["cases"]:
        ds = gold["datasets"][case["data"]]
        data = RVData(ds["X"], ds["Y"], ds["YERR"], ds["FLAG"])
        blocks = []
        for n in range(len(case["planets"])):
            b = KeplerianBlock(case["parameterisation"], n + 1)
            for p in b:
                p.limits, p.prior = [-1e9, 1e9], "Uniform"
            blocks.append(b)
        for b in (OffsetBlock(data.nins), JitterBlock(data.nins)):
            for p in b:
                p.limits, p.prior = [-1e9, 1e9], "Uniform"
            blocks.append(b)
        m = ReddModel(blocks, data)
        th = np.array(sum(case["planets"], []) + case["offsets"] + case["jitters"])
        got = m.lnl_batch(th[None, :])[0]
        assert abs(got - case["lnl"]) <= REL * abs(case["lnl"]), (case["parameterisation"], got, case["lnl"])
        mu, err2 = m.my_model(th)
        assert np.allclose(mu, case["mu"], rtol=0, atol=1e-9)
        mixed = m.lnl_batch(th[None, :], "mixed")[0]
        assert abs(mixed - case["lnl"]) < 1e-6


def test_prior_bit_exact():
    rng = np.random.default_rng(3)
    for p in [0, 1, 6, 7]:
        sim = make_sim("GJ876", 2, parameterisation=p, conditions=[("Amplitude 1", "prior", "Normal"),
                                                                  ("Amplitude 1", "prargs", [10.0, 4.0])])
        m = sim.model
        lim = m.limits_free()
        th = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(size=(400, len(lim)))
        # push one coordinate of every third row just outside its support
        for r in range(0, 400, 3):
            j = rng.integers(len(lim))
            th[r, j] = lim[j, 1] + 0.01 * (lim[j, 1] - lim[j, 0]) if r % 2 else lim[j, 0] - 1e-9
        o = oracle_for(m)
        ref = np.array([o.my_prior(t) for t in th])
        got = m.lnprior_batch(th)
        # Same support, and the same sum to within 2 ulp: the reference squares
        # numpy scalars with `** 2`, i.e. libm pow, which is 1 ulp off the
        # correctly rounded x*x in ~0.1% of cases (the device squares exactly).
        assert np.array_equal(np.isfinite(got), np.isfinite(ref))
        fin = np.isfinite(ref)
        assert np.all(np.abs(got[fin] - ref[fin]) <= 2 * np.spacing(np.abs(ref[fin]))), p
        assert np.isfinite(ref).sum() > 20 and (~np.isfinite(ref)).sum() > 20


def test_prior_beta():
    rng = np.random.default_rng(4)
    sim = make_sim("GJ876", 1, parameterisation=6, conditions=[("Eccentricity 1", "prior", "Beta"),
                                                              ("Eccentricity 1", "prargs", [0.867, 3.03])])
    m = sim.model
    lim = m.limits_free()
    th = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(size=(300, len(lim)))
    o = oracle_for(m)
    ref = np.array([o.my_prior(t) for t in th])
    got = m.lnprior_batch(th)
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-13)


def test_my_model_matches_oracle():
    rng = np.random.default_rng(5)
    for moav, sem in [(None, "reference"), ({"order": 1, "global": False}, "intended"),
                      ({"order": 1, "global": True}, "reference")]:
        sim = make_sim("GJ876", 2, parameterisation=1, moav=moav, semantics=sem)
        m = sim.model
        o = oracle_for(m)
        th_all = _draws(m, 40, rng)
        keep = [t for t in th_all if np.isfinite(o.my_likelihood(t))][:5]
        assert len(keep) == 5
        for th in keep:
            mu, e2 = m.my_model(th)
            mu_o, e2_o = o.my_model(th)
            assert np.allclose(mu, mu_o, rtol=1e-11, atol=1e-9)
            assert np.allclose(e2, e2_o, rtol=1e-14, atol=0)
            assert abs(m.my_likelihood(th) - o.my_likelihood(th)) <= REL * abs(o.my_likelihood(th))


def test_kat_on_gpu():
    """lnL at the 51 Peg k=0 MAP and k=1 optimum from the GPU kernel."""
    sim0 = make_sim("51Peg", 0)
    assert abs(sim0.model.my_likelihood([-0.152, 36.062]) - (-1308.796)) < 2e-3
    sim1 = make_sim("51Peg", 1, parameterisation=1)
    v = sim1.model.my_likelihood([4.230732, 55.9957, 1.7274, 0.09596, 0.06079, 3.6940, 2.9475])
    assert abs(v - (-869.4598)) < 5e-3


def test_nonfinite_maps_to_minus_inf():
    sim = make_sim("51Peg", 1, parameterisation=0)
    m = sim.model
    before = m.nonfinite_count()
    th = np.array([[4.23, 50.0, 1.0, 1.5, 0.3, 0.0, 5.0],    # e > 1: NaN in the reference
                   [4.23, 50.0, 1.0, 0.3, 0.3, 0.0, 5.0]])
    out = m.lnl_batch(th)
    assert out[0] == -np.inf and np.isfinite(out[1])
    assert m.nonfinite_count() == before + 1


def test_large_config_mixed_vs_fp64():
    """Bench-sized workload: 3 planets, 4 instruments, 1530 epochs, MA(1)."""
    from astroemperor_amd import synth
    d = synth.c2()
    m = d["model"]
    rng = np.random.default_rng(9)
    th = _draws(m, 2048, rng)
    a = m.lnl_batch(th, "fp64")
    b = m.lnl_batch(th, "mixed")
    fin = np.isfinite(a)
    assert np.array_equal(fin, np.isfinite(b))
    dm = np.abs(a[fin] - b[fin])
    print(f"c2 mixed vs fp64: max {dm.max():.2e} p99 {np.percentile(dm, 99):.2e}", flush=True)
    assert dm.max() < 1e-6
    o = oracle_for(m)
    idx = rng.choice(np.where(fin)[0], 16, replace=False)
    ref = np.array([o.my_likelihood(t) for t in th[idx]])
    assert np.max(np.abs(a[idx] - ref) / np.abs(ref)) <= REL


@pytest.mark.parametrize("lo,hi", [(0.0, 0.5), (0.5, 0.75), (0.75, 0.95), (0.95, 0.99)])
@pytest.mark.parametrize("wpt", [1, 4])
def test_eccentricity_bands_mixed_vs_fp64(lo, hi, wpt):
    """Every planet at e in [lo, hi]: one band per Kepler path of the mixed
    path (the per-vector table with an FP64 Newton step up to e = 0.5, the
    table with an FP64 Halley step up to 0.75, the table with an FP32 and
    an FP64 Halley step up to 0.95, the FP32 Markley start with the same two
    steps above), on draws that fit the data badly, which magnifies every
    model error: still within 1e-6 of the FP64 path.  At 1 wave per vector
    a workgroup holds 8 vectors and keeps the tables as nodes, at 4 it holds
    2 and keeps them per interval; the two forms evaluate the same FP32
    cubic up to the rounding of M / h, so they agree far inside the bar."""
    from astroemperor_amd import synth
    m = synth.c2()["model"]
    rng = np.random.default_rng(17)
    th = _draws(m, 4096, rng)
    ecol = [i for i, n in enumerate(m.names_free()) if n.startswith("Eccentricity")]
    th[:, ecol] = rng.uniform(lo, hi, size=(len(th), len(ecol)))
    m.set_waves_per_theta(wpt)
    try:
        a = m.lnl_batch(th, "fp64")
        b = m.lnl_batch(th, "mixed")
        if wpt == 4:
            m.set_waves_per_theta(1)
            b1 = m.lnl_batch(th, "mixed")
    finally:
        m.set_waves_per_theta(0)
    fin = np.isfinite(a)
    assert fin.sum() > 1000 and np.array_equal(fin, np.isfinite(b))
    dm = np.abs(a[fin] - b[fin])
    print(f"e in [{lo}, {hi}] wpt {wpt} mixed vs fp64: max {dm.max():.2e} p99 {np.percentile(dm, 99):.2e}",
          flush=True)
    assert dm.max() < 1e-6
    if wpt == 4:
        dn = np.abs(b1[fin] - b[fin])
        print(f"e in [{lo}, {hi}] nodes vs intervals: max {dn.max():.2e}", flush=True)
        assert dn.max() < 1e-6
