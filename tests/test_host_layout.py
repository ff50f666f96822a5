"""Host logic: data staging, block layout, SmartSetter defaults, conditions,
descriptor contents (CPU; no device calls)."""
import numpy as np
import pytest

from tests.conftest import make_sim, oracle_for


def test_loader_matches_oracle(peg51, golden_dir):
    from astroemperor_amd import DataWrapper
    d = DataWrapper("51Peg", read_loc=golden_dir + "/").get_data__()
    assert d.ndat == 256 and d.nins == 1
    assert np.array_equal(d.X, peg51["X"]) and np.array_equal(d.Y, peg51["Y"])
    assert np.array_equal(d.YERR, peg51["YERR"]) and np.array_equal(d.FLAG, peg51["FLAG"])
    assert d.X[0] == 0.0 and np.all(np.diff(d.X) >= 0)
    assert abs(d.Y.mean()) < 1e-9                        # qol_utils.py:84-85


def test_multi_instrument_loader(golden_dir):
    from astroemperor_amd import DataWrapper
    d = DataWrapper("GJ876", read_loc=golden_dir + "/").get_data__()
    assert d.nins == 8 and d.ndat == 752          # comment lines in the .vels files are skipped
    # flags follow sorted filenames; every instrument is mean-subtracted
    for k in range(1, 9):
        assert abs(d.Y[d.FLAG == k].mean()) < 1e-8


def test_smartsetter_defaults_51peg(peg51):
    sim = make_sim("51Peg", 1, parameterisation=1)
    m = sim.model
    names = m.names_free()
    assert names == ["Period 1", "Amplitude 1", "Phase 1", "Ecc_sin 1", "Ecc_cos 1", "Offset 1", "Jitter 1"]
    lim = m.limits_free()
    span = peg51["X"].max() - peg51["X"].min()
    std = np.std(peg51["Y"])
    amax = np.abs(peg51["Y"]).max()
    assert np.allclose(lim[0], [1.5, span]) and np.allclose(lim[1], [1e-6, 2 * std])
    assert np.allclose(lim[2], [0, 2 * np.pi]) and np.allclose(lim[3], [-1, 1])
    assert np.allclose(lim[5], [-amax, amax]) and np.allclose(lim[6], [1e-5, amax])
    pri = [p.prior for p in m.params()]
    assert pri == ["Uniform"] * 6 + ["Normal"]
    # derived eccentricity prior for parameterisation 1 (block_repo.py:662-683)
    spec = m.spec()
    assert spec["blocks"][0]["additional"][0]["prior"] == "Uniform"
    assert spec["blocks"][0]["additional"][0]["limits"] == [0, 1]


def test_block_order_and_fixed(peg51):
    sim = make_sim("51Peg", 2, parameterisation=0, moav={"order": 1, "global": False},
                   conditions=[("Eccentricity 1", "fixed", 0.0), ("MATimescale 1 Order 1", "fixed", 5.0)])
    m = sim.model
    types = [b["type"] for b in m.spec()["blocks"]]
    assert types == ["Keplerian", "Keplerian", "Offset", "Jitter", "MOAV"]
    assert m.ndim_full == 10 + 1 + 1 + 2 and m.ndim__ == 12
    assert list(m.A_) == [3, 13]
    o = oracle_for(m)
    th = np.arange(12, dtype=float) + 1.0
    full = o.full_theta(th)
    assert full[3] == 0.0 and full[13] == 5.0 and full[0] == 1.0 and full[4] == 4.0


def test_conditions_and_parameterisations(golden_dir):
    for p in [0, 1, 2, 3, 4, 6, 7]:
        sim = make_sim("51Peg", 1, parameterisation=p,
                       conditions=[(sim_name(p), "limits", [3.0, 5.0])])
        m = sim.model
        assert m.params()[0].limits == [3.0, 5.0]
        desc, keep = m.descriptor()
        assert desc.n_planets == 1 and desc.parameterisation == p and desc.ndim == 7


def sim_name(p):
    return "lPeriod 1" if p in (2, 6, 7) else "Period 1"


def test_descriptor_priors_normalised():
    sim = make_sim("51Peg", 1, parameterisation=1)
    desc, keep = sim.model.descriptor()
    terms = keep[6]
    # Uniform: log(1/(hi-lo)); Normal: mu, s, logZ, log(s sqrt(2 pi))
    assert terms[0].kind == 0 and abs(terms[0].a0 - np.log(1 / (terms[0].hi - terms[0].lo))) < 1e-15
    j = terms[6]
    assert j.kind == 1 and j.a0 == 5 and j.a1 == 5 and abs(j.a3 - np.log(5 * np.sqrt(2 * np.pi))) < 1e-15
    from scipy.stats import norm
    assert abs(j.a2 - np.log(norm.cdf((j.hi - 5) / 5) - norm.cdf((j.lo - 5) / 5))) < 1e-15
    ecc = keep[7]
    assert ecc[0].kind == 0 and ecc[0].lo == 0 and ecc[0].hi == 1


def test_unsupported_prior_raises():
    sim = make_sim("51Peg", 1, parameterisation=1, conditions=[("Period 1", "prior", "GaussianMixture")])
    with pytest.raises(NotImplementedError):
        sim.model.descriptor()


def test_philox_known_answers():
    """Random123 known-answer vectors for Philox-4x32-10 (kat_vectors)."""
    from astroemperor_amd import _abi
    L = _abi.lib()
    cases = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
             ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
             ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
              [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, want in cases:
        c = np.array(ctr, dtype=np.uint32)
        k = np.array(key, dtype=np.uint32)
        o = np.zeros(4, dtype=np.uint32)
        _abi.check(L.emp_philox4x32_10(c.ctypes.data, k.ctypes.data, o.ctypes.data))
        assert list(o) == want


def test_descriptor_cosine_block_offsets():
    """off_sinusoid / off_magnetic: the block starts, -1 when absent; a model
    smaller than a magnetic-cycle block without one creates."""
    from tests.conftest import make_sim
    m = make_sim("51Peg", 0).model
    desc, _ = m.descriptor()
    assert desc.off_sinusoid == -1 and desc.off_magnetic == -1
    assert desc.ndim_full < 5
    m2 = make_sim("51Peg", 1, sinusoid=1, magnetic_cycle=1).model
    desc, _ = m2.descriptor()
    assert desc.off_sinusoid == m2.block("Sinusoid").slice.start
    assert desc.off_magnetic == m2.block("MagneticCycle").slice.start
    assert desc.off_magnetic + 5 == desc.ndim_full


def test_descriptor_activity_columns():
    """off_sai / n_sai and the [N, n_sai] columns with their per-instrument counts."""
    from tests.conftest import make_sim
    m = make_sim("GJ876", 1, switch_SA=True).model
    desc, keep = m.descriptor()
    assert desc.off_sai == m.block("StellarActivity").slice.start and desc.n_sai == 10
    assert [desc.sai_counts[k] for k in range(8)] == [0, 0, 0, 0, 6, 0, 0, 4]
    assert desc.sai[3 * 10 + 4] == m.data.SA[3, 4]
    m0 = make_sim("GJ876", 1).model
    desc, _ = m0.descriptor()
    assert desc.off_sai == -1 and desc.n_sai == 0 and not desc.sai


def test_backend_views_of_a_saved_run(tmp_path):
    """sampler.backend / backend[t] as EMPEROR reads reddemcee's backend
    (emp/emp.py:734-761), over a saved run: rows in steps (2 per sweep here),
    the ladder history one row per step."""
    from astroemperor_amd.backend import FORMAT, SavedRun
    rng = np.random.default_rng(0)
    T, S, R, W, D = 3, 4, 8, 6, 2
    arrays = dict(chain=rng.normal(size=(T, R, W, D)), log_like=rng.normal(size=(T, R, W)),
                  log_prob=rng.normal(size=(T, R, W)), beta_history=np.repeat(rng.uniform(size=(T, S)), 2, axis=1),
                  accepted=rng.integers(0, 8, size=(T, W)), tsw_history=rng.uniform(size=(S, T - 1)),
                  smd_history=rng.uniform(size=(S, T - 1)), betas=np.array([1.0, 0.5, 0.0]))
    np.savez(tmp_path / "r.npz", format=np.array(FORMAT), iteration=np.array(R), **arrays)
    run = SavedRun(tmp_path / "r.npz")
    assert (run.nrows, run.nsweeps) == (R, S)
    bk = run.backend
    assert bk.iteration == R and len(bk) == T
    assert np.array_equal(bk.tsw_history, arrays["tsw_history"])
    assert np.array_equal(bk.smd_history, arrays["smd_history"])
    for t in (0, 2, -1):
        tt = t % T
        v = bk[t]
        assert np.array_equal(v.get_chain(), arrays["chain"][tt])
        assert np.array_equal(v.get_chain(discard=2, thin=3, flat=True), arrays["chain"][tt, 2::3].reshape(-1, D))
        assert np.array_equal(v.get_log_like(flat=True), arrays["log_like"][tt].ravel())
        assert np.array_equal(v.get_log_prob(discard=1), arrays["log_prob"][tt, 1:])
        assert np.array_equal(v.get_betas(), arrays["beta_history"][tt])
        assert np.array_equal(v.get_betas(discard=4), arrays["beta_history"][tt, 4:])
        assert np.array_equal(v.accepted, arrays["accepted"][tt])
    with pytest.raises(IndexError):
        bk[T]


def test_sampler_refuses_a_backend_object():
    from astroemperor_amd import PTSampler
    from tests.conftest import make_sim
    m = make_sim("51Peg", 1, parameterisation=1).model
    with pytest.raises(NotImplementedError):
        PTSampler(16, m.ndim__, m, ntemps=2, backend=object())
