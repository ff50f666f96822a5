"""Batched independent systems (one persistent launch per sweep) on the GPU.

* system 0 of a batch follows the single-system engine bit for bit (same
  seed, initial state, ladder and waves per walker), with epochs resident in
  LDS (N <= 512) and streamed (GJ876, N = 752, with the moving average);
* another system of the batch replays against the oracle's moves with the
  system index mixed into the random-number counters;
* the stored lnL of every walker of a ragged batch matches the oracle
  (FP64 1e-10 relative, mixed 1e-6 absolute);
* the autotuner leaves the chains where they were; graph replay equals
  direct launches; recorded samples equal the state they recorded.
"""
import numpy as np
import pytest

from tests.conftest import make_sim, oracle_for
from tests.test_gpu_pt import M32, _feistel, _philox, _Post, _u01

pytestmark = pytest.mark.gpu


def _models_small(n, seed=5):
    from astroemperor_amd import synth
    rng = np.random.default_rng(seed)
    return [synth._system(rng)[0] for _ in range(n)]


def _p0(models, T, W, seed):
    from astroemperor_amd.simulation import test_init
    rng = np.random.default_rng(seed)
    return [test_init(m, T, W, rng) for m in models]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["51Peg", "GJ876_ma"])
def test_system0_matches_single_engine(precision, case):
    from astroemperor_amd import PTSampler, SystemsPT
    if case == "51Peg":
        m0 = make_sim("51Peg", 1, parameterisation=1).model
    else:
        m0 = make_sim("GJ876", 2, parameterisation=1, moav={"order": 1, "global": False},
                      semantics="intended").model
    others = _models_small(2)
    T, W, seed, n = 4, 32, 0xfeed_beef_1234, 6
    p0 = _p0([m0] + others, T, W, 11)
    betas = np.array([1.0, 0.5, 0.2, 0.05])
    m0.set_waves_per_theta(1)
    try:
        ref = PTSampler(W, m0.ndim__, m0, ntemps=T, betas=betas, seed=seed, adapt_tau=20, adapt_nu=2,
                        precision=precision)
        ref.run_mcmc(p0[0], nsweeps=n)
        st = ref.get_last_sample()
    finally:
        m0.set_waves_per_theta(0)
    sysb = SystemsPT([m0] + others, T, W, p0, betas=betas, seed=seed, adapt_tau=20, adapt_nu=2,
                     precision=precision)
    sysb.set_tuning(waves_per_wg=16, waves_per_theta=1)
    sysb.run(n)
    x, ll, lp = sysb.state(0)
    assert np.array_equal(x, st.coords)
    assert np.array_equal(ll, st.log_like)
    assert np.array_equal(lp, st.log_prior)
    assert np.array_equal(sysb.betas[0], st.betas)
    assert np.array_equal(sysb.accepted[0], (ref.acceptance_fraction * n).round().astype(np.int64))
    assert np.allclose(sysb.get_tsw()[0], ref.get_tsw().mean(0), rtol=0, atol=1e-15)
    assert sysb.nsweeps == n


@pytest.mark.parametrize("mode", [1, 2])
def test_system0_ladder_modes_match_single_engine(mode):
    """The swap-distance and specific-heat ladders in the batched engine:
    system 0's ladder and state equal the single engine's bit for bit (the
    per-temperature spread is one wave per temperature in both)."""
    from astroemperor_amd import PTSampler, SystemsPT
    m0 = make_sim("51Peg", 1, parameterisation=1).model
    others = _models_small(2)
    T, W, seed, n = 5, 32, 0x5eed_0002, 8
    p0 = _p0([m0] + others, T, W, 13)
    betas = np.array([1.0, 0.6, 0.3, 0.1, 0.0])
    m0.set_waves_per_theta(1)
    try:
        ref = PTSampler(W, m0.ndim__, m0, ntemps=T, betas=betas, seed=seed, adapt_tau=20, adapt_nu=2,
                        adapt_mode=mode, precision="fp64")
        ref.run_mcmc(p0[0], nsweeps=n)
        st = ref.get_last_sample()
    finally:
        m0.set_waves_per_theta(0)
    sysb = SystemsPT([m0] + others, T, W, p0, betas=betas, seed=seed, adapt_tau=20, adapt_nu=2,
                     adapt_mode=mode, precision="fp64")
    sysb.set_tuning(waves_per_wg=16, waves_per_theta=1)
    sysb.run(n)
    x, ll, _ = sysb.state(0)
    assert np.array_equal(x, st.coords)
    assert np.array_equal(ll, st.log_like)
    assert np.array_equal(sysb.betas[0], st.betas)
    assert not np.array_equal(st.betas, betas)


def _replay_sweep_sys(o, X, lnl, lnp, betas, sweep, seed, sysmix, a=2.0):
    from oracle import emp_oracle as orc
    T, W, D = X.shape
    W2 = W // 2
    k0, k1 = seed & M32, (seed >> 32) & M32
    hi = (sweep >> 32) ^ sysmix
    post = _Post(o)
    for half in (0, 1):
        for t in range(T):
            active = list(range(half * W2, (half + 1) * W2))
            partner, uz, ua = [], [], []
            for k in active:
                wid = t * W + k
                r1 = _philox([wid, sweep & M32, hi, half], [k0, k1])
                r2 = _philox([wid, sweep & M32, hi, half | 0x80000000], [k0, k1])
                uz.append(_u01(r1[0], r1[1]))
                partner.append((1 - half) * W2 + int(_u01(r1[2], r1[3]) * W2))
                ua.append(_u01(r2[0], r2[1]))
            orc.stretch_halfstep(X[t], lnl[t], lnp[t], betas[t], active, partner, uz, ua, post, a)
    for h in range(T - 1, 0, -1):
        c = h - 1
        perm = [_feistel(i, W, k0 ^ (sweep & M32) ^ 0x5bd1e995,
                         k1 ^ ((h * 0x27d4eb2f) & M32) ^ hi) for i in range(W)]
        u = []
        for i in range(W):
            r = _philox([i, sweep & M32, hi, 0x40000000 | h], [k0, k1])
            u.append(_u01(r[0], r[1]))
        orc.swap_pair(X[h], lnl[h], lnp[h], X[c], lnl[c], lnp[c], betas[h], betas[c], perm, u)


def test_other_system_replays_oracle():
    from astroemperor_amd import SystemsPT
    models = _models_small(3, seed=8)
    T, W, seed, n = 3, 16, 0x1234_5678_9abc, 3
    betas = np.array([1.0, 0.3, 0.05])
    p0 = _p0(models, T, W, 12)
    s = SystemsPT(models, T, W, p0, betas=betas, seed=seed, adapt_mode=-1, precision="fp64",
                  record_every=1, record_temps=T, max_records=n)
    s.run(n, use_graph=False)
    sys_i = 2
    chain = s.get_chain(sys_i)                  # [n, T, W, D]
    o = oracle_for(models[sys_i])
    X = np.array(p0[sys_i])
    lnl = np.array([[o.my_likelihood(x) for x in X[t]] for t in range(T)])
    lnp = np.array([[o.my_prior(x) for x in X[t]] for t in range(T)])
    for sweep in range(n):
        _replay_sweep_sys(o, X, lnl, lnp, betas, sweep, seed, sys_i)
        assert np.allclose(chain[sweep], X, rtol=1e-12, atol=1e-12), sweep
    assert np.allclose(s.get_log_like(sys_i)[-1], lnl, rtol=1e-10, atol=0)


def _lpt_order(models):
    """The queue order the library builds: longest first by N (51 k + 46 (+23
    with an active MA)), stable."""
    cost = [m.data.ndat * (m.kplan__ * 51.0 + 46.0 +
                           (23.0 if m.block("MOAV") is not None and
                            (m.semantics == "intended" or m.block("MOAV").is_global) else 0.0))
            for m in models]
    return sorted(range(len(models)), key=lambda i: -cost[i])


@pytest.mark.parametrize("waves_per_wg", [8, 16])
def test_queue_pops_many_systems_per_workgroup(waves_per_wg):
    """More systems than persistent workgroups (one per compute unit, so each
    workgroup pops about three systems per sweep, reusing its LDS tile and
    slot memory): systems that sit past the first pop in the longest-first
    queue replay the oracle sweep by sweep, and every system's stored lnL
    matches the oracle."""
    from astroemperor_amd import SystemsPT, device_info
    n_cu = device_info()["n_cu"]
    S = 3 * n_cu + 17
    models = _models_small(S, seed=61)
    T, W, seed, n = 2, 16, 0x0dd_ba11_5eed, 3
    betas = np.array([1.0, 0.25])
    p0 = _p0(models, T, W, 14)
    s = SystemsPT(models, T, W, p0, betas=betas, seed=seed, adapt_mode=-1, precision="fp64",
                  record_every=1, record_temps=T, max_records=n)
    s.set_tuning(waves_per_wg=waves_per_wg, waves_per_theta=1, wgs_per_cu=1, order="lpt")
    s.run(n, use_graph=False)
    assert s.nsweeps == n
    order = _lpt_order(models)
    grid = min(S, n_cu)
    for pos in (grid, 2 * grid + 3, S - 1):          # second, third and last pops
        sys_i = order[pos]
        chain = s.get_chain(sys_i)
        o = oracle_for(models[sys_i])
        X = np.array(p0[sys_i])
        lnl = np.array([[o.my_likelihood(x) for x in X[t]] for t in range(T)])
        lnp = np.array([[o.my_prior(x) for x in X[t]] for t in range(T)])
        for sweep in range(n):
            _replay_sweep_sys(o, X, lnl, lnp, betas, sweep, seed, sys_i)
            assert np.allclose(chain[sweep], X, rtol=1e-12, atol=1e-12), (pos, sys_i, sweep)
        assert np.allclose(s.get_log_like(sys_i)[-1], lnl, rtol=1e-10, atol=0), (pos, sys_i)
    rng = np.random.default_rng(3)
    for sys_i in rng.choice(S, 24, replace=False):
        x, ll, _ = s.state(int(sys_i))
        o = oracle_for(models[sys_i])
        t, w = int(rng.integers(0, T)), int(rng.integers(0, W))
        want = o.my_likelihood(x[t, w])
        assert abs(ll[t, w] - want) <= 1e-10 * abs(want), (sys_i, ll[t, w], want)
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_lnl_matches_oracle(precision):
    from astroemperor_amd import SystemsPT
    models = _models_small(6, seed=21)
    T, W = 4, 16
    s = SystemsPT(models, T, W, _p0(models, T, W, 3), seed=5, precision=precision)
    s.run(4)
    rng = np.random.default_rng(0)
    tol = (1e-10, 0.0) if precision == "fp64" else (0.0, 1e-6)
    for i, m in enumerate(models):
        x, ll, lp = s.state(i)
        o = oracle_for(m)
        for t, w in zip(rng.integers(0, T, 12), rng.integers(0, W, 12)):
            want = o.my_likelihood(x[t, w])
            if np.isfinite(want):
                assert np.isclose(ll[t, w], want, rtol=tol[0], atol=tol[1]), (i, t, w, ll[t, w], want)
            wp = o.my_prior(x[t, w])
            assert abs(lp[t, w] - wp) <= 2 * np.spacing(abs(wp)) or lp[t, w] == wp


def test_autotune_keeps_chains_and_graph_equals_direct():
    from astroemperor_amd import SystemsPT
    models = _models_small(40, seed=33)
    T, W = 4, 32
    p0 = _p0(models, T, W, 4)
    a = SystemsPT(models, T, W, p0, seed=9)
    a.run(2)
    best = a.autotune(trial_sweeps=1)
    assert best["ms_per_sweep"] > 0 and a.nsweeps == 2
    a.run(3, use_graph=True)
    b = SystemsPT(models, T, W, p0, seed=9)
    b.run(2, use_graph=False)            # the default shape, as a's first sweeps
    b.set_tuning(best["waves_per_wg"], best["waves_per_theta"], best["wgs_per_cu"], best["order"])
    b.run(3, use_graph=False)
    assert a.nsweeps == b.nsweeps == 5
    for i in (0, 17, 39):
        for u, v in zip(a.state(i), b.state(i)):
            assert np.array_equal(u, v), i
    assert np.array_equal(a.betas, b.betas)
    assert np.array_equal(a.accepted, b.accepted)
    assert np.array_equal(a.get_tsw(), b.get_tsw())


def test_records_equal_state():
    from astroemperor_amd import SystemsPT
    models = _models_small(5, seed=44)
    T, W = 3, 16
    s = SystemsPT(models, T, W, _p0(models, T, W, 6), seed=2, record_every=2, record_temps=2, max_records=3)
    s.run(5)
    assert s.nsweeps == 5
    for i in range(5):
        ch = s.get_chain(i)
        assert ch.shape == (2, 2, W, models[i].ndim__)
    s.run(1)                                    # sweep 6 writes record 3
    for i in range(5):
        ch, ll = s.get_chain(i), s.get_log_like(i)
        x, l, _ = s.state(i)
        assert ch.shape[0] == 3 and np.array_equal(ch[-1], x[:2]) and np.array_equal(ll[-1], l[:2])
    s.run(4)                                    # capacity reached: records stop at 3
    assert s.get_chain(0).shape[0] == 3
    assert s.n_lnl_evaluated > 0
    assert s.log_like().shape == (5, T, W) and s.get_tsw().shape == (5, T - 1)
