"""Edge cases of the GPU path against the oracle: the largest models the
kernels take (8 Keplerians, 16 instruments, 104 parameters), tiny data
(instruments of two epochs, fewer epochs than a wave), epochs tied in time
under the global MA, a long series (30000 epochs, 59 tiles), batches of 0, 1
and a count that leaves a workgroup part-filled, and the smallest PT
ensembles (one temperature; two walkers)."""
import numpy as np
import pytest

from astroemperor_amd import synth
from tests.conftest import oracle_for
from tests.test_gpu_lnl import _check_lnl, _draws

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wpt", [1, 8])
def test_largest_model(wpt):
    m = synth._multi(90, 8, [25] * 16, 3000.0, 1, (1, 2, 1, 1), ma=True)["model"]
    assert m.kplan__ == 8 and m.nins__ == 16 and m.ndim_full == 104
    m.set_waves_per_theta(wpt)
    e, dm = _check_lnl(m, _draws(m, 24, np.random.default_rng(wpt)), f"8 planets 16 ins wpt{wpt}")
    print(f"largest model wpt={wpt}: fp64 rel {e:.2e}, |mixed - fp64| {dm:.2e}", flush=True)
    assert dm < 1e-6


def test_tiny_data():
    m = synth._multi(91, 1, [2, 2, 3], 500.0, 1, (1, 2, 1, 1), ma=True)["model"]
    assert m.data.ndat == 7
    for wpt in (1, 2, 4, 8):
        m.set_waves_per_theta(wpt)
        _check_lnl(m, _draws(m, 16, np.random.default_rng(wpt)), f"tiny wpt{wpt}")


def test_tied_epochs_global_ma():
    from astroemperor_amd import JitterBlock, KeplerianBlock, MOAVBlock, OffsetBlock, ReddModel, RVData, SmartSetter
    rng = np.random.default_rng(3)
    t = np.sort(2450000.0 + rng.uniform(0, 800, 120))
    t2 = t[::3].copy()                                  # instrument 2 observes at instrument 1's times
    rv = [rng.normal(0, 5, len(t)), rng.normal(0, 5, len(t2))]
    err = [rng.uniform(1, 3, len(t)), rng.uniform(1, 3, len(t2))]
    data = RVData.from_arrays([t, t2], rv, err)
    assert len(np.unique(data.X)) < data.ndat
    ss = SmartSetter(data.as_table())
    blocks = []
    for b, args in ((KeplerianBlock(1, 1), dict(prargs=[[0, 1], [0, 0.1]])), (OffsetBlock(2), {}),
                    (JitterBlock(2), {}), (MOAVBlock(2, {"order": 2, "global": True}), {})):
        ss(b, **args)
        blocks.append(b)
    m = ReddModel(blocks, data, semantics="reference")
    for wpt in (1, 4):
        m.set_waves_per_theta(wpt)
        _check_lnl(m, _draws(m, 16, np.random.default_rng(wpt)), f"ties wpt{wpt}")


def test_long_series():
    m = synth._multi(92, 2, [30000], 20000.0, 1, (1, 2, 1, 1), ma=True)["model"]
    assert m.data.ndat == 30000
    o = oracle_for(m)
    th = _draws(m, 4, np.random.default_rng(0))
    ref = np.array([o.my_likelihood(x) for x in th])
    for wpt in (1, 8):
        m.set_waves_per_theta(wpt)
        got = m.lnl_batch(th, "fp64")
        fin = np.isfinite(ref)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.max(np.abs(got[fin] - ref[fin]) / np.abs(ref[fin])) <= 1e-10
        assert np.max(np.abs(m.lnl_batch(th, "mixed")[fin] - got[fin])) < 1e-6


def test_batch_sizes():
    m = synth.c1()["model"]
    o = oracle_for(m)
    th = _draws(m, 13, np.random.default_rng(1))
    assert m.lnl_batch(th[:0]).shape == (0,)
    for n in (1, 13):
        got = m.lnl_batch(th[:n])
        ref = np.array([o.my_likelihood(x) for x in th[:n]])
        fin = np.isfinite(ref)
        assert np.allclose(got[fin], ref[fin], rtol=1e-10, atol=0)
        lp = m.lnprior_batch(th[:n])
        assert np.allclose(lp, [o.my_prior(x) for x in th[:n]], rtol=1e-14, atol=0)


@pytest.mark.parametrize("T,W", [(1, 16), (3, 2)])
def test_smallest_ensembles(T, W):
    from astroemperor_amd import PTSampler
    from astroemperor_amd.simulation import test_init
    m = synth.c1()["model"]
    s = PTSampler(W, m.ndim__, m, ntemps=T, seed=5)
    s.run_mcmc(test_init(m, T, W, np.random.default_rng(0)), nsweeps=5)
    st = s.get_last_sample()
    o = oracle_for(m)
    for t in range(T):
        for w in range(W):
            assert np.isclose(st.log_like[t, w], o.my_likelihood(st.coords[t, w]), rtol=1e-10, atol=0)
    assert s.get_tsw().shape == (5, T - 1)
