"""The gradient move (ensemble-preconditioned Metropolis-adjusted Langevin) and
the move mix on the device.

* replay: after stretch sweeps have brought the ensemble near the posterior,
  gradient sweeps (alone, and mixed 4:1 with stretch sweeps) are replayed on
  the host: the device's random numbers regenerated from the documented Philox
  counters, the oracle's langevin_halfstep with lnL from the oracle and grad lnL
  from the FP64 torch reference; positions, lnL, ln prior and acceptances must
  agree;
* tempered Gaussian: the offsets-only model, whose posterior at inverse
  temperature beta is N(m, v / beta), sampled by the gradient move alone
  (started at stationarity) and by the 1:4 mix (from a dispersed start);
* graph replay equals direct launches, and checkpoint/restore resumes bitwise
  with the gradient move in the mix.
"""
import numpy as np
import pytest
import torch

from tests.conftest import make_sim, oracle_for
from tests.test_gpu_pt import M32, _feistel, _gaussian_model, _philox, _u01
from tests.torch_reference import lnl_torch

pytestmark = pytest.mark.gpu


class _PostG:
    def __init__(self, o, model):
        self.o, self.model = o, model

    def prior(self, th):
        return self.o.my_prior(th)

    def like(self, th):
        return self.o.my_likelihood(th)

    def grad(self, th):
        t = torch.tensor(th, dtype=torch.float64, requires_grad=True)
        lnl_torch(self.model, t).backward()
        return t.grad.numpy()


def _move_of(sweep, seed, w_stretch, w_grad):
    if w_grad <= 0:
        return 0
    if w_stretch <= 0:
        return 1
    r = _philox([0, sweep & M32, sweep >> 32, 0x10000000], [seed & M32, (seed >> 32) & M32])
    return 1 if _u01(r[0], r[1]) * (w_stretch + w_grad) < w_grad else 0


def _normals(t, W, k, sweep, half, D, key):
    out = np.empty(D)
    for j in range(D):
        r = _philox([t * W + k, sweep & M32, sweep >> 32, 0x20000000 | (half << 8) | (j >> 1)], key)
        rad = np.sqrt(-2.0 * np.log(_u01(r[0], r[1])))
        ang = 2.0 * np.pi * _u01(r[2], r[3])
        out[j] = rad * (np.sin(ang) if j & 1 else np.cos(ang))
    return out


def _replay(o, post, X, lnl, lnp, betas, sweep, seed, move, eps, a=2.0):
    from oracle import emp_oracle as orc
    T, W, D = X.shape
    W2 = W // 2
    key = [seed & M32, (seed >> 32) & M32]
    acc = np.zeros((T, W), dtype=int)
    if move == 1:
        grad = np.array([[post.grad(x) for x in X[t]] for t in range(T)])
    for half in (0, 1):
        for t in range(T):
            active = list(range(half * W2, (half + 1) * W2))
            other = list(range((1 - half) * W2, (2 - half) * W2))
            if move == 1:
                xi = np.array([_normals(t, W, k, sweep, half, D, key) for k in active])
                ua = []
                for k in active:
                    r = _philox([t * W + k, sweep & M32, sweep >> 32, 0x20000000 | (half << 8) | 255], key)
                    ua.append(_u01(r[0], r[1]))
                got = orc.langevin_halfstep(X[t], lnl[t], lnp[t], grad[t], betas[t], active, other, xi, ua,
                                            post, eps[t] if np.ndim(eps) else eps)
            else:
                partner, uz, ua = [], [], []
                for k in active:
                    wid = t * W + k
                    r1 = _philox([wid, sweep & M32, sweep >> 32, half], key)
                    r2 = _philox([wid, sweep & M32, sweep >> 32, half | 0x80000000], key)
                    uz.append(_u01(r1[0], r1[1]))
                    partner.append((1 - half) * W2 + int(_u01(r1[2], r1[3]) * W2))
                    ua.append(_u01(r2[0], r2[1]))
                got = orc.stretch_halfstep(X[t], lnl[t], lnp[t], betas[t], active, partner, uz, ua, post, a)
            acc[t, active] += got
    for h in range(T - 1, 0, -1):
        c = h - 1
        perm = [_feistel(i, W, key[0] ^ (sweep & M32) ^ 0x5bd1e995,
                         key[1] ^ ((h * 0x27d4eb2f) & M32) ^ (sweep >> 32)) for i in range(W)]
        u = []
        for i in range(W):
            r = _philox([i, sweep & M32, sweep >> 32, 0x40000000 | h], key)
            u.append(_u01(r[0], r[1]))
        orc.swap_pair(X[h], lnl[h], lnp[h], X[c], lnl[c], lnp[c], betas[h], betas[c], perm, u)
    return acc


@pytest.mark.parametrize("mix", ["gradient", "mixed_1_4"])
def test_gradient_move_replay_matches_oracle(mix):
    from astroemperor_amd import GradientMove, PTSampler, StretchMove
    from astroemperor_amd.simulation import test_init
    sim = make_sim("51Peg", 1, parameterisation=1,
                   conditions=[("Period 1", "limits", [4.0, 4.5]), ("Amplitude 1", "limits", [40, 70])])
    m = sim.model
    T, W, seed = 3, 32, 0x2468_ace0_1357 + (mix == "gradient")
    # one step for every temperature, or one per temperature
    eps = 0.5 if mix == "gradient" else np.array([0.45, 0.6, 0.8])
    betas = np.array([1.0, 0.3, 0.05])
    ws, wg = (0.0, 1.0) if mix == "gradient" else (4.0, 1.0)
    burn, S = 400, 6
    s = PTSampler(W, m.ndim__, m, ntemps=T, betas=betas, adapt_mode=-1, seed=seed)
    s.run_mcmc(test_init(m, T, W, np.random.default_rng(5)), nsweeps=burn, capacity=burn + S)
    moves = [(GradientMove(eps), 1.0)] if mix == "gradient" else [(StretchMove(), ws), (GradientMove(eps), wg)]
    s.set_moves(moves)
    acc0 = s._get(6, np.int64)[..., 0, 0].copy()
    X = s.get_last_sample().coords.copy()
    s.run_mcmc(None, nsweeps=S, use_graph=(mix == "gradient"))
    chain = s.get_chain()
    acc1 = s._get(6, np.int64)[..., 0, 0]
    o = oracle_for(m)
    post = _PostG(o, m)
    lnl = np.array([[o.my_likelihood(x) for x in X[t]] for t in range(T)])
    lnp = np.array([[o.my_prior(x) for x in X[t]] for t in range(T)])
    acc = np.zeros((T, W), dtype=int)
    kinds = []
    for sweep in range(burn, burn + S):
        mv = _move_of(sweep, seed, ws, wg)
        kinds.append(mv)
        acc += _replay(o, post, X, lnl, lnp, betas, sweep, seed, mv, eps)
        assert np.allclose(chain[:, sweep], X, rtol=1e-9, atol=1e-9), (sweep, mv)
    assert np.array_equal(acc1 - acc0, acc), (acc1 - acc0, acc)
    assert np.allclose(s.get_log_like()[:, -1], lnl, rtol=1e-10, atol=0)
    grad_acc = sum(1 for k in kinds if k == 1)
    assert grad_acc >= 1, kinds
    if mix == "gradient":
        assert acc.sum() > 0
    else:
        assert np.array_equal(s.gradient_steps, eps)


def test_gradient_step_adaptation_reaches_target():
    """adapt_gradient_steps on a tempered Gaussian: after the burn-in the
    acceptance of gradient sweeps at every temperature sits near the target,
    and the move mix comes back with those steps held fixed."""
    from astroemperor_amd import GradientMove, PTSampler, StretchMove
    rng = np.random.default_rng(3)
    m, mean, var = _gaussian_model(rng)
    betas = np.array([1.0, 0.4, 0.1, 0.02])
    T, W = len(betas), 128
    p0 = mean + rng.normal(size=(T, W, 3)) * np.sqrt(var / betas[:, None, None])
    s = PTSampler(W, 3, m, ntemps=T, betas=betas, adapt_mode=-1, seed=21,
                  moves=[(StretchMove(), 4.0), (GradientMove(8.0), 1.0)])
    s.run_mcmc(p0, nsweeps=4, capacity=200)
    steps = s.adapt_gradient_steps(nsweeps=40, target=0.5, batch=2)
    assert steps.shape == (T,) and np.array_equal(s.gradient_steps, steps)
    kinds = [type(mv).__name__ for mv, _ in s.moves]
    assert kinds == ["StretchMove", "GradientMove"] and [w for _, w in s.moves] == [4.0, 1.0]
    s.set_moves([(GradientMove(steps), 1.0)])
    a0 = s.accepted.sum(1)
    s.run_mcmc(None, nsweeps=20)
    acc = (s.accepted.sum(1) - a0) / (20.0 * W)
    print(f"steps {steps}, acceptance {acc}", flush=True)
    assert np.all(np.abs(acc - 0.5) < 0.15), acc
    s.close()


def test_gradient_move_gaussian_moments():
    """Gradient move alone, started at stationarity, then the 1:4 mix from a
    dispersed start: both keep N(m, v / beta) at every temperature."""
    from astroemperor_amd import GradientMove, PTSampler, StretchMove
    rng = np.random.default_rng(7)
    m, mean, var = _gaussian_model(rng)
    betas = np.array([1.0, 0.5, 0.2, 0.05])
    T, W, S, burn = len(betas), 64, 3000, 300
    for label, moves, p0 in [
        ("gradient", [(GradientMove(), 1.0)],
         mean + rng.normal(size=(T, W, 3)) * np.sqrt(var / betas[:, None, None])),
        ("mixed", [(StretchMove(), 4.0), (GradientMove(), 1.0)],
         mean + rng.normal(size=(T, W, 3)) * np.sqrt(var) * 3),
    ]:
        s = PTSampler(W, 3, m, ntemps=T, betas=betas, adapt_mode=-1, seed=11, moves=moves)
        s.run_mcmc(p0, nsweeps=S)
        chain = s.get_chain(discard=burn)
        for t, b in enumerate(betas):
            x = chain[t].reshape(-1, 3)
            assert np.all(np.abs(x.mean(0) - mean) < 6 * np.sqrt(var / b / 3000)), (label, t, x.mean(0), mean)
            assert np.all(np.abs(x.var(0) / (var / b) - 1) < 0.08), (label, t, x.var(0) / (var / b))
        af = s.acceptance_fraction
        print(f"{label}: acceptance {af.mean():.3f}", flush=True)
        assert np.all(af > 0.2)
        s.close()


def test_gradient_mix_graph_equals_direct_and_resumes(tmp_path):
    from astroemperor_amd import GradientMove, PTSampler, StretchMove
    from astroemperor_amd.simulation import test_init
    sim = make_sim("51Peg", 1, parameterisation=1,
                   conditions=[("Period 1", "limits", [4.0, 4.5]), ("Amplitude 1", "limits", [40, 70])])
    m = sim.model
    T, W = 2, 32
    p0 = test_init(m, T, W, np.random.default_rng(3))
    moves = [(StretchMove(), 1.0), (GradientMove(0.3), 1.0)]
    out = []
    for use_graph in (True, False):
        s = PTSampler(W, m.ndim__, m, ntemps=T, seed=77, moves=moves)
        s.run_mcmc(p0, nsweeps=24, use_graph=use_graph, capacity=34)
        out.append((s.get_chain(), s.get_log_like(), s.acceptance_fraction, s.n_lnl_evaluated))
        if use_graph:
            s.checkpoint(tmp_path / "ck.bin")
            s.run_mcmc(None, nsweeps=10)
            s2 = PTSampler(W, m.ndim__, m, ntemps=T, seed=77, moves=moves)
            s2.run_mcmc(p0, nsweeps=0, capacity=34)
            s2.restore(tmp_path / "ck.bin")
            s2.run_mcmc(None, nsweeps=10)
            # a fresh engine restored from the checkpoint holds the whole
            # history, the rows before the checkpoint included
            assert s2.iteration == s.iteration == 34
            for get in ("get_chain", "get_log_like", "get_log_prior", "get_betas", "get_tsw", "get_smd"):
                a, b = getattr(s, get)(), getattr(s2, get)()
                assert np.isfinite(b).all() and np.array_equal(a, b), get
            assert np.array_equal(s.lnl_moments(4)[0], s2.lnl_moments(4)[0])
            s2.close()
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
    assert out[0][3] == out[1][3]
