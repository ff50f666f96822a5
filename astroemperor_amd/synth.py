"""Synthetic workloads with fixed seeds (numpy PCG64(20261202 + id)).

Each builder returns a dict with the time-sorted ``RVData``, a ready
``ReddModel`` (blocks, priors and conditions set the way a user would set
them on ``Simulation``), the true free parameter vector and the sampler
setup ``(ntemps, nwalkers, nsweeps, nsteps)``.

* ``c1``: 1 Keplerian, 1 instrument, N = 256 over 3000 d (BJD 2450000+);
  P = 4.2308 d, K = 55.9 m/s, e = 0.01, w = 1.0, M0 = 0.5, offset -3 m/s,
  sigma ~ U(1, 3) plus 2 m/s jitter; parameterisation 1; uniform priors;
  5 x 100 x 500.
* ``c2``: 3 Keplerians, 4 instruments with 80/150/400/900 epochs over
  6000 d; ln P uniform in [ln 2, ln 3000], K in [1, 50], e ~ Beta(0.867,
  3.03); per-instrument offset, jitter ~ U(0, 5) and MA(1) coefficient
  ~ U(-0.5, 0.5) with a fixed 5 d timescale (intended MA semantics);
  parameterisation 6; 10 x 256 x 2000.
* ``c4``: 5 Keplerians, N = 2000, 2 instruments, parameterisation 1.
* ``c5``: 7 Keplerians, N = 10000 over 5 instruments with MA(1);
  16 x 1024 x 5000.
* ``c3``: batched uneven systems: S = 4096 independent systems, N_s =
  clip(round(LogNormal(ln 200, 0.6)), 50, 2000) epochs over 3000 d, k_s ~
  U{1..4} Keplerians (ln P ~ U(ln 2, ln 1000), K ~ U(1, 50), e ~ Beta(0.867,
  3.03)), I_s ~ U{1..3} instruments sharing the epochs by a Dirichlet(1)
  draw (at least 5 each); parameterisation 1 with the uniform priors of
  ``c1``; each system 8 x 64 walkers.
"""
from __future__ import annotations

import numpy as np

from .blocks import JitterBlock, KeplerianBlock, MOAVBlock, OffsetBlock, SmartSetter
from .data import RVData
from .model import ReddModel

BASE_SEED = 20261202


def kepler_rv(t, P, K, e, w, M0, t0):
    """Noise-free RV of one planet, used only to synthesise data."""
    M = 2 * np.pi / P * (t - t0) + M0
    E = M.copy()
    for _ in range(60):
        E = E - (E - e * np.sin(E) - M) / (1 - e * np.cos(E))
    f = 2 * np.arctan2(np.sqrt(1 + e) * np.sin(E / 2), np.sqrt(1 - e) * np.cos(E / 2))
    return K * (np.cos(f + w) + e * np.cos(w))


def _instruments(rng, counts, span):
    bjd = [np.sort(2450000.0 + rng.uniform(0, span, n)) for n in counts]
    sig = [rng.uniform(1, 3, n) for n in counts]
    return bjd, sig


def _ma_noise(rng, t, sig, jit, coef, tau):
    """r_i = eps_i + coef exp(-dt/tau) eps_{i-1}, eps_i ~ N(0, sig^2 + jit^2)."""
    eps = rng.normal(0, np.sqrt(sig ** 2 + jit ** 2))
    r = eps.copy()
    r[1:] += coef * np.exp(-np.diff(t) / tau) * eps[:-1]
    return r


def _build(blocks_spec, data, conds, semantics):
    ss = SmartSetter(data.as_table())
    blocks = []
    for kind, arg in blocks_spec:
        if kind == "kep":
            b = KeplerianBlock(arg["param"], arg["n"])
            ss(b, prargs=[[0, 1], [0, 0.1]])
        elif kind == "off":
            b = OffsetBlock(data.nins)
            ss(b)
        elif kind == "jit":
            b = JitterBlock(data.nins)
            ss(b, [0, 1], [5, 5])
        elif kind == "ma":
            b = MOAVBlock(data.nins, arg)
            ss(b)
        blocks.append(b)
    for name, attr, val in conds:
        hit = False
        for b in blocks:
            for p in b:
                if p.name == name:
                    setattr(p, attr, val)
                    hit = True
        if not hit:
            raise KeyError(name)
    return ReddModel(blocks, data, semantics=semantics)


def c1():
    rng = np.random.default_rng(BASE_SEED + 1)
    (bjd,), (sig,) = _instruments(rng, [256], 3000.0)
    P, K, e, w, M0 = 4.2308, 55.9, 0.01, 1.0, 0.5
    rv = kepler_rv(bjd, P, K, e, w, M0, bjd.min()) - 3.0 + rng.normal(0, np.sqrt(sig ** 2 + 2.0 ** 2))
    data = RVData.from_arrays([bjd], [rv], [sig], labels=["synth_c1"])
    model = _build([("kep", {"param": 1, "n": 1}), ("off", None), ("jit", None)], data,
                   [("Jitter 1", "prior", "Uniform")], "reference")
    # truth in the model's frame: the loader subtracts the mean of the raw RVs
    S, C = np.sqrt(e) * np.sin(w), np.sqrt(e) * np.cos(w)
    truth = np.array([P, K, M0, S, C, -3.0 - float(np.mean(rv)), 2.0])
    return dict(data=data, model=model, truth=truth, setup=(5, 100, 500, 1))


def _multi(cid, kplan, counts, span, param, setup, ma=True):
    rng = np.random.default_rng(BASE_SEED + cid)
    bjd, sig = _instruments(rng, counts, span)
    t0 = min(b.min() for b in bjd)
    planets = []
    for _ in range(kplan):
        P = float(np.exp(rng.uniform(np.log(2), np.log(3000))))
        K = float(rng.uniform(1, 50))
        e = float(rng.beta(0.867, 3.03))
        w = float(rng.uniform(0, 2 * np.pi))
        M0 = float(rng.uniform(0, 2 * np.pi))
        planets.append((P, K, e, w, M0))
    I = len(counts)
    gam = rng.uniform(-10, 10, I)
    jit = rng.uniform(0, 5, I)
    coef = rng.uniform(-0.5, 0.5, I)
    rvs = []
    for i in range(I):
        mu = sum(kepler_rv(bjd[i], *pl, t0) for pl in planets) + gam[i]
        rvs.append(mu + _ma_noise(rng, bjd[i], sig[i], jit[i], coef[i] if ma else 0.0, 5.0))
    data = RVData.from_arrays(bjd, rvs, sig, labels=[f"ins{i + 1}" for i in range(I)])
    blocks = [("kep", {"param": param, "n": n + 1}) for n in range(kplan)] + [("off", None), ("jit", None)]
    if ma:
        blocks.append(("ma", {"order": 1, "global": False}))
    conds = []
    for n in range(1, kplan + 1):
        if param == 6:
            conds += [(f"lPeriod {n}", "limits", [float(np.log(2)), float(np.log(3000))]),
                      (f"Amplitude {n}", "limits", [1.0, 50.0]),
                      (f"Eccentricity {n}", "prior", "Beta"),
                      (f"Eccentricity {n}", "prargs", [0.867, 3.03])]
        else:
            conds += [(f"Period {n}", "limits", [2.0, 3000.0]),
                      (f"Amplitude {n}", "limits", [1.0, 50.0])]
    for i in range(1, I + 1 if ma else 1):
        conds += [(f"MACoefficient {i} Order 1", "limits", [-0.5, 0.5]),
                  (f"MATimescale {i} Order 1", "fixed", 5.0)]
    model = _build(blocks, data, conds, "intended")
    truth = []
    for (P, K, e, w, M0) in planets:
        if param == 6:
            truth += [np.log(P), K, M0, e, w]
        else:
            truth += [P, K, M0, np.sqrt(e) * np.sin(w), np.sqrt(e) * np.cos(w)]
    # offsets relative to each instrument's mean-subtracted data
    truth += list(gam - np.array([np.mean(r) for r in rvs]))
    truth += list(jit) + (list(coef) if ma else [])
    return dict(data=data, model=model, truth=np.array(truth), setup=setup, planets=planets)


def c2():
    return _multi(2, 3, [80, 150, 400, 900], 6000.0, 6, (10, 256, 2000, 1))


def c4():
    return _multi(4, 5, [1000, 1000], 6000.0, 1, (1, 65536, 0, 1), ma=False)


def c5():
    return _multi(5, 7, [1000, 1500, 2000, 2500, 3000], 6000.0, 6, (16, 1024, 5000, 1))


def _system(rng, span=3000.0):
    N = int(np.clip(np.round(rng.lognormal(np.log(200.0), 0.6)), 50, 2000))
    k = int(rng.integers(1, 5))
    I = int(rng.integers(1, 4))
    share = rng.dirichlet(np.ones(I))
    counts = 5 + rng.multinomial(N - 5 * I, share)
    bjd, sig = _instruments(rng, list(counts), span)
    t0 = min(b.min() for b in bjd)
    planets = []
    for _ in range(k):
        planets.append((float(np.exp(rng.uniform(np.log(2), np.log(1000)))), float(rng.uniform(1, 50)),
                        float(rng.beta(0.867, 3.03)), float(rng.uniform(0, 2 * np.pi)),
                        float(rng.uniform(0, 2 * np.pi))))
    gam = rng.uniform(-10, 10, I)
    jit = rng.uniform(0, 5, I)
    rvs = []
    for i in range(I):
        mu = sum(kepler_rv(bjd[i], *pl, t0) for pl in planets) + gam[i]
        rvs.append(mu + rng.normal(0, np.sqrt(sig[i] ** 2 + jit[i] ** 2)))
    data = RVData.from_arrays(bjd, rvs, sig, labels=[f"ins{i + 1}" for i in range(I)])
    blocks = [("kep", {"param": 1, "n": n + 1}) for n in range(k)] + [("off", None), ("jit", None)]
    conds = [(f"Jitter {i + 1}", "prior", "Uniform") for i in range(I)]
    model = _build(blocks, data, conds, "reference")
    truth = []
    for (P, K, e, w, M0) in planets:
        truth += [P, K, M0, np.sqrt(e) * np.sin(w), np.sqrt(e) * np.cos(w)]
    truth += list(gam - np.array([np.mean(r) for r in rvs])) + list(jit)
    return model, np.array(truth)


def c3(n_systems=4096):
    rng = np.random.default_rng(BASE_SEED + 3)
    models, truths = [], []
    for _ in range(n_systems):
        m, t = _system(rng)
        models.append(m)
        truths.append(t)
    return dict(models=models, truths=truths, setup=(8, 64, 1000, 1))


def initial_positions(model, ntemps, nwalkers, rng):
    """[T, W, D] uniform in the prior box, with every (S, C) pair of a
    parameterisation 1/2/4/7 Keplerian inside the unit disk (e < 1)."""
    lim = model.limits_free()
    D = len(lim)
    x = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(size=(ntemps, nwalkers, D))
    names = model.names_free()
    for j, nm in enumerate(names):
        if nm.startswith("Ecc_sin") and j + 1 < D and names[j + 1].startswith("Ecc_cos"):
            r = np.sqrt(rng.uniform(size=(ntemps, nwalkers)))
            th = rng.uniform(0, 2 * np.pi, size=(ntemps, nwalkers))
            x[..., j], x[..., j + 1] = 0.99 * r * np.sin(th), 0.99 * r * np.cos(th)
    return x


def prior_draws(model, n, rng):
    """Draws inside the prior box of every free parameter (uniform over the
    limits), re-drawing those with a -inf prior (S^2 + C^2 > 1 etc.)."""
    lim = model.limits_free()
    out = lim[:, 0] + (lim[:, 1] - lim[:, 0]) * rng.uniform(size=(n, len(lim)))
    return out


CONFIGS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c5": c5}
