"""Reconstructed reference CPU path, timed as bench.py's cpu_baseline.

TEST/BENCH INFRASTRUCTURE ONLY (oracle package): imported by bench.py's
cpu_baseline leg, never by the product.

The reference samples with reddemcee on a ``multiprocessing.Pool(ncores)``
(support/pools/01.pool:1-12): every stretch half-step of every temperature
maps the per-walker log-probability over the pool, and each worker evaluates
the generated numpy ``my_prior``/``my_likelihood`` (emp/emp.py:182-254,
support/likelihoods/00.like) with ``kepler.solve`` in C++.  kepler.py and
reddemcee are not available here, so this restatement uses:

* the oracle's numpy ``OracleModel`` (the generated script's arithmetic) with
  the oracle's C Kepler solver in place of kepler.solve;
* an emcee-style red/blue stretch driver (a = 2) with adjacent swaps in numpy;
* ``multiprocessing.Pool(ncores)`` (fork) and one ``pool.map`` per half-step
  per temperature, the reference's IPC shape.
"""
from __future__ import annotations

import multiprocessing as mp
import os
import platform
import time

import numpy as np

from .emp_oracle import OracleModel

_MODEL = None


def _init(spec, X, Y, YERR, FLAG):
    global _MODEL
    _MODEL = OracleModel(spec, X, Y, YERR, FLAG)


def _logpost(args):
    theta, beta = args
    lp = _MODEL.my_prior(theta)
    if not np.isfinite(lp):
        return -np.inf, -np.inf
    ll = _MODEL.my_likelihood(theta)
    if not np.isfinite(ll):
        ll = -np.inf
    return lp, ll


def cpu_model_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run(spec, X, Y, YERR, FLAG, p0, betas, nsweeps, ncores=None, seed=0, a=2.0):
    """Runs `nsweeps` PT sweeps; returns timing and the final state."""
    ncores = ncores or os.cpu_count()
    T, W, D = p0.shape
    W2 = W // 2
    rng = np.random.default_rng(seed)
    X_ = p0.copy()
    ctx = mp.get_context("fork")
    with ctx.Pool(ncores, initializer=_init, initargs=(spec, X, Y, YERR, FLAG)) as pool:
        # initial evaluation (not timed)
        res = pool.map(_logpost, [(X_[t, w], betas[t]) for t in range(T) for w in range(W)])
        lnp = np.array([r[0] for r in res]).reshape(T, W)
        lnl = np.array([r[1] for r in res]).reshape(T, W)
        n_maps = 0
        t0 = time.perf_counter()
        for _ in range(nsweeps):
            for t in range(T):
                for half in (0, 1):
                    act = np.arange(half * W2, (half + 1) * W2)
                    comp = np.arange((1 - half) * W2, (2 - half) * W2)
                    z = ((a - 1.0) * rng.uniform(size=W2) + 1.0) ** 2 / a
                    j = comp[rng.integers(0, W2, size=W2)]
                    Yp = X_[t, j] + z[:, None] * (X_[t, act] - X_[t, j])
                    out = pool.map(_logpost, [(y, betas[t]) for y in Yp])
                    n_maps += 1
                    lp_new = np.array([o[0] for o in out])
                    ll_new = np.array([o[1] for o in out])
                    b = betas[t]
                    with np.errstate(invalid="ignore"):
                        dl = np.where(b == 0, 0.0, b * (ll_new - lnl[t, act]))
                    lnalpha = (D - 1) * np.log(z) + dl + lp_new - lnp[t, act]
                    ok = np.isfinite(lp_new) & (np.log(rng.uniform(size=W2)) < lnalpha)
                    X_[t, act[ok]] = Yp[ok]
                    lnl[t, act[ok]] = ll_new[ok]
                    lnp[t, act[ok]] = lp_new[ok]
            for h in range(T - 1, 0, -1):
                c = h - 1
                perm = rng.permutation(W)
                lr = (betas[c] - betas[h]) * (lnl[h] - lnl[c, perm])
                ok = np.log(rng.uniform(size=W)) < lr
                hi, ci = np.where(ok)[0], perm[ok]
                X_[h, hi], X_[c, ci] = X_[c, ci].copy(), X_[h, hi].copy()
                lnl[h, hi], lnl[c, ci] = lnl[c, ci].copy(), lnl[h, hi].copy()
                lnp[h, hi], lnp[c, ci] = lnp[c, ci].copy(), lnp[h, hi].copy()
        dt = time.perf_counter() - t0
    return dict(seconds=dt, sweeps=nsweeps, sweeps_per_s=nsweeps / dt,
                lnl_evals_per_s=nsweeps * T * W / dt, cores=ncores, pool_maps=n_maps,
                cpu_model=cpu_model_name(), os_cpu_count=os.cpu_count(), final=(X_, lnl, lnp))
