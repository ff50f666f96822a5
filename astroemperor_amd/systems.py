"""Many independent systems, one PT ensemble each, advanced together.

The reference fits one star per run: it builds ``reddemcee.PTSampler`` over
the star's ``my_likelihood``/``my_prior`` (emp/emp.py:576-593) and runs
``run_mcmc`` (support/endit_reddemcee.scr:3).  A survey of S stars is S such
runs.  :class:`SystemsPT` runs all of them at once on the device: S models
(any mix of epochs, instruments, planets and free parameters) with the same
T x W ensemble shape, one kernel launch per sweep for the whole batch.

Per system the moves are those of :class:`~astroemperor_amd.sampler.PTSampler`
(stretch halves, hottest-first adjacent swaps, ladder adaptation in every mode),
and system 0 of a batch follows exactly the trajectory the single-system
sampler would with the same seed, waves per walker and initial state.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi
from .sampler import default_ladder, ladder_mode


class SystemsPT:
    """S independent parallel-tempering ensembles over ``models``.

    p0: list of [T, W, D_s] arrays (or one [S, T, W, D] array when every
    system has the same ndim).  betas: [T] (shared start) or [S, T].
    record_every > 0 keeps the coldest ``record_temps`` temperatures every
    ``record_every`` sweeps, up to ``max_records`` records.
    """

    WHAT = {"x": 7, "lnl": 8, "lnp": 9, "betas": 10, "accepted": 6, "tsw": 4, "smd": 5,
            "chain": 0, "rec_lnl": 1, "rec_lnp": 2, "nsweeps": 11}

    def __init__(self, models, ntemps, nwalkers, p0, betas=None, nsteps=1, precision="mixed", seed=None,
                 a=2.0, adapt_mode=0, adapt_tau=1000.0, adapt_nu=1.0, record_every=0, record_temps=1,
                 max_records=0):
        self.models = list(models)
        S = len(self.models)
        if S < 1:
            raise ValueError("need at least one system")
        self.S, self.ntemps, self.nwalkers = S, int(ntemps), int(nwalkers)
        self.ndims = np.array([m.ndim__ for m in self.models], dtype=np.int64)
        T, W = self.ntemps, self.nwalkers
        if isinstance(p0, np.ndarray) and p0.ndim == 4:
            p0 = list(p0)
        if len(p0) != S:
            raise ValueError(f"p0 must hold {S} systems")
        flat = []
        for s, (x, D) in enumerate(zip(p0, self.ndims)):
            x = np.asarray(x, dtype=np.float64)
            if x.shape != (T, W, D):
                raise ValueError(f"p0[{s}] must be [{T}, {W}, {D}], got {x.shape}")
            flat.append(x.ravel())
        p0c = np.ascontiguousarray(np.concatenate(flat))
        if betas is None:
            betas = default_ladder(T, int(np.median(self.ndims)))
        betas = np.asarray(betas, dtype=np.float64)
        if betas.ndim == 1:
            betas = np.broadcast_to(betas, (S, T))
        if betas.shape != (S, T):
            raise ValueError(f"betas must be [{T}] or [{S}, {T}]")
        betas = np.ascontiguousarray(betas)
        prec = self.models[0]._prec(precision)
        d = _abi.EmpSystemsDesc()
        d.n_systems, d.ntemps, d.nwalkers, d.nsteps = S, T, W, int(nsteps)
        d.precision = prec
        d.adapt_mode = ladder_mode(adapt_mode)
        d.a = float(a)
        self.seed = int(seed) if seed is not None else int.from_bytes(os.urandom(8), "little")
        d.seed = self.seed & 0xFFFFFFFFFFFFFFFF
        d.adapt_tau, d.adapt_nu = float(adapt_tau), float(adapt_nu)
        d.record_every, d.record_temps, d.max_records = int(record_every), int(record_temps), int(max_records)
        self.record_temps = int(record_temps) if record_every else 0
        self.max_records = int(max_records) if record_every else 0
        descs = (_abi.EmpModelDesc * S)()
        keep = []
        for s, m in enumerate(self.models):
            desc, k = m.descriptor()
            descs[s] = desc
            keep.append(k)
        h = C.c_void_p()
        _abi.check(_abi.lib().emp_systems_create(descs, C.byref(d), betas.ctypes.data, p0c.ctypes.data,
                                                 C.byref(h)))
        self._h = h
        del keep

    def close(self):
        if getattr(self, "_h", None):
            _abi.lib().emp_systems_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- running ---------------------------------------------------------------
    def run(self, nsweeps, use_graph=True):
        _abi.check(_abi.lib().emp_systems_run(self._h, int(nsweeps), int(bool(use_graph))))

    def autotune(self, trial_sweeps=2):
        """Times every launch configuration that fits (the chains are left
        untouched) and keeps the fastest; returns it as a dict."""
        t = _abi.EmpSystemsTuning()
        _abi.check(_abi.lib().emp_systems_autotune(self._h, int(trial_sweeps), C.byref(t)))
        return self._tdict(t)

    @staticmethod
    def _tdict(t):
        return {"waves_per_wg": t.waves_per_wg, "waves_per_theta": t.waves_per_theta,
                "wgs_per_cu": t.wgs_per_cu, "order": "lpt" if t.order == 0 else "input",
                "ms_per_sweep": t.ms_per_sweep}

    @property
    def tuning(self):
        t = _abi.EmpSystemsTuning()
        _abi.check(_abi.lib().emp_systems_get_tuning(self._h, C.byref(t)))
        return self._tdict(t)

    def set_tuning(self, waves_per_wg=16, waves_per_theta=1, wgs_per_cu=1, order="lpt"):
        t = _abi.EmpSystemsTuning(int(waves_per_wg), int(waves_per_theta), int(wgs_per_cu),
                                  0 if order == "lpt" else 1, -1.0)
        _abi.check(_abi.lib().emp_systems_set_tuning(self._h, C.byref(t)))

    def last_ms_per_sweep(self):
        v = C.c_double()
        _abi.check(_abi.lib().emp_systems_last_timing(self._h, C.byref(v)))
        return v.value

    # -- getters ------------------------------------------------------------------
    def _get(self, what, system, count, dtype=np.float64):
        shape = (C.c_int64 * 4)()
        buf = np.empty(max(int(count), 1), dtype=dtype)
        _abi.check(_abi.lib().emp_systems_get(self._h, what, int(system), buf.ctypes.data, buf.nbytes, shape))
        n = int(np.prod(list(shape)))
        return buf[:n].reshape([int(v) for v in shape])

    def _squeeze(self, a, ndims):
        return a.reshape(a.shape[:ndims])

    @property
    def nsweeps(self):
        return int(self._get(11, -1, 2, np.int64).ravel()[0])

    @property
    def n_lnl_evaluated(self):
        return int(self._get(11, -1, 2, np.int64).ravel()[1])

    def state(self, system):
        """(positions [T, W, D_s], lnL [T, W], ln prior [T, W]) of one system."""
        T, W, D = self.ntemps, self.nwalkers, int(self.ndims[system])
        x = self._squeeze(self._get(7, system, T * W * D), 3)
        ll = self._squeeze(self._get(8, system, T * W), 2)
        lp = self._squeeze(self._get(9, system, T * W), 2)
        return x, ll, lp

    def log_like(self):
        """Current lnL of every walker, [S, T, W]."""
        return self._squeeze(self._get(8, -1, self.S * self.ntemps * self.nwalkers), 3)

    def log_prior(self):
        return self._squeeze(self._get(9, -1, self.S * self.ntemps * self.nwalkers), 3)

    @property
    def betas(self):
        return self._squeeze(self._get(10, -1, self.S * self.ntemps), 2)

    @property
    def accepted(self):
        return self._squeeze(self._get(6, -1, self.S * self.ntemps * self.nwalkers, np.int64), 3)

    @property
    def acceptance_fraction(self):
        return self.accepted / max(self.nsweeps, 1)

    def get_tsw(self):
        """Mean swap acceptance per adjacent pair over the sweeps run, [S, T-1]."""
        return self._squeeze(self._get(4, -1, self.S * max(self.ntemps - 1, 1)), 2)

    def get_smd(self):
        return self._squeeze(self._get(5, -1, self.S * max(self.ntemps - 1, 1)), 2)

    def get_chain(self, system):
        """Recorded coldest temperatures of one system, [R, record_temps, W, D_s]."""
        D = int(self.ndims[system])
        return self._get(0, system, self.max_records * self.record_temps * self.nwalkers * D)

    def get_log_like(self, system):
        """Recorded lnL of one system, [R, record_temps, W]."""
        a = self._get(1, system, self.max_records * self.record_temps * self.nwalkers)
        return self._squeeze(a, 3)
