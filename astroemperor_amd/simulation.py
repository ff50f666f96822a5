"""``Simulation``: the reference's user-facing run API on the device engine.

Mirrors the parts of ``astroemperor.Simulation`` (emp/emp.py:2240-2916) that
drive the RV hot path: ``load_data``, ``set_engine('reddemcee')`` with its
``engine_config``/``run_config`` dictionaries, ``add_condition``, the model
switches (``keplerian_parameterisation``, ``use_c``, ``moav``,
``acceleration``, ``sinusoid``, ``magnetic_cycle``, ``switch_SA`` (set before
``load_data``), ``switch_jitter``,
``eccentricity_*``, ``jitter_*``),
``run()`` and ``autorun(k_start, k_end)`` with the BIC stopping rule.  Instead
of writing and executing a script in a subprocess (emp/emp.py:137-175,
2559-2582), the model is built in-process and sampled on the GPU.

Extra ``engine_config`` keys: ``'backend'`` ('hip_fp64' or 'hip_mixed'),
``'seed'`` and ``'init'``: 'prior' (default, set_init/test_init as the
reference) or 'map' (walkers drawn around the posterior mode that a batched
L-BFGS search finds, with the Laplace covariance; ``autograd.map_init``).  ``'moves'``, which the reference declares but never passes to
reddemcee (emp/emp.py:2373, 576-593), reaches the sampler here (e.g.
``[(StretchMove(), 0.8), (GradientMove(), 0.2)]``).  ``semantics`` selects 'reference' (default) or 'intended' MA.
"""
from __future__ import annotations

import itertools
import os
import time

import numpy as np

from .blocks import (AccelerationBlock, JitterBlock, KeplerianBlock, MagneticCycleBlock, MOAVBlock,
                     OffsetBlock, SAIBlock, SinusoidBlock, SmartSetter)
from .data import DataWrapper, RVData
from .model import ReddModel
from .postprocess import StatsMixin
from .sampler import PTSampler


def set_init(model, ntemps, nwalkers, rng):
    """Initial positions (emp/emp.py:617-647): per free parameter a stratified
    uniform spread over init_pos (or the limits), shrunk by 0.707 for 'is_hou'
    parameters, shuffled per temperature."""
    free = [p for p in model.params() if p.fixed is None]
    pos = np.zeros((ntemps, nwalkers, len(free)))
    for t in range(ntemps):
        for j, p in enumerate(free):
            r = 0.707 if getattr(p, "is_hou", False) else 1.0
            ip = getattr(p, "init_pos", [None, None]) or [None, None]
            b = p.limits[0] if ip[0] is None else np.round(ip[0], 8)
            a = p.limits[1] if ip[1] is None else np.round(ip[1], 8)
            m = (a + b) / 2
            rr = (a - b) / 2 * r
            dist = np.sort(rng.uniform(0, 1, nwalkers))
            pos[t][:, j] = rr * (2 * dist - 1) + m
            rng.shuffle(pos[t, :, j])
    return pos


def test_init(model, ntemps, nwalkers, rng, max_repeats=100):
    """emp/emp.py:650-679: redraw walkers whose prior is -inf, up to 100 rounds."""
    p0 = set_init(model, ntemps, nwalkers, rng)
    for _ in range(max_repeats):
        lp = model.lnprior_batch(p0.reshape(-1, p0.shape[-1])).reshape(ntemps, nwalkers)
        bad = ~np.isfinite(lp)
        if not bad.any():
            return p0
        fresh = set_init(model, ntemps, nwalkers, rng)
        p0[bad] = fresh[bad]
    print("COULDNT FIND VALID INITIAL POSITION")
    return p0


class Simulation(StatsMixin):
    def __init__(self):
        self.time_init = time.time()
        self.read_loc = ""
        self.starname = None
        self.keplerian_parameterisation = 0
        self.use_c = False
        self.moav = {"order": 0, "global": False}
        self.acceleration = 0
        self.sinusoid = 0
        self.magnetic_cycle = 0
        self.switch_SA = False
        self.cornums = []
        self.switch_jitter = True
        self.eccentricity_limits = [0, 1]
        self.eccentricity_prargs = [0, 0.1]
        self.jitter_limits = [0, 1]
        self.jitter_prargs = [5, 5]
        self.cores__ = os.cpu_count()
        self.instrument_names_RV = None
        self.starmass = 1.0
        self.semantics = "reference"
        self.conds = []
        self.blocks__ = []
        self.kplanets__ = 0
        self.nins__ = 0
        self.model = None
        self.sampler = None
        self.switch_first_run = True
        self.results = []
        self._init_stats()

    # -- data (emp/emp.py:2263-2312) ----------------------------------------
    def load_data(self, folder_name):
        self.starname = folder_name
        self.data_wrapper = DataWrapper(folder_name, read_loc=self.read_loc)
        self._set_data(self.data_wrapper.get_data__())

    def load_arrays(self, bjd_list, rv_list, erv_list, names=None, starname="arrays", sa_list=None):
        """Extension: per-instrument arrays instead of a datafiles folder
        (sa_list: raw activity-index columns per instrument, or None)."""
        self.starname = starname
        self._set_data(RVData.from_arrays(bjd_list, rv_list, erv_list, labels=names, sa_list=sa_list))

    def _set_data(self, data):
        # emp.py:2305-2312: activity columns are kept only with switch_SA
        self.cornums = list(data.nsai)
        if not self.switch_SA:
            data = data.without_activity()
            self.cornums = [0] * data.nins
        self.my_data = data
        self.nins__ = data.nins
        if self.instrument_names_RV is None:
            self.instrument_names_RV = list(data.labels)
        self.SmartSetter = SmartSetter(data.as_table())

    # -- engine (emp/emp.py:2315-2388) ------------------------------------
    def set_engine(self, eng):
        if eng != "reddemcee":
            raise NotImplementedError("the device engine implements reddemcee's parallel tempering")
        self.engine_config = {"setup": np.array([5, 100, 500, 2]), "ntemps": 5, "betas": None,
                              "moves": None, "tsw_history": True, "smd_history": True,
                              "adapt_tau": 1000, "adapt_nu": 1, "adapt_mode": 0, "progress": True,
                              "backend": "hip_fp64", "seed": None, "init": "prior"}
        self.run_config = {"burnin": None, "adaptation_batches": None, "adaptation_nsweeps": None,
                           "thin": 1, "discard": 0.5, "logger_level": "CRITICAL"}

    # -- blocks (emp/emp.py:1212-1320) --------------------------------------
    def add_keplerian(self):
        self.kplanets__ += 1
        kb = KeplerianBlock(self.keplerian_parameterisation, self.kplanets__, self.use_c)
        self.SmartSetter(kb, prargs=[self.eccentricity_limits, self.eccentricity_prargs])
        self.blocks__.insert(self.kplanets__ - 1, kb)

    def _append(self, block, *args):
        self.SmartSetter(block, *args)
        self.blocks__.append(block)

    def add_offset(self):
        self._append(OffsetBlock(self.nins__))

    def add_jitter(self):
        self._append(JitterBlock(self.nins__), self.jitter_limits, self.jitter_prargs)

    def add_moav(self):
        self._append(MOAVBlock(self.nins__, self.moav))

    def add_acceleration(self):
        self._append(AccelerationBlock(self.acceleration))

    def add_sai(self):
        self._append(SAIBlock(self.nins__, self.cornums))

    def add_sinusoid(self):
        self._append(SinusoidBlock(self.sinusoid))

    def add_magnetic_cycle(self):
        self._append(MagneticCycleBlock(self.magnetic_cycle))

    def _autorun_add_RV_ins(self):                               # emp.py:2628-2650
        if self.acceleration:
            self.add_acceleration()
        self.add_offset()
        if self.switch_jitter:
            self.add_jitter()
        if self.moav["order"]:
            self.add_moav()
        if self.switch_SA and sum(self.cornums) > 0:
            self.add_sai()
        if self.sinusoid:
            self.add_sinusoid()
        if self.magnetic_cycle:
            self.add_magnetic_cycle()

    # -- conditions (emp/emp.py:2790-2810) -----------------------------------
    def add_condition(self, cond):
        self.conds.append(cond)

    def apply_conditions(self):
        applied = []
        for b in self.blocks__:
            for p in b:
                for c in self.conds:
                    if p.name == c[0]:
                        setattr(p, c[1], c[2])
                        if c not in applied:
                            applied.append(c)
        for c in applied:
            while c in self.conds:
                self.conds.remove(c)

    def update_model(self):
        self.model = ReddModel(self.blocks__, self.my_data, semantics=self.semantics)
        return self.model

    # -- run (emp/emp.py:2463-2527) -------------------------------------------
    def run(self):
        t0 = time.time()
        self.apply_conditions()
        self.update_model()
        ntemps, nwalkers, nsweeps, nsteps = [int(v) for v in self.engine_config["setup"]]
        niter = nsteps * nsweeps
        burnin = self.run_config["burnin"]
        self.reddemcee_discard = (int(burnin * niter) if isinstance(burnin, float)
                                  else int(burnin) if isinstance(burnin, int) else 0)
        self.reddemcee_thin = self.run_config["thin"]
        seed = self.engine_config.get("seed")
        rng = np.random.default_rng(seed)
        if self.engine_config.get("init", "prior") == "map":
            from .autograd import map_init
            p1, self.map_mode, self.map_lnpost = map_init(self.model, ntemps, nwalkers, rng,
                                                          betas=self.engine_config["betas"])
        else:
            p1 = test_init(self.model, ntemps, nwalkers, rng)
        prec = "mixed" if self.engine_config.get("backend") == "hip_mixed" else "fp64"
        self.sampler = PTSampler(nwalkers, self.model.ndim__, self.model.my_likelihood, self.model.my_prior,
                                 ntemps=ntemps, betas=self.engine_config["betas"],
                                 tsw_history=self.engine_config["tsw_history"],
                                 smd_history=self.engine_config["smd_history"],
                                 adapt_tau=self.engine_config["adapt_tau"],
                                 adapt_nu=self.engine_config["adapt_nu"],
                                 adapt_mode=self.engine_config["adapt_mode"], precision=prec,
                                 seed=None if seed is None else int(rng.integers(0, 2 ** 63)),
                                 moves=self.engine_config.get("moves"))
        self.sampler.D_ = np.diff(self.model.limits_free()).ravel()      # emp.py:595-602
        adapt_n = self.run_config.get("adaptation_nsweeps")
        if self.run_config.get("adaptation_batches") and adapt_n:       # endit_freeze1.scr:1-16
            self.sampler.run_mcmc(p1, nsweeps=self.reddemcee_discard, nsteps=nsteps, capacity=nsweeps)
            self.sampler.select_adjustment("00")
            self.sampler.run_mcmc(None, nsweeps=nsweeps - self.reddemcee_discard, nsteps=nsteps)
        else:
            self.sampler.run_mcmc(p1, nsweeps=nsweeps, nsteps=nsteps)
        self._postprocess()
        self.time_run = time.time() - t0

    def _autorun_add_blocks(self):                                # emp.py:2614-2625
        if self.switch_first_run:
            self._autorun_add_RV_ins()
            for _ in range(self.k_start):
                self.add_keplerian()

    def _load_next_model(self):                                   # emp.py:2773-2778
        self.add_keplerian()
        self.update_model()

    def autorun(self, k_start=0, k_end=10):
        """emp.py:2414-2456: run with k_start Keplerians, then add one at a time
        while the comparison criterion (BIC by default) improves by more than
        the tolerance (5), starting each new run from the previous run's
        planet ranges and constraints (method 'range' by default)."""
        self.k_start, self.k_end = k_start, k_end
        assert self.k_start <= self.k_end, f"Invalid keplerian starting point: ({k_start}, {k_end})"
        while self.k_start <= self.k_end:
            self._autorun_add_blocks()
            self._stat_holder_update()
            self.run()
            self.switch_first_run = False
            self.k_start += 1
            if self._check_continue_run():
                break
            self._set_init_pos()
            self._constrain_next_run()
            self._load_next_model()
        return self.results
