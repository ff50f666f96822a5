"""Parallel-tempering sampler with the reddemcee ``PTSampler`` interface.

The reference constructs ``reddemcee.PTSampler(nwalkers, ndim, my_likelihood,
my_prior, ntemps=, pool=, backend=, betas=, tsw_history=, smd_history=,
adapt_tau=, adapt_nu=, adapt_mode=)`` (emp/emp.py:576-593), runs
``sampler.run_mcmc(p1, nsweeps=, nsteps=, progress=)``
(support/endit_reddemcee.scr:3) and reads the results through
``get_chain/get_log_like/get_log_prob(discard, thin, flat)``, ``betas``,
``acceptance_fraction``, ``get_tsw``, ``get_smd`` and ``get_autocorr_time``
(emp/emp.py:1375-1429, 1969-1990).  This class keeps those names and
argument meanings; the sweep itself runs on the device (libemp_mi355x).

Moves (published methods; reddemcee/emcee are not vendored by the reference):
  * red/blue affine-invariant stretch move, a = 2 (Goodman & Weare 2010);
    the two halves are walkers [0, W/2) and [W/2, W);
  * adjacent-temperature swaps, hottest pair first, walker i of the hotter
    chain paired with a keyed pseudo-random permutation of the colder chain,
    accepted with min(1, exp((beta_c - beta_h)(lnL_h - lnL_c)));
  * ladder adaptation: adapt_mode 0, equal swap acceptance (Vousden, Farr &
    Mandel 2016); 1, equal mean |log swap ratio|; 2, specific heat (equal
    thermodynamic length, the mode tests/02_51peg_advanced.ipynb labels
    "Specific Heat"); -1, frozen.  reddemcee's own modes 1 and 2 are not in
    the reference, so these two are this package's definitions.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi
from .moves import move_descs, normalise_moves


def ladder_mode(adapt_mode):
    """adapt_mode as the engine takes it: None means 0; -1 (frozen), 0, 1, 2."""
    m = 0 if adapt_mode is None else int(adapt_mode)
    if m not in (-1, 0, 1, 2):
        raise NotImplementedError(f"adapt_mode {adapt_mode!r}: -1 (frozen), 0 (equal swap acceptance), "
                                  "1 (equal swap distance) or 2 (specific heat)")
    return m


def default_ladder(ntemps, ndim):
    """Geometric ladder ending at beta = 0 (the reference's adapted ladders end
    at or near 0, tests/00_mini_test.ipynb:243).  Step sqrt(2 / ndim)-ish in
    log temperature, as Vousden et al. (2016) suggest for ~25% swap rates."""
    if ntemps == 1:
        return np.array([1.0])
    step = 1.0 + np.sqrt(2.0 / max(ndim, 1))
    betas = step ** -np.arange(ntemps - 1, dtype=np.float64)
    return np.concatenate([betas, [0.0]])


def integrated_time(x, c=5, tol=50, quiet=True):
    """Integrated autocorrelation time of x[niter, nwalkers, ndim]
    (Sokal 1989 automatic window, as emcee.autocorr.integrated_time)."""
    x = np.atleast_1d(x)
    if x.ndim == 1:
        x = x[:, None, None]
    if x.ndim == 2:
        x = x[:, :, None]
    n_t, n_w, n_d = x.shape
    tau_est = np.empty(n_d)
    windows = np.empty(n_d, dtype=int)
    for d in range(n_d):
        f = np.zeros(n_t)
        for k in range(n_w):
            y = x[:, k, d] - np.mean(x[:, k, d])
            n = 1 << int(np.ceil(np.log2(2 * n_t)))
            fy = np.fft.fft(y, n=n)
            acf = np.fft.ifft(fy * np.conjugate(fy))[:n_t].real
            f += acf / acf[0] if acf[0] != 0 else acf
        f /= n_w
        taus = 2.0 * np.cumsum(f) - 1.0
        m = np.arange(len(taus)) < c * taus
        windows[d] = np.argmin(m) if np.any(~m) else len(taus) - 1
        tau_est[d] = taus[windows[d]]
    flag = tol * tau_est > n_t
    if np.any(flag) and not quiet:
        raise RuntimeError(f"chain shorter than {tol} times the autocorrelation time: {tau_est}")
    return tau_est


class PTState:
    def __init__(self, coords, log_like, log_prior, betas):
        self.coords, self.log_like, self.log_prior, self.betas = coords, log_like, log_prior, betas


class PTSampler:
    """reddemcee-compatible PT sampler over a :class:`ReddModel`.

    ``log_likelihood_fn`` / ``log_prior_fn`` must be the ``my_likelihood`` /
    ``my_prior`` callables of one ReddModel (bound methods), or the model
    itself may be passed as ``log_likelihood_fn``.  ``pool`` is accepted and
    ignored (the device replaces the process pool, support/pools/01.pool).
    ``moves`` takes emcee's forms (see :mod:`astroemperor_amd.moves`); the
    default is the stretch move with scale ``a``.  ``backend`` must be None
    (EMPEROR's default, emp/emp.py:570-573): the history stays on the device,
    is read back through ``sampler.backend`` as reddemcee's is, and
    ``save()`` writes it to disk; an HDF5 backend object raises
    NotImplementedError.
    """

    def __init__(self, nwalkers, ndim, log_likelihood_fn, log_prior_fn=None, ntemps=None, pool=None,
                 backend=None, betas=None, tsw_history=True, smd_history=True, adapt_tau=1000,
                 adapt_nu=1, adapt_mode=0, precision="fp64", seed=None, a=2.0, record=True, moves=None):
        from .model import ReddModel
        model = log_likelihood_fn if isinstance(log_likelihood_fn, ReddModel) else \
            getattr(log_likelihood_fn, "__self__", None)
        if not isinstance(model, ReddModel):
            raise TypeError("log_likelihood_fn must be a ReddModel or its my_likelihood method: "
                            "the device engine evaluates the model itself")
        if log_prior_fn is not None and getattr(log_prior_fn, "__self__", model) is not model:
            raise TypeError("log_prior_fn must belong to the same ReddModel")
        if ndim != model.ndim__:
            raise ValueError(f"ndim={ndim} but the model has {model.ndim__} free parameters")
        adapt_mode = ladder_mode(adapt_mode)
        if backend is not None:
            raise NotImplementedError("backend must be None: the history is kept on the device "
                                      "(read it through sampler.backend, write it with save())")
        self.model = model
        self.nwalkers = int(nwalkers)
        self.ndim = int(ndim)
        if betas is not None:
            betas = np.asarray(betas, dtype=np.float64)
            ntemps = len(betas) if ntemps is None else ntemps
            if len(betas) != ntemps:
                raise ValueError("betas must have ntemps entries")
        self.ntemps = int(ntemps or 1)
        self._betas0 = betas if betas is not None else default_ladder(self.ntemps, self.ndim)
        self.tsw_history_bool = bool(tsw_history)
        self.smd_history_bool = bool(smd_history)
        self.adapt_tau, self.adapt_nu = float(adapt_tau), float(adapt_nu)
        self.adapt_mode = adapt_mode
        self.precision = precision
        self.a = float(a)
        self.moves = normalise_moves(moves, self.a)
        self.seed = int(seed) if seed is not None else int.from_bytes(os.urandom(8), "little")
        self.record = bool(record)
        self.D_ = None
        self._pt = None
        self._capacity = 0
        self._nsteps = None
        self._iterations = 0
        self._chain_parts = []

    # -- engine lifetime -------------------------------------------------
    def _create(self, p0, nsteps, capacity):
        L = _abi.lib()
        d = _abi.EmpPTDesc()
        d.ntemps, d.nwalkers, d.nsteps = self.ntemps, self.nwalkers, int(nsteps)
        d.precision = self.model._prec(self.precision)
        d.a = self.a
        d.seed = self.seed & 0xFFFFFFFFFFFFFFFF
        d.adapt_mode = self.adapt_mode
        d.record = int(self.record)
        d.adapt_tau, d.adapt_nu = self.adapt_tau, self.adapt_nu
        d.max_sweeps = int(capacity)
        p0 = np.ascontiguousarray(p0, dtype=np.float64)
        betas = np.ascontiguousarray(self._betas0, dtype=np.float64)
        h = C.c_void_p()
        _abi.check(L.emp_pt_create(self.model.device.handle, C.byref(d), betas.ctypes.data,
                                   p0.ctypes.data, C.byref(h)))
        self._pt = h
        self._capacity = capacity
        self._nsteps = int(nsteps)
        self._apply_moves()

    def _apply_moves(self):
        descs = move_descs(self.moves, self.ndim)
        _abi.check(_abi.lib().emp_pt_set_moves(self._pt, C.addressof(descs), len(descs)))
        for m, w in self.moves:
            if getattr(m, "steps", None) is not None and w > 0:
                if len(m.steps) != self.ntemps:
                    raise ValueError(f"GradientMove has {len(m.steps)} steps for {self.ntemps} temperatures")
                _abi.check(_abi.lib().emp_pt_set_gradient_steps(self._pt, m.steps.ctypes.data))

    def close(self):
        if self._pt:
            _abi.lib().emp_pt_destroy(self._pt)
            self._pt = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_mcmc(self, initial_state, nsweeps, nsteps=1, progress=False, use_graph=True, capacity=None):
        """Run `nsweeps` sweeps from `initial_state` ([T, W, D] array or a
        PTState); continuing runs pass the previous state (or None)."""
        if self._pt is not None and int(nsteps) != self._nsteps:
            raise ValueError(f"nsteps is fixed at {self._nsteps} for this sampler")
        if initial_state is not None:
            p0 = initial_state.coords if isinstance(initial_state, PTState) else initial_state
            p0 = np.asarray(p0, dtype=np.float64)
            if p0.shape != (self.ntemps, self.nwalkers, self.ndim):
                raise ValueError(f"initial state must be [{self.ntemps}, {self.nwalkers}, {self.ndim}]")
        if self._pt is None:
            if initial_state is None:
                raise ValueError("the first run needs an initial state")
            self._create(p0, nsteps, capacity or int(nsweeps))
        elif initial_state is not None and not isinstance(initial_state, PTState):
            _abi.check(_abi.lib().emp_pt_set_state(self._pt, p0.ctypes.data, None))
        _abi.check(_abi.lib().emp_pt_run(self._pt, int(nsweeps), int(bool(use_graph))))
        self._iterations += int(nsweeps)
        return self.get_last_sample()

    def set_moves(self, moves):
        """Replace the move mix (emcee's forms); takes effect from the next sweep."""
        self.moves = normalise_moves(moves, self.a)
        if self._pt is not None:
            self._apply_moves()

    @property
    def gradient_steps(self):
        """The gradient move's step at each temperature [T] (once the move
        has been enabled)."""
        return self._get(12).ravel()

    def adapt_gradient_steps(self, nsweeps=16, target=0.5, batch=1, step=None):
        """Burn-in for the gradient move: runs `nsweeps` gradient-only sweeps
        in batches of `batch` and after each batch moves every temperature's
        step towards the acceptance `target` (in log step: x10 or /10 until
        the target is bracketed, then bisection), then restores the move mix
        with the adapted steps held fixed.  The sweeps are ordinary sweeps of
        the chain.  Returns the steps [T]."""
        if self._pt is None:
            raise ValueError("run the sampler first (it needs an initial state)")
        if not (0.0 < target < 1.0) or batch < 1:
            raise ValueError("target must lie in (0, 1) and batch be >= 1")
        from .moves import GradientMove
        mix = self.moves
        grad = [m for m, w in mix if isinstance(m, GradientMove)]
        g0 = grad[0] if grad else GradientMove(step)
        if g0.steps is not None:
            logs = np.log(g0.steps.copy())
        else:
            logs = np.full(self.ntemps, np.log(g0.scale(self.ndim) if step is None else float(step)))
        lo = np.full(self.ntemps, -np.inf)
        hi = np.full(self.ntemps, np.inf)
        done = 0
        while done < nsweeps:
            n = min(batch, nsweeps - done)
            self.set_moves([(GradientMove(np.exp(logs)), 1.0)])
            a0 = self.accepted.sum(1)
            self.run_mcmc(None, nsweeps=n, nsteps=self._nsteps)
            acc = (self.accepted.sum(1) - a0) / float(n * self._nsteps * self.nwalkers)
            for t in range(self.ntemps):
                if acc[t] > target:
                    lo[t] = logs[t]
                else:
                    hi[t] = logs[t]
                if np.isfinite(lo[t]) and np.isfinite(hi[t]):
                    logs[t] = 0.5 * (lo[t] + hi[t])
                else:
                    logs[t] += np.log(10.0) if acc[t] > target else -np.log(10.0)
            done += n
        steps = np.exp(logs)
        new_mix = [(GradientMove(steps), w) if isinstance(m, GradientMove) else (m, w) for m, w in mix]
        self.set_moves(new_mix)
        return steps

    def select_adjustment(self, code):
        """reddemcee's select_adjustment('00') freezes the ladder
        (support/endit_freeze1.scr:11)."""
        if code != "00":
            raise NotImplementedError("only '00' (frozen ladder)")
        self.adapt_mode = -1
        if self._pt is not None:
            _abi.check(_abi.lib().emp_pt_set_adaptation(self._pt, -1, self.adapt_tau, self.adapt_nu))

    # -- getters (reddemcee names) -----------------------------------------
    def _get(self, what, dtype=np.float64):
        L = _abi.lib()
        shape = (C.c_int64 * 4)()
        T, W, D = self.ntemps, self.nwalkers, self.ndim
        S = max(self._capacity, 1)
        R = S * max(self._nsteps or 1, 1)          # history rows: one per step
        cap = {0: R * T * W * D, 1: R * T * W, 2: R * T * W, 3: S * T, 4: S * T, 5: S * T,
               6: T * W, 7: T * W * D, 8: T * W, 9: T * W, 10: T, 11: 2, 12: T}[what]
        buf = np.empty(cap, dtype=dtype)
        _abi.check(L.emp_pt_get(self._pt, what, buf.ctypes.data, buf.nbytes, shape))
        n = int(np.prod(list(shape)))
        return buf[:n].reshape([int(s) for s in shape])

    @property
    def iteration(self):
        """Stretch steps taken per walker: sweeps x nsteps (one history row each)."""
        return int(self._get(11, np.int64).ravel()[0]) * self._nsteps if self._pt else 0

    @property
    def n_lnl_evaluated(self):
        """Likelihoods actually evaluated (proposals inside the prior support)."""
        return int(self._get(11, np.int64).ravel()[1]) if self._pt else 0

    def _sel(self, arr, discard, thin, flat):
        arr = arr[int(discard)::int(thin)]               # [S', T, ...]
        arr = np.moveaxis(arr, 0, 1)                      # [T, S', ...]
        if flat:
            s = arr.shape
            arr = arr.reshape(s[0], s[1] * s[2], *s[3:])
        return arr

    def get_chain(self, discard=0, thin=1, flat=False):
        return self._sel(self._get(0), discard, thin, flat)

    def get_log_like(self, discard=0, thin=1, flat=False):
        return self._sel(self._get(1)[..., 0], discard, thin, flat)

    def get_log_prior(self, discard=0, thin=1, flat=False):
        return self._sel(self._get(2)[..., 0], discard, thin, flat)

    def get_log_prob(self, discard=0, thin=1, flat=False):
        """Tempered log posterior beta_t lnL + ln prior, each row (step) with the
        ladder of its sweep."""
        ll = self._get(1)[..., 0]
        lp = self._get(2)[..., 0]
        b = self.get_betas()[:ll.shape[0], :, None]
        with np.errstate(invalid="ignore"):
            tempered = np.where(b == 0, 0.0, b * ll)
        return self._sel(tempered + lp, discard, thin, flat)

    def get_betas(self, discard=0, thin=1):
        """The ladder of every history row [rows, T]: each sweep's ladder
        (after its update) repeated for its nsteps rows, so that ``discard``
        counts steps as EMPEROR passes it (emp/emp.py:962, while get_tsw and
        get_smd take discard // nsteps, 1986-1989)."""
        per_sweep = self._get(3)[..., 0, 0]
        return np.repeat(per_sweep, max(self._nsteps or 1, 1), axis=0)[int(discard)::int(thin)]

    @property
    def betas(self):
        return self._get(10).ravel()

    def get_tsw(self, discard=0):
        return self._get(4)[..., 0, 0][int(discard):]

    def get_smd(self, discard=0):
        return self._get(5)[..., 0, 0][int(discard):]

    @property
    def accepted(self):
        """Accepted stretch proposals per temperature and walker [T, W]."""
        return self._get(6, np.int64)[..., 0, 0]

    @property
    def acceptance_fraction(self):
        return self.accepted.astype(np.float64) / max(self.iteration, 1)

    @property
    def backend(self):
        """The history as reddemcee's backend is read (emp/emp.py:734-761):
        ``backend.iteration``, ``backend.tsw_history`` / ``smd_history`` and
        per temperature ``backend[t].get_chain()`` and the other getters."""
        from .backend import RunView
        return RunView(self)

    def get_last_sample(self):
        return PTState(self._get(7)[..., 0], self._get(8)[..., 0, 0], self._get(9)[..., 0, 0], self.betas)

    def get_autocorr_time(self, discard=0, thin=1, quiet=True, tol=50, c=5):
        chain = self.get_chain(discard=discard, thin=thin)      # [T, S, W, D]
        return np.array([integrated_time(chain[t], c=c, tol=tol, quiet=quiet) * thin
                         for t in range(self.ntemps)])

    # -- chain storage (reddemcee's PTHDFBackend, emp/emp.py:722-762) ------------
    def save(self, path):
        """Writes the recorded history to a .npz archive (see
        :mod:`astroemperor_amd.backend`); read it back with SavedRun."""
        from .backend import save_run
        save_run(self, path)

    # -- evidence (reddemcee's get_evidence_*; emp/emp.py:1432-1447) -------------
    def lnl_moments(self, discard=0):
        """Per temperature (ladder order): mean lnL, stepping-stone log ratio
        to the next colder temperature, and its variance, over the recorded
        sweeps from `discard` on (device reduction)."""
        T = self.ntemps
        mean, lse, var = np.empty(T), np.empty(T), np.empty(T)
        _abi.check(_abi.lib().emp_pt_lnl_moments(self._pt, int(discard), mean.ctypes.data, lse.ctypes.data,
                                                  var.ctypes.data))
        return mean, lse, var

    def get_evidence_ti(self, discard=0, pchip=True):
        """(log Z, error) by thermodynamic integration over the current ladder."""
        from .evidence import ti_log_evidence
        mean, _, _ = self.lnl_moments(discard)
        return ti_log_evidence(self.betas, mean, pchip=pchip)

    def get_evidence_ss(self, discard=0):
        """(log Z, error) by stepping stones over the current ladder."""
        from .evidence import ss_log_evidence
        mean, lse, var = self.lnl_moments(discard)
        return ss_log_evidence(self.betas, mean, lse, var)

    def get_evidence_hybrid(self, discard=0):
        """reddemcee's hybrid estimator is not available here; EMPEROR then
        falls back to get_evidence_ti (emp/emp.py:1443-1446)."""
        raise NotImplementedError("the hybrid evidence estimator is not available; use 'ti' or 'ss'")

    # -- checkpoint / resume ---------------------------------------------------
    def checkpoint(self, path):
        _abi.check(_abi.lib().emp_pt_checkpoint(self._pt, str(path).encode()))

    def restore(self, path):
        _abi.check(_abi.lib().emp_pt_restore(self._pt, str(path).encode()))

    # -- timing ------------------------------------------------------------------
    def last_timing(self):
        a, b = C.c_double(), C.c_double()
        _abi.check(_abi.lib().emp_pt_last_timing(self._pt, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_swap_path(self, path="auto"):
        """Swap-round implementation ('auto', 'lds', 'stream' or 'global'; the
        results are identical): for comparisons and tests."""
        code = {"auto": _abi.EMP_SWAP_AUTO, "lds": _abi.EMP_SWAP_LDS, "stream": _abi.EMP_SWAP_STREAM,
                "global": _abi.EMP_SWAP_GLOBAL}[path]
        _abi.check(_abi.lib().emp_pt_set_swap_path(self._pt, code))

    def autotune(self, trials=10):
        """Pick the fastest waves per walker from timed graph-replayed sweeps
        (state left unchanged); returns (waves_per_walker, ms per sweep)."""
        w, t = C.c_int32(), C.c_double()
        _abi.check(_abi.lib().emp_pt_autotune(self._pt, int(trials), C.byref(w), C.byref(t)))
        return w.value, t.value

    def time_halfstep(self, n=20):
        v = C.c_double()
        _abi.check(_abi.lib().emp_pt_time_halfstep(self._pt, int(n), C.byref(v)))
        return v.value

    def time_gradstep(self, n=4):
        """Mean duration (ms) of n adjoint launches over the proposals of
        gradient sweeps, each timed with events on the engine stream; runs
        ceil(n / (2 nsteps)) ordinary gradient sweeps."""
        v = C.c_double()
        _abi.check(_abi.lib().emp_pt_time_gradstep(self._pt, int(n), C.byref(v)))
        return v.value
