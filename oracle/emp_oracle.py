"""CPU oracle for EMPEROR's batched Keplerian radial-velocity likelihood.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The product path in
``astroemperor_amd`` never imports it and fails loudly when its HIP library is
missing.

This is a plain FP64 numpy restatement of the Python code that the reference
generates at run time (``temp_script_*.py``), assembled from its text
templates.  Every function cites the template or source line it follows
(paths relative to the reference repository root, ``src/astroemperor/``
abbreviated to ``emp/``).

Pinning (see the tests that exercise each function):

* ``load_rv_files``            -- reference fixture ``51peg.vels`` plus the
                                  recorded KAT lnL values (tests/test_oracle_kat.py).
* ``kepler_solve``             -- 50-digit mpmath roots (tests/golden/kepler_mpmath.npz)
                                  and the identities E(0)=0, E(pi)=pi, E(-M)=-E(M).
* ``keplerian_term``           -- mpmath evaluation of the literal template
                                  formula (tests/golden/lnl_mpmath.json) and
                                  closed forms at e below the omega threshold.
* ``OracleModel.my_likelihood`` -- same mpmath vectors, the 51 Peg KATs.
* ``OracleModel.my_prior``     -- scipy.stats closed forms.
* ``ma_residuals``             -- a literal scalar loop of moav01 on tiny N.
* Sinusoid / MagneticCycle terms of ``my_model`` -- a sinusoid equals the
                                  circular (e = 0, omega = 0) Keplerian; a
                                  magnetic cycle equals two sinusoids at P and
                                  2P; spot values; unfiltered by the MA
                                  (tests/test_oracle_cosine_terms.py).
* ``stretch_step`` / ``swap_step`` -- injected-random unit tests and the
                                  tempered-Gaussian moment tests.
* ``ladder_update`` / ``lnl_spread`` -- fixed points: equal acceptances or
                                  distances leave the ladder alone; under
                                  sig(beta) = s / beta (a Gaussian likelihood)
                                  a geometric ladder is the exact fixed point of
                                  mode 2, and a perturbed one moves back toward
                                  it; the spread against numpy's std(ddof=1)
                                  (tests/test_oracle_golden.py).  Modes 1 and 2
                                  follow this package's own definitions:
                                  parity with reddemcee's is unpinned.
* ``langevin_factor`` / ``langevin_logq`` / ``langevin_halfstep`` -- the
                                  proposal density against scipy's multivariate
                                  normal, detailed balance on random pairs, and
                                  the stationary moments of correlated Gaussian
                                  targets (tests/test_oracle_langevin.py).
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_KEPLER_LIB = None


# ---------------------------------------------------------------------------
# Data staging -- emp/qol_utils.py:61-100 (DataWrapper.mk_RV) and
# emp/qol_utils.py:192-198 (get_data__).
# ---------------------------------------------------------------------------
def load_rv_files(paths):
    """Read ``.vels`` files the way ``DataWrapper`` does.

    * each file: ``np.loadtxt`` -> columns BJD, RV, eRV (qol_utils.py:63);
    * RV mean subtracted when ``|mean| > 1e-6`` (qol_utils.py:84-85);
    * ``Flag`` = 1-based file index in the given (sorted-filename) order
      (qol_utils.py:95);
    * concatenated, sorted by BJD, ``common_t = min(BJD)`` subtracted
      (qol_utils.py:193-197).

    Returns dict(X, Y, YERR, FLAG, common_t, ndata).
    """
    frames = []
    for k, p in enumerate(paths):
        d = np.loadtxt(p)
        if d.ndim == 1:
            d = d[None, :]
        bjd, rv, erv = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
        mean = rv.mean()
        if abs(mean) > 1e-6:
            rv = rv - mean
        frames.append((bjd, rv, erv, np.full(len(bjd), k + 1, dtype=np.int64)))
    bjd = np.concatenate([f[0] for f in frames])
    rv = np.concatenate([f[1] for f in frames])
    erv = np.concatenate([f[2] for f in frames])
    flag = np.concatenate([f[3] for f in frames])
    order = np.argsort(bjd, kind="quicksort")
    bjd, rv, erv, flag = bjd[order], rv[order], erv[order], flag[order]
    common_t = bjd.min()
    return dict(X=bjd - common_t, Y=rv, YERR=erv, FLAG=flag, common_t=common_t,
                ndata=[int(f[0].size) for f in frames])


# ---------------------------------------------------------------------------
# Kepler solve -- stand-in for kepler.solve (kepler.py 0.0.7,
# requirements.txt:143; call sites support/models/kep0*.model).
# ---------------------------------------------------------------------------
def _load_kepler_lib():
    global _KEPLER_LIB
    if _KEPLER_LIB is None:
        path = os.path.join(_HERE, "libkepler_oracle.so")
        if os.path.exists(path):
            lib = ctypes.CDLL(path)
            lib.kepler_oracle_solve.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_int64]
            lib.kepler_oracle_solve.restype = None
            _KEPLER_LIB = lib
        else:
            _KEPLER_LIB = False
    return _KEPLER_LIB


def _kepler_solve_numpy(M, e):
    """Vectorised bracketed Newton (same method as kepler_oracle.c)."""
    M = np.asarray(M, dtype=np.float64)
    e = np.broadcast_to(np.asarray(e, dtype=np.float64), M.shape)
    two_pi = 2.0 * np.pi
    m = np.fmod(M, two_pi)
    m = np.where(m < 0, m + two_pi, m)
    upper = m > np.pi
    mh = np.where(upper, two_pi - m, m)
    lo = mh.copy()
    hi = np.minimum(mh + e, np.pi)
    E = mh + 0.85 * e
    E = np.where(E > hi, 0.5 * (lo + hi), E)
    for _ in range(100):
        s, c = np.sin(E), np.cos(E)
        f = E - e * s - mh
        hi = np.where(f > 0, E, hi)
        lo = np.where(f > 0, lo, E)
        En = E - f / (1.0 - e * c)
        bad = ~((En > lo) & (En < hi))
        En = np.where(bad, 0.5 * (lo + hi), En)
        d = np.abs(En - E)
        E = En
        if np.all(d <= 4e-16 * np.maximum(E, 1.0)):
            break
    E = E - (E - e * np.sin(E) - mh) / (1.0 - e * np.cos(E))
    E = np.where((e == 0) | (mh == 0) | (mh == np.pi), mh, E)
    return np.where(upper, two_pi - E, E)


def kepler_solve(M, e):
    """E with E - e sin E = M, elementwise; E in [0, 2 pi)."""
    M = np.ascontiguousarray(M, dtype=np.float64)
    e = np.ascontiguousarray(np.broadcast_to(np.asarray(e, dtype=np.float64), M.shape))
    lib = _load_kepler_lib()
    if lib:
        out = np.empty_like(M)
        lib.kepler_oracle_solve(M.ctypes.data, e.ctypes.data, out.ctypes.data, M.size)
        return out
    return _kepler_solve_numpy(M, e)


# ---------------------------------------------------------------------------
# Keplerian term -- support/models/kep00.model ... kep07.model and
# support/models/clib/ckep00-04.model (use_c path, no small-e guard).
# ---------------------------------------------------------------------------
def _omega_from_sc(S, C, thresh, guard=True):
    """kep01.model:3-9 / kep02.model:5-12: e = S^2 + C^2, w from arccos."""
    ecc = S ** 2 + C ** 2
    if guard and ecc < thresh:
        return ecc, 0
    with np.errstate(invalid="ignore", divide="ignore"):
        w = np.arccos(C / (ecc ** 0.5))
        if S < 0:
            w = 2 * np.pi - w
    return ecc, w


def keplerian_term(th, X, parameterisation, use_c=False):
    """Contribution of one planet, ``A*(cos(f+w) + ecc*cos(w))``.

    th: the planet's 5 parameters in the block's order (block_repo.py:23-31).
    Follows the template for `parameterisation` line by line.
    """
    p = parameterisation
    # numpy scalars, as theta[slice] elements are in the generated script
    a0, a1, a2, a3, a4 = [np.float64(v) for v in th]
    if p == 0:      # kep00.model:2-8
        per, A, phase, ecc, w = a0, a1, a2, a3, a4
        M = (2. * np.pi / per) * X + phase
    elif p == 1:    # kep01.model:1-15
        per, A, phase = a0, a1, a2
        ecc, w = _omega_from_sc(a3, a4, 1e-6, guard=not use_c)
        M = (2. * np.pi / per) * X + phase
    elif p == 2:    # kep02.model:1-22
        per = np.exp(a0)
        A = a1 ** 2 + a2 ** 2
        ecc, w = _omega_from_sc(a3, a4, 1e-5, guard=not use_c)
        with np.errstate(invalid="ignore", divide="ignore"):
            phase = np.arccos(a2 / (A ** 0.5))
            if a1 < 0:
                phase = 2 * np.pi - np.arccos(a2 / (A ** 0.5))
        M = (2. * np.pi / per) * X + phase
    elif p == 3:    # kep03.model:1-7
        per, A, tp, ecc, w = a0, a1, a2, a3, a4
        M = (2. * np.pi / per) * (X - tp)
    elif p == 4:    # kep04.model:1-16
        per, A, tp = a0, a1, a2
        ecc, w = _omega_from_sc(a3, a4, 1e-5, guard=not use_c)
        M = (2. * np.pi / per) * (X - tp)
    elif p == 6:    # kep06.model:2-10
        per = np.exp(a0)
        A, phase, ecc, w = a1, a2, a3, a4
        M = (2. * np.pi / per) * X + phase
    elif p == 7:    # kep07.model:2-18
        per = np.exp(a0)
        A, phase = a1, a2
        ecc, w = _omega_from_sc(a3, a4, 1e-6, guard=not use_c)
        M = (2. * np.pi / per) * X + phase
    else:
        raise ValueError(f"unsupported parameterisation {p}")
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        E = kepler_solve(M, np.repeat(ecc, len(M)))
        f = np.arctan(((1. + ecc) ** 0.5 / (1. - ecc) ** 0.5) * np.tan(E / 2.)) * 2.
        return A * (np.cos(f + w) + ecc * np.cos(w))


# ---------------------------------------------------------------------------
# Moving average -- support/models/moav01.model:3-15 (global, real
# recurrence) and moav00.model:4-19 (per instrument; a no-op in the reference
# because ``model0[mask][i] += MA`` writes into a copy).
# ---------------------------------------------------------------------------
def ma_residuals(t, res, coefs, taus):
    """Literal moav01 loop: returns (residuals_after, ma_added).

    residuals[i] -= sum_c coef_c*exp(-|t_i - t_{i-1-c}|/tau_c)*residuals[i-1-c]
    with the already-updated residuals (``res_`` aliases ``residuals``).
    """
    r = np.array(res, dtype=np.float64, copy=True)
    added = np.zeros_like(r)
    q = len(coefs)
    for i in range(len(r)):
        for c in range(q):
            if i > c:
                dt = abs(t[i] - t[i - 1 - c])
                MA = coefs[c] * np.exp(-dt / taus[c]) * r[i - 1 - c]
                added[i] += MA
                r[i] -= MA
    return r, added


# ---------------------------------------------------------------------------
# Priors -- support/priors/*.prior; normalisation emp/emp_model.py:82-120.
# ---------------------------------------------------------------------------
def prior_uniform(x, limits, logZ):          # Uniform.prior:1-5
    return logZ if limits[0] <= x <= limits[1] else -np.inf


def prior_normal(x, limits, args):           # Normal.prior:1-9
    lo, hi = limits
    mu, s, logZ = args
    if lo <= x <= hi:
        return -0.5 * ((x - mu) / s) ** 2 - np.log(s * np.sqrt(2 * np.pi)) - logZ
    return -np.inf


def prior_jeffreys(x, limits, args=None):    # Jeffreys.prior:1-6 (uniform in x)
    lo, hi = limits
    if lo <= x <= hi:
        return np.log(1 / (hi - lo))
    return -np.inf


def prior_beta(x, limits, args):             # Beta.prior:1-6
    from scipy.stats import beta as betapdf
    if limits[0] <= x <= limits[1]:
        with np.errstate(divide="ignore"):
            return np.log(betapdf.pdf(x, args[0], args[1]))
    return -np.inf


def prior_fixed(x, limits=None, args=None):  # Fixed.prior:1-2
    return 0.


PRIORS = {"Uniform": prior_uniform, "Normal": prior_normal, "Jeffreys": prior_jeffreys,
          "Beta": prior_beta, "Fixed": prior_fixed}


def normalise_prargs(prior, limits, prargs):
    """_check_prargs, emp/emp_model.py:82-99."""
    from scipy.stats import norm
    if prior == "Uniform":
        lo, hi = limits
        return np.log(1 / (hi - lo))
    if prior == "Normal":
        lo, hi = limits
        mu, s = prargs[0], prargs[1]
        a, b = (lo - mu) / s, (hi - mu) / s
        return [mu, s, np.log(norm.cdf(b) - norm.cdf(a))]
    return prargs


# ---------------------------------------------------------------------------
# The generated model, likelihood and prior -- emp/emp_model.py:706-781
# (my_model), support/likelihoods/00.like:1-5 (my_likelihood),
# emp/emp.py:182-254 (my_prior).
# ---------------------------------------------------------------------------
class OracleModel:
    """Spec-driven restatement of the generated ``my_model``/``my_likelihood``/``my_prior``.

    ``spec`` is the plain-dict model description that the host library
    produces (``astroemperor_amd.model.ReddModel.spec()``):

    * ``blocks``: list of dicts in block order with keys ``type``
      ('Keplerian', 'Acceleration', 'Offset', 'Jitter', 'MOAV',
      'StellarActivity', 'Sinusoid', 'MagneticCycle'), ``start``,
      ``len`` (full-vector slice), and for Keplerians ``parameterisation``,
      ``use_c`` and ``additional`` (derived-eccentricity priors);
      MOAV blocks carry ``order`` and ``global``.
    * ``params``: list (full order) of dicts with ``name``, ``prior``,
      ``limits``, ``prargs`` (already normalised), ``fixed``.
    * ``nins``, ``semantics`` ('reference' or 'intended').
    """

    def __init__(self, spec, X, Y, YERR, FLAG, SAI=None):
        self.spec = spec
        # activity columns SAI{n}_ = my_data.iloc[:, 3 + n] (emp_model.py:425-432)
        self.SAI = None if SAI is None else np.asarray(SAI, dtype=np.float64)
        self.X = np.asarray(X, dtype=np.float64)
        self.Y = np.asarray(Y, dtype=np.float64)
        self.YERR = np.asarray(YERR, dtype=np.float64)
        self.FLAG = np.asarray(FLAG)
        self.ndat = len(self.X)
        self.nins = int(spec["nins"])
        self.masks = [(self.FLAG == n + 1) for n in range(self.nins)]
        fixed = [p.get("fixed") for p in spec["params"]]
        self.A_ = np.array([i for i, f in enumerate(fixed) if f is not None], dtype=int)
        self.C_ = np.array([i for i, f in enumerate(fixed) if f is None], dtype=int)
        self.mod_fixed_ = fixed
        self.likelihood_constant = -0.5 * np.log(2 * np.pi) * self.ndat   # 00.like:1

    def full_theta(self, theta):
        """emp_model.py:709-711: ``for a in A_: theta = np.insert(theta, a, mod_fixed_[a])``."""
        theta = np.asarray(theta, dtype=np.float64)
        for a in self.A_:
            theta = np.insert(theta, a, self.mod_fixed_[a])
        return theta

    def my_model(self, theta):
        th = self.full_theta(theta)
        X, Y = self.X, self.Y
        model0 = np.zeros(self.ndat)                                  # emp_model.py:713
        err20 = self.YERR ** 2                                         # emp_model.py:714
        for b in self.spec["blocks"]:
            sl = th[b["start"]:b["start"] + b["len"]]
            typ = b["type"]
            if typ == "Keplerian":                                     # kep0p.model
                model0 = model0 + keplerian_term(sl, X, b["parameterisation"], b.get("use_c", False))
            elif typ == "Offset":                                      # offset00.model:3
                for nin in range(self.nins):
                    model0 = model0 + sl[nin] * self.masks[nin]
            elif typ == "Acceleration":                                # acc.model:2
                model0 = model0 + np.polyval(np.concatenate([sl, [0]]), (X - X[0]))
            elif typ == "MOAV":                                        # emp_model.py:757-769
                residuals = Y - model0
                q = int(b["order"])
                if q <= 0:
                    continue
                if b["global"]:                                        # moav01.model:3-15
                    coefs = [sl[2 * c] for c in range(q)]
                    taus = [sl[2 * c + 1] for c in range(q)]
                    _, added = ma_residuals(X, residuals, coefs, taus)
                    model0 = model0 + added
                elif self.spec.get("semantics", "reference") == "intended":
                    # moav01 recurrence applied to each instrument's time-ordered
                    # subsequence with that instrument's own (coef, tau) per order.
                    for nin in range(self.nins):
                        m = self.masks[nin]
                        th_ma = sl[q * 2 * nin:q * 2 * (nin + 1)]
                        coefs = [th_ma[2 * c] for c in range(q)]
                        taus = [th_ma[2 * c + 1] for c in range(q)]
                        _, added = ma_residuals(X[m], residuals[m], coefs, taus)
                        model0[m] = model0[m] + added
                # reference semantics: moav00.model:18-19 writes into boolean-mask
                # copies, so neither model0 nor the residuals change.
            elif typ == "Jitter":                                      # jitter00.model:3
                for nin in range(self.nins):
                    err20 = err20 + self.masks[nin] * sl[nin] ** 2
            elif typ == "StellarActivity":                             # emp_model.py:736-745
                for n in range(b["len"]):                              # sai00.model:3
                    model0 = model0 + sl[n] * self.SAI[:, n]
            elif typ == "Sinusoid":                                    # sinusoid00.model:2-7
                per, A, phase = sl
                model0 = model0 + A * np.cos(2. * np.pi / per * X + phase)
            elif typ == "MagneticCycle":                               # magneticcycle00.model:2-10
                per, A1, A2, phase1, phase2 = sl
                M1 = np.pi / per * X + phase1
                M2 = 2. * np.pi / per * X + phase2
                model0 = model0 + (A1 * np.cos(M1) + A2 * np.cos(M2))
            else:
                raise ValueError(f"unsupported block {typ}")
        return model0, err20

    def my_likelihood(self, theta):
        """support/likelihoods/00.like:3-5."""
        model, err2 = self.my_model(theta)
        with np.errstate(invalid="ignore", divide="ignore"):
            return float(-0.5 * (np.sum((self.Y - model) ** 2 / err2 + np.log(err2)))
                         + self.likelihood_constant)

    def my_prior(self, theta):
        """emp/emp.py:182-254 as generated for reddemcee."""
        th = self.full_theta(theta)
        params = self.spec["params"]
        lp = 0.
        for b in self.spec["blocks"]:
            for j in range(b["start"], b["start"] + b["len"]):
                p = params[j]
                if p.get("fixed") is not None:
                    continue
                lp += PRIORS[p["prior"]](th[j], p["limits"], p["prargs"])
            if lp == -np.inf:                                         # emp.py:203-206
                return lp
            for add in b.get("additional", []):                       # emp.py:208-248
                if add["kind"] == "Ecc":
                    sl = th[b["start"]:b["start"] + b["len"]]
                    x = sl[3] ** 2 + sl[4] ** 2
                    lp += PRIORS[add["prior"]](x, add["limits"], add["prargs"])
        return lp

    def log_prob(self, theta, beta=1.0):
        lp = self.my_prior(theta)
        if not np.isfinite(lp):
            return -np.inf, -np.inf, lp
        ll = self.my_likelihood(theta)
        return lp + beta * ll, ll, lp


# ---------------------------------------------------------------------------
# Sampler moves with injected randoms.  reddemcee/emcee are third-party and
# absent; these follow the published methods named at their call sites
# (emp/emp.py:576-593 builds reddemcee.PTSampler without `moves`, so emcee's
# default StretchMove with a=2 applies; swaps/ladder per Vousden, Farr &
# Mandel 2016, whose adaptation knobs are adapt_tau/adapt_nu, emp.py:2370-2380).
# ---------------------------------------------------------------------------
def stretch_z(u, a=2.0):
    """Goodman & Weare (2010) eq. 9 draw: z = ((a-1) u + 1)^2 / a."""
    return ((a - 1.0) * u + 1.0) ** 2 / a


def stretch_halfstep(X, lnl, lnp, beta, active, partner, u_z, u_acc, logpost_fn, a=2.0):
    """One half of the red/blue stretch move at one temperature.

    X[W, D], lnl[W], lnp[W] are updated in place for the walkers in `active`.
    partner[k] is the complementary walker index for active walker k; u_z and
    u_acc are the uniforms that the sampler drew.  Returns accepted mask.
    """
    D = X.shape[1]
    acc = np.zeros(len(active), dtype=bool)
    for n, k in enumerate(active):
        z = stretch_z(u_z[n], a)
        j = partner[n]
        Y = X[j] + z * (X[k] - X[j])
        lp_new = logpost_fn.prior(Y)
        if not np.isfinite(lp_new):
            continue
        ll_new = logpost_fn.like(Y)
        if not np.isfinite(ll_new):
            ll_new = -np.inf
        tempered_new = beta * ll_new if (beta != 0 or np.isfinite(ll_new)) else 0.0
        tempered_old = beta * lnl[k] if (beta != 0 or np.isfinite(lnl[k])) else 0.0
        if beta == 0:
            tempered_new = tempered_old = 0.0
        lnalpha = (D - 1) * np.log(z) + tempered_new - tempered_old + lp_new - lnp[k]
        if np.log(u_acc[n]) < lnalpha:
            X[k] = Y
            lnl[k] = ll_new
            lnp[k] = lp_new
            acc[n] = True
    return acc


def swap_pair(Xh, lnlh, lnph, Xc, lnlc, lnpc, beta_h, beta_c, perm, u):
    """Adjacent-temperature swap between hotter chain h and colder chain c.

    Walker i of the hot chain is paired with walker perm[i] of the cold chain
    and the pair is exchanged with probability
    min(1, exp((beta_c - beta_h) * (lnl_h[i] - lnl_c[perm[i]]))).
    Returns (n_accepted, mean |log acceptance ratio| over pairs).
    """
    W = len(lnlh)
    dbeta = beta_c - beta_h
    nacc = 0
    dist = 0.0
    for i in range(W):
        j = perm[i]
        lr = dbeta * (lnlh[i] - lnlc[j])
        dist += abs(lr)
        if np.log(u[i]) < lr:
            Xh[i], Xc[j] = Xc[j].copy(), Xh[i].copy()
            lnlh[i], lnlc[j] = lnlc[j], lnlh[i]
            lnph[i], lnpc[j] = lnpc[j], lnph[i]
            nacc += 1
    return nacc, dist / W


def lnl_spread(lnl_by_temp):
    """Per-temperature sample standard deviation (n - 1) of the finite lnL of
    each row of lnl_by_temp[T, W]; 0 with fewer than two finite values."""
    out = np.zeros(len(lnl_by_temp))
    for t, row in enumerate(np.asarray(lnl_by_temp, dtype=np.float64)):
        f = row[np.isfinite(row)]
        if len(f) > 1:
            out[t] = np.sqrt(np.sum((f - f.mean()) ** 2) / (len(f) - 1))
    return out


def ladder_update(betas, swap_rates, iteration, adapt_tau, adapt_nu, mode=0, smd=None, spread=None):
    """Ladder dynamics after one sweep; the coldest (beta = 1) and hottest
    temperatures stay fixed.

    EMPEROR passes engine_config['adapt_mode'] to reddemcee
    (emp/emp.py:806, 2378) and labels mode 2 "Specific Heat"
    (tests/02_51peg_advanced.ipynb); reddemcee itself is not in the
    reference.  S_i = ln(T_i - T_{i-1}), i = 1..T-2, moves by kappa times a
    balance term, kappa = tau / (t + tau) / nu:
      mode 0 (Vousden, Farr & Mandel 2016): A_{i-1} - A_i with the pair
        acceptances `swap_rates`;
      mode 1: (d_i - d_{i-1}) / (d_i + d_{i-1}) with the pair mean |log swap
        ratio| `smd`;
      mode 2: the same with the thermodynamic length of each pair,
        l_k = (beta_k - beta_{k+1}) (sig_k + sig_{k+1}) / 2, sig the
        per-temperature lnL spread `spread` (lnl_spread of the post-swap
        state);
    the normalised terms are 0 when the two distances sum to 0 or overflow.
    """
    betas = np.array(betas, dtype=np.float64)
    T = len(betas)
    if T < 3 or mode < 0:
        return betas
    kappa = adapt_tau / (iteration + adapt_tau) / adapt_nu
    if mode == 0:
        A = np.asarray(swap_rates, dtype=np.float64)
        dS = kappa * (A[:-1] - A[1:])
    else:
        if mode == 1:
            d = np.asarray(smd, dtype=np.float64)
        else:
            sig = np.asarray(spread, dtype=np.float64)
            d = (betas[:-1] - betas[1:]) * 0.5 * (sig[:-1] + sig[1:])
        lo, hi = d[:-1], d[1:]
        tot = lo + hi
        ok = (tot > 0) & np.isfinite(tot)
        dS = np.zeros(T - 2)
        dS[ok] = kappa * (hi[ok] - lo[ok]) / tot[ok]
    with np.errstate(divide="ignore"):
        temps = 1.0 / betas
    dT = np.diff(temps[:-1]) * np.exp(dS)
    newT = temps[0] + np.cumsum(dT)
    betas[1:-1] = 1.0 / newT
    return betas


# ---------------------------------------------------------------------------
# Gradient move with injected randoms.  No reference counterpart: EMPEROR
# leaves reddemcee's `moves` at None (emp/emp.py:2373), so this is the
# published method itself -- Metropolis-adjusted Langevin (Roberts & Tweedie
# 1996) preconditioned by the covariance of the complementary half of the
# ensemble (Leimkuhler, Matthews & Weare 2018).
# ---------------------------------------------------------------------------
def langevin_factor(C, ridge=1e-6):
    """Preconditioner from the complementary walkers C[n, D]: the Cholesky
    factor L of their population covariance with the diagonal times 1 + ridge.
    Parameters without spread (zero variance) get a unit pivot, no coupling,
    and are inactive (they do not move).  A factorisation that fails falls back
    to the diagonal of standard deviations.  Returns (L, active mask)."""
    C = np.asarray(C, dtype=np.float64)
    mu = C.mean(0)
    Z = C - mu
    S = Z.T @ Z / len(C)
    var = np.diag(S).copy()
    act = var > 0
    A = S.copy()
    A[~act, :] = 0.0
    A[:, ~act] = 0.0
    A[np.diag_indices_from(A)] = np.where(act, var * (1.0 + ridge), 1.0)
    try:
        L = np.linalg.cholesky(A)
    except np.linalg.LinAlgError:
        L = np.diag(np.where(act, np.sqrt(np.where(act, var, 1.0)), 1.0))
    return L, act


def langevin_logq(x, y, gx, L, act, beta, eps):
    """log q(y | x) up to the constant shared by both directions: y - x =
    L((eps^2/2) beta L^T g(x) + eps xi) gives -|xi|^2 / 2 over the active
    parameters; a non-finite g(x) switches the drift off."""
    drift = np.all(np.isfinite(gx))
    v = L.T @ gx if drift else np.zeros_like(x)
    r = np.linalg.solve(L, y - x)
    xi = np.where(act, r / eps - 0.5 * eps * beta * v, 0.0)
    return -0.5 * float(np.sum(xi ** 2))


def langevin_halfstep(X, lnl, lnp, grad, beta, active, complement, xi, u_acc, logpost_fn, eps):
    """One half of the gradient move at one temperature.

    X[W, D], lnl[W], lnp[W] and grad[W, D] (grad lnL at X) are updated in
    place for the walkers in `active`; `complement` lists the other half.
    xi[n, D] are the normals and u_acc[n] the uniform that the sampler drew for
    active walker n.  logpost_fn.prior / .like / .grad give ln prior, lnL and
    grad lnL.  The proposal is y = x + L u, u = (eps^2/2) beta L^T g(x) + eps xi,
    accepted with probability min(1, exp(beta (lnL(y) - lnL(x)) + ln pi(y) -
    ln pi(x) + log q(x | y) - log q(y | x))).  Returns the accepted mask."""
    L, act = langevin_factor(X[complement])
    acc = np.zeros(len(active), dtype=bool)
    for n, k in enumerate(active):
        x = X[k].copy()
        g = grad[k]
        drift = np.all(np.isfinite(g))
        v = L.T @ g if drift else np.zeros_like(x)
        u = np.where(act, 0.5 * eps * eps * beta * v + eps * xi[n], 0.0)
        y = x + L @ u
        lq_f = -0.5 * float(np.sum(np.where(act, xi[n], 0.0) ** 2))
        lp_new = logpost_fn.prior(y)
        if not lp_new > -np.inf:
            continue
        ll_new = logpost_fn.like(y)
        if not (ll_new == ll_new) or ll_new == np.inf:
            ll_new = -np.inf
        g_new = np.asarray(logpost_fn.grad(y), dtype=np.float64)
        lq_r = langevin_logq(y, x, g_new, L, act, beta, eps)
        tn = 0.0 if beta == 0 else beta * ll_new
        to = 0.0 if beta == 0 else beta * lnl[k]
        with np.errstate(invalid="ignore"):
            lnalpha = (tn - to) + (lp_new - lnp[k]) + (lq_r - lq_f)
        if np.log(u_acc[n]) < lnalpha:
            X[k] = y
            lnl[k] = ll_new
            lnp[k] = lp_new
            grad[k] = g_new
            acc[n] = True
    return acc
