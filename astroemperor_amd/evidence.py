"""Evidence estimators over a parallel-tempering run.

EMPEROR asks the sampler for ``get_evidence_{method}(discard=...)`` and keeps
(log Z, error) (emp/emp.py:1432-1447), falling back to ``get_evidence_ti`` with
``pchip=False`` when the chosen method fails.  reddemcee is third-party and
absent; these are the published classic estimators, fed by the device
reduction ``emp_pt_lnl_moments`` (mean lnL per temperature and the
stepping-stone log ratios):

* thermodynamic integration, log Z = int_0^1 <lnL>_beta d beta (Lartillot &
  Philippe 2006), by the trapezoid rule over the ladder or by integrating a
  monotone cubic (PCHIP, Fritsch & Carlson 1980) through <lnL>(beta).  A
  ladder that stops short of beta = 0 takes the hottest chain's mean there.
  The error is the difference from the same rule over every other
  temperature.
* stepping stones (Xie et al. 2011), log Z = sum_t log <exp((beta_{t-1} -
  beta_t) lnL)>_{beta_t}, with the trapezoid over [0, beta_min] when the
  hottest beta is above 0.  The error adds the stones' delta-method variances
  (autocorrelation ignored, so it is a lower bound).
"""
from __future__ import annotations

import numpy as np


def _ascending(betas, mean):
    betas = np.asarray(betas, dtype=np.float64)
    mean = np.asarray(mean, dtype=np.float64)
    i = np.argsort(betas)
    b, m = betas[i], mean[i]
    if b[0] > 0.0:
        b = np.concatenate([[0.0], b])
        m = np.concatenate([[m[0]], m])
    return b, m


def _integral(b, m, pchip):
    if pchip and len(b) >= 3:
        from scipy.interpolate import PchipInterpolator
        return float(PchipInterpolator(b, m).integrate(b[0], b[-1]))
    return float(np.sum(0.5 * (m[1:] + m[:-1]) * np.diff(b)))


def ti_log_evidence(betas, mean_lnl, pchip=True):
    """(log Z, error) by thermodynamic integration."""
    b, m = _ascending(betas, mean_lnl)
    z = _integral(b, m, pchip)
    keep = np.zeros(len(b), dtype=bool)
    keep[::2] = True
    keep[-1] = True
    z2 = _integral(b[keep], m[keep], pchip)
    return z, abs(z - z2)


def ss_log_evidence(betas, mean_lnl, lse, var):
    """(log Z, error) by stepping stones.  betas, mean_lnl, lse, var are in
    ladder order (coldest first), as emp_pt_lnl_moments returns them: lse[t]
    is the stone from betas[t] to betas[t - 1]."""
    betas = np.asarray(betas, dtype=np.float64)
    if np.any(np.diff(betas) >= 0):
        raise ValueError("stepping stones need a strictly decreasing ladder (coldest first)")
    z = float(np.sum(np.asarray(lse, dtype=np.float64)[1:]))
    if betas[-1] > 0.0:
        z += betas[-1] * float(mean_lnl[-1])
    return z, float(np.sqrt(np.sum(np.asarray(var, dtype=np.float64)[1:])))
