"""Run statistics, model comparison and constrain-next-run.

Host mirror of the parts of EMPEROR's post-processing that the ``autorun``
loop needs (emp/emp.py:1112-1212 criteria and statistics, 1401-1560 samples
and fits, 1720-1800 per-parameter values, 2414-2456 and 2654-2778 the loop's
continue test, initial positions and constraints).  The chains come from the
device engine; the model at the point estimate and the lnL at the posterior
mean are evaluated on the device.
"""
from __future__ import annotations

import numpy as np


def hdi_of_samples(samples, cred_mass=0.9):
    """Shortest interval holding floor(cred_mass * n) + 1 of the sorted
    samples (emp/math_utils.py:72-85)."""
    lo, hi = hdi_of_chain(np.asarray(samples, dtype=np.float64)[:, None], cred_mass)[0]
    return lo, hi


def hdi_of_chain(chain, cred_mass=0.9):
    """Per column of a [steps, dims] chain: (low, high) of the shortest such
    interval (emp/math_utils.py:88-115); the first of equally short ones."""
    s = np.sort(np.asarray(chain, dtype=np.float64), axis=0)
    n = s.shape[0]
    k = int(np.floor(cred_mass * n))
    widths = s[k:] - s[:n - k]
    i = np.argmin(widths, axis=0)
    cols = np.arange(s.shape[1])
    return np.stack([s[i, cols], s[i + k, cols]], axis=1)


class StatsMixin:
    """Criteria, statistics and the autorun bookkeeping of ``Simulation``."""

    STATS_POSI = ["chi2", "chi2_red", "AIC", "BIC", "DIC", "HQIC", "RMSE", "RMSi", "Weights"]
    STATS_NEGA = ["post_max", "like_max", "BayesFactor", "Evidence"]
    _MIN_CRIT = ("chi2", "chi2_red", "AIC", "BIC", "DIC", "HQIC", "RMSE")
    _MAX_CRIT = ("post_max", "like_max", "BayesFactor", "Evidence")

    def _init_stats(self):
        for s in self.STATS_POSI:
            setattr(self, s, np.inf)
        for s in self.STATS_NEGA:
            setattr(self, s, -np.inf)
        self.evidence = (-np.inf, np.inf)
        self.use_fit = "max_post"                                        # emp.py:105
        self.constrain = {"method": "range", "sigma": 3, "types": ["Keplerian", "Jitter"],
                          "tol": 1e-4, "known_methods": ["sigma", "GM", "range", "None"]}
        self.switch_evidence = True
        # EMPEROR's default (emp.py:129).  reddemcee's hybrid estimator is not
        # available here: get_evidence_hybrid raises, and the flow falls back
        # to classic thermodynamic integration as emp.py:1443-1446 does.
        self.evidence_method = "hybrid"
        self._crit_tolerance = 5
        self._crit_current = "BIC"
        self._crit_msg = ""

    # -- criteria (emp.py:1112-1171) ---------------------------------------
    def set_comparison_criteria(self, val):
        if val in self._MIN_CRIT or val in self._MAX_CRIT or val == "Pass":
            self._crit_current = val

    def set_tolerance(self, val):
        self._crit_tolerance = val

    def criteria_compare(self, new, old):
        c = self._crit_current
        if c == "Pass":
            self._crit_msg = "Criteria set to PASS"
            return True
        if c in self._MAX_CRIT:
            self._crit_msg = f"{np.round(new, 3)} - {np.round(old, 3)} > {self._crit_tolerance}"
            return new - old > self._crit_tolerance
        self._crit_msg = f"{np.round(old, 3)} - {np.round(new, 3)} > {self._crit_tolerance}"
        return old - new > self._crit_tolerance

    def _stat_holder_update(self):                                      # emp.py:1173-1185
        for s in self.STATS_POSI + ["like_max", "post_max", "BayesFactor"]:
            setattr(self, f"stats_old_{s}", getattr(self, s))
        self.stats_old_Evidence = self.evidence[0]

    def _check_continue_run(self):
        """True when the loop must stop (emp.py:2654-2668)."""
        new = getattr(self, self._crit_current) if self._crit_current != "Pass" else 0.0
        old = getattr(self, f"stats_old_{self._crit_current}", 0.0)
        if not self.criteria_compare(new, old):
            return True
        return self.k_start > self.k_end

    # -- samples, fit, parameter values (emp.py:1401-1560, 1720-1753) -------
    def _postprocess(self):
        kw = dict(discard=self.reddemcee_discard, thin=self.reddemcee_thin, flat=True)
        self.chain = self.sampler.get_chain(**kw)
        self.likes = self.sampler.get_log_like(**kw)
        self.posts = self.sampler.get_log_prob(**kw)
        if self.switch_evidence:
            self.evidence = self._get_evidence()
        self.best_loc_post = int(np.argmax(self.posts[0]))
        self.best_loc_like = int(np.argmax(self.likes[0]))
        c0 = self.chain[0]
        self.post_max = float(self.posts[0][self.best_loc_post])
        self.like_max = float(self.likes[0][self.best_loc_post])
        self.prior_max_post = self.post_max - self.like_max
        self.sigmas = np.std(c0, axis=0)
        self.fit_max = c0[self.best_loc_post]
        self.fit_mean = np.mean(c0, axis=0)
        self.fit_median = np.median(c0, axis=0)
        self.fit_low1, self.fit_high1 = hdi_of_chain(c0, 0.36).T
        self.fit_low2, self.fit_high2 = hdi_of_chain(c0, 0.90).T
        self.fit_low3, self.fit_high3 = hdi_of_chain(c0, 0.95).T
        self.fit_maxlike = c0[self.best_loc_like]
        self.ajuste = {"max_post": self.fit_max, "max_like": self.fit_maxlike, "mean": self.fit_mean,
                       "median": self.fit_median}.get(self.use_fit, self.fit_max)
        self._set_p_values()
        self._calculate_statistics()
        self.results.append(dict(k=self.kplanets__, like_max=self.like_max, post_max=self.post_max,
                                 BIC=self.BIC, AIC=self.AIC, HQIC=self.HQIC, DIC=self.DIC,
                                 evidence=tuple(self.evidence), fit=np.asarray(self.ajuste).tolist()))

    def _get_evidence(self):
        d = self.reddemcee_discard
        try:
            z = getattr(self.sampler, f"get_evidence_{self.evidence_method}")(discard=d)[:2]
            if z[0] == -np.inf:
                z = self.sampler.get_evidence_ti(discard=d)[:2]
            if z[0] == -np.inf:
                z = self.sampler.get_evidence_ti(discard=d, pchip=False)[:2]
        except Exception:
            print("Thermodynamic Integration interpolation failed! Using classic method...")
            z = self.sampler.get_evidence_ti(discard=d, pchip=False)[:2]
        return float(z[0]), float(z[1])

    def _set_p_values(self):
        j = 0
        for b in self.blocks__:
            for p in b:
                if p.fixed is not None:
                    self._update_p_fixed(p)
                else:
                    self._update_p_free(p, j)
                    j += 1

    def _update_p_free(self, p, j):                                      # emp.py:1720-1753
        p.cpointer = j
        p.value = float(self.ajuste[j])
        p.sigma = float(self.sigmas[j])
        p.value_max = float(self.fit_max[j])
        p.value_mean = float(self.fit_mean[j])
        p.value_median = float(self.fit_median[j])
        p.value_max_lk = float(self.fit_maxlike[j])
        p.value_low1, p.value_high1 = float(self.fit_low1[j]), float(self.fit_high1[j])
        p.value_low2, p.value_high2 = float(self.fit_low2[j]), float(self.fit_high2[j])
        p.value_low3, p.value_high3 = float(self.fit_low3[j]), float(self.fit_high3[j])
        a = p.value_low1 if p.value_low1 < p.value_max else p.value_max - p.sigma
        b = p.value_high1 if p.value_high1 > p.value_max else p.value_max + p.sigma
        p.value_range = [max(a, p.limits[0]), min(b, p.limits[1])]

    @staticmethod
    def _update_p_fixed(p):                                              # emp.py:1756-1767
        p.sigma = np.nan
        p.value_max = p.value_mean = p.value_median = p.value_max_lk = p.value
        for s in (1, 2, 3):
            setattr(p, f"value_low{s}", np.nan)
            setattr(p, f"value_high{s}", np.nan)
        p.value_range = [np.nan, np.nan]

    # -- statistics (emp.py:1186-1208, 1543-1558) ---------------------------
    def _calculate_statistics(self):
        m = self.model
        ymod, err2 = m.my_model(self.ajuste)
        residuals = self.my_data.Y - ymod
        n, k = self.my_data.ndat, m.ndim__
        self.dof = n - k
        self.chi2 = float(np.sum(residuals ** 2 / err2))
        self.chi2_red = self.chi2 / self.dof
        self.RMSE = float(np.sqrt(np.sum(residuals ** 2) / n))
        flags = self.my_data.FLAG
        self.RMSi = [float(np.sqrt(np.sum(residuals[flags == i + 1] ** 2) / np.sum(flags == i + 1)))
                     for i in range(self.nins__)]
        # emp.py:1195 weighs by len(FLAGS_ == i + 1), the length of the mask (n)
        w = np.array([n / r ** 2 for r in self.RMSi])
        self.Total_Weight = float(np.sum(w))
        self.Weights = w / self.Total_Weight
        self.AIC = 2 * k - 2 * self.like_max
        self.BIC = np.log(n) * k - 2 * self.like_max
        self.dic_aux = float(np.var(-2 * self.likes[0]))
        self.DIC = -2 * float(m.my_likelihood(self.fit_mean)) + self.dic_aux
        self.Evidence = self.evidence[0]
        self.post_true = self.post_max - self.Evidence
        self.BayesFactor = self.like_max - self.Evidence
        self.HQIC = 2 * k * np.log(np.log(n)) - 2 * self.like_max

    # -- next run (emp.py:2671-2778) -----------------------------------------
    def _set_init_pos(self):
        for b in self.blocks__:
            if b.type_ == "Keplerian":
                for p in b:
                    if p.fixed is None:
                        p.init_pos = list(p.value_range)

    def _constrain_next_run(self):
        method = self.constrain["method"]
        if method == "None":
            return
        if method == "GM":
            raise NotImplementedError("Gaussian-mixture priors are not on the device path; "
                                      "use constrain method 'range', 'sigma' or 'None'")
        if method not in self.constrain["known_methods"]:
            raise ValueError(f"constrain method {method!r} not identified")
        getattr(self, f"_apply_constrain_{method}")()

    def _constrained_limits(self, b, p, limf, limc):
        limc = min(limc, p.limits[1])
        if limf < p.limits[0] or b.type_ == "Jitter":
            limf = p.limits[0]
        if (limc - limf) / abs(p.value) >= self.constrain["tol"]:
            self.add_condition([p.name, "limits", [limf, limc]])

    def _apply_constrain_sigma(self):                                   # emp.py:2713-2731
        for b in self.blocks__:
            if b.type_ in self.constrain["types"]:
                for p in b:
                    if p.fixed is None:
                        s = self.constrain["sigma"]
                        self._constrained_limits(b, p, p.value - s * p.sigma, p.value + s * p.sigma)

    def _apply_constrain_range(self):                                   # emp.py:2752-2770
        for b in self.blocks__:
            if b.type_ in self.constrain["types"]:
                for p in b:
                    if p.fixed is None:
                        s = self.constrain["sigma"]
                        self._constrained_limits(b, p, getattr(p, f"value_low{s}"), getattr(p, f"value_high{s}"))
