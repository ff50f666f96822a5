"""Parameter blocks and data-driven default priors.

Host-side mirror of the reference's model-spec layer: parameter metadata
(emp/param_repo.py:15-175), ``Parameter``/``Parameter_Block``
(emp/emp_model.py:16-160), the block factories (emp/block_repo.py:22-240)
and ``SmartSetter`` (emp/block_repo.py:505-867).  Only the blocks on the RV
hot path are provided: Keplerian (all parameterisations the RV templates
support), Acceleration, Offset, Jitter, MOAV, StellarActivity (SAI),
Sinusoid and MagneticCycle.

These objects carry names, limits, priors and fixed values; the numerical
work happens in the HIP library, which receives a flat descriptor built by
:mod:`astroemperor_amd.model`.
"""
from __future__ import annotations

import copy
import itertools
import math

import numpy as np

# ---------------------------------------------------------------------------
# Parameter metadata -- names and flags as in emp/param_repo.py:18-175.
# ---------------------------------------------------------------------------
_DEFAULTS = {
    "prior": "Uniform", "limits": [None, None], "init_pos": [None, None],
    "value": -np.inf, "fixed": None, "prargs": None, "type": None,
    "ptformargs": None, "sigma": None, "display_prior": "",
}

_META = {
    "dPeriod": ("Period", "P", False, False),
    "dAmplitude": ("Amplitude", "K", False, False),
    "dPhase": ("Phase", "M_0", True, False),
    "dEccentricity": ("Eccentricity", "e", False, False),
    "dLongitude": ("Longitude", "w", True, False),
    "dlPeriod": ("lPeriod", "lnP", False, False),
    "dAmp_sin": ("Amp_sin", "K_sin", False, True),
    "dAmp_cos": ("Amp_cos", "K_cos", False, True),
    "dEcc_sin": ("Ecc_sin", "e_sin", False, True),
    "dEcc_cos": ("Ecc_cos", "e_cos", False, True),
    "dT_0": ("T_0", "T_0", False, False),
    "dM0": ("M0", "M_0", False, False),
    "dOffset": ("Offset", "gamma_0", False, False),
    "dJitter": ("Jitter", "J", False, False),
    "dMACoefficient": ("MACoefficient", "Phi", False, False),
    "dMATimescale": ("MATimescale", "tau", False, False),
    "dAcceleration": ("Acceleration", "gamma", False, False),
    "dStaract": ("Staract", "SA", False, False),
    "dSMA": ("Semi-Major Axis", "SMA", False, False),
    "dMinM": ("Minimum Mass", "MinM", False, False),
}


def make_parameter(key):
    name, mini, circ, hou = _META[key]
    d = copy.deepcopy(_DEFAULTS)
    d.update(name=name, mininame=mini, is_circular=circ, is_hou=hou)
    return d


class Parameter:
    """Attribute bag for one parameter (emp/emp_model.py:16-31)."""

    def __init__(self, attributes):
        self.name = None
        for k, v in attributes.items():
            setattr(self, k, v)

    def __repr__(self):
        return str(self.name)


# Parameter lists per Keplerian parameterisation, emp/block_repo.py:23-31.
KEPLERIAN_PARAMS = {
    0: ["dPeriod", "dAmplitude", "dPhase", "dEccentricity", "dLongitude"],
    1: ["dPeriod", "dAmplitude", "dPhase", "dEcc_sin", "dEcc_cos"],
    2: ["dlPeriod", "dAmp_sin", "dAmp_cos", "dEcc_sin", "dEcc_cos"],
    3: ["dPeriod", "dAmplitude", "dT_0", "dEccentricity", "dLongitude"],
    4: ["dPeriod", "dAmplitude", "dT_0", "dEcc_sin", "dEcc_cos"],
    6: ["dlPeriod", "dAmplitude", "dPhase", "dEccentricity", "dLongitude"],
    7: ["dlPeriod", "dAmplitude", "dPhase", "dEcc_sin", "dEcc_cos"],
}
SC_PARAMETERISATIONS = (1, 2, 4, 7)   # e and w derived from (S, C)


class Parameter_Block:
    """A block of parameters with its free/fixed bookkeeping (emp/emp_model.py:33-200)."""

    def __init__(self, params, attributes):
        self.list_ = list(params)
        self.additional_priors = []
        self.additional_parameters = []
        self.C_, self.A_ = [], []
        for k, v in attributes.items():
            setattr(self, k, v)
        self.refresh_block()

    # emp_model.py:58-79
    def refresh_block(self):
        fixed = [p.fixed for p in self]
        self.C_ = [i for i, f in enumerate(fixed) if f is None]
        self.A_ = [i for i, f in enumerate(fixed) if f is not None]
        for i in self.A_:
            p = self[i]
            p.value = p.fixed
            p.prior = "Fixed"
            p.limits = [np.nan, np.nan]
        self.ndim_ = len(self.C_)

    # emp_model.py:82-120 (normalising constants of the priors)
    @staticmethod
    def _normalise(p):
        if p.prior == "Uniform":
            lo, hi = p.limits
            p.prargs = float(np.log(1 / (hi - lo)))
        elif p.prior == "Normal":
            from scipy.stats import norm
            lo, hi = p.limits
            mu, s = p.prargs[0], p.prargs[1]
            a, b = (lo - mu) / s, (hi - mu) / s
            p.prargs = [mu, s, float(np.log(norm.cdf(b) - norm.cdf(a)))]

    def check_prargs(self):
        for p in self:
            self._normalise(p)
        for p in self.additional_parameters:
            if getattr(p, "has_prior", False):
                self._normalise(p)

    def add_additional_parameters(self, p):
        self.additional_parameters.append(p)

    def get_attr(self, name):
        return [getattr(p, name) for p in self]

    def set_attr(self, name, values):
        for p, v in zip(self, values):
            setattr(p, name, v)

    def __len__(self):
        return len(self.list_)

    def __iter__(self):
        return iter(self.list_)

    def __getitem__(self, n):
        return self.list_[n]

    def __repr__(self):
        return f"{self.name_}({self.ndim_}): {self.list_}"


def _named(key, suffix):
    d = make_parameter(key)
    if suffix:
        d["name"] += suffix
        d["mininame"] += suffix
    return Parameter(d)


def KeplerianBlock(parameterisation, number, use_c=False):
    """emp/block_repo.py:22-72 (RV parameterisations only)."""
    if parameterisation not in KEPLERIAN_PARAMS:
        raise ValueError(f"Keplerian parameterisation {parameterisation} is not an RV "
                         f"parameterisation (supported: {sorted(KEPLERIAN_PARAMS)})")
    if use_c and parameterisation > 4:
        raise ValueError("use_c supports parameterisations 0-4 only (emp/block_repo.py:63-67)")
    suffix = f" {number}" if number else ""
    params = [_named(k, suffix) for k in KEPLERIAN_PARAMS[parameterisation]]
    return Parameter_Block(params, {
        "name_": f"KeplerianBlock {number}", "type_": "Keplerian", "class_": "Model",
        "parameterisation": parameterisation, "number_": number, "use_c": bool(use_c),
        "dynamics_bool": False})


def AccelerationBlock(order):
    """emp/block_repo.py:137-173: polynomial of `order` in (X - X[0])."""
    params = []
    for i in range(order):
        d = make_parameter("dAcceleration")
        if i > 0:
            d["name"] = f"Acceleration Order {i + 1}"
        params.append(Parameter(d))
    return Parameter_Block(params, {"name_": f"AccelerationBlock o{order}",
                                    "type_": "Acceleration", "class_": "Data",
                                    "number_": order})


def OffsetBlock(nins):
    """emp/block_repo.py:103-134."""
    params = [_named("dOffset", f" {i + 1}") for i in range(nins)]
    return Parameter_Block(params, {"name_": "OffsetBlock", "type_": "Offset",
                                    "class_": "Data", "number_": nins})


def JitterBlock(nins):
    """emp/block_repo.py:176-203."""
    params = []
    for i in range(nins):
        d = make_parameter("dJitter")
        d["name"] += f" {i + 1}"
        params.append(Parameter(d))
    return Parameter_Block(params, {"name_": "JitterBlock", "type_": "Jitter",
                                    "class_": "Data", "number_": nins})


def MOAVBlock(nins, moav):
    """emp/block_repo.py:206-240: (coef, timescale) per instrument (or one set
    when global) per order, ordered instrument-major then order."""
    order, is_global = int(moav["order"]), bool(moav["global"])
    reps = 1 if is_global else nins
    params = []
    for i, j in itertools.product(range(reps), range(order)):
        c = make_parameter("dMACoefficient")
        c["name"] += f" {i + 1} Order {j + 1}"
        t = make_parameter("dMATimescale")
        t["name"] += f" {i + 1} Order {j + 1}"
        params.extend((Parameter(c), Parameter(t)))
    return Parameter_Block(params, {"name_": "MOAVBlock", "type_": "MOAV", "class_": "Data",
                                    "number_": nins, "moav": order, "is_global": is_global})


def SAIBlock(nins, cornums):
    """emp/block_repo.py:243-272: one coefficient per activity column,
    'Staract i j' for column j of instrument i, instrument-major."""
    params = []
    for i in range(nins):
        for j in range(int(cornums[i])):
            d = make_parameter("dStaract")
            d["name"] += f" {i + 1} {j + 1}"
            d["mininame"] += f"_{i + 1} {j + 1}"
            params.append(Parameter(d))
    b = Parameter_Block(params, {"name_": "StellarActivityBlock", "type_": "StellarActivity",
                                 "class_": "Data", "number_": nins})
    b.cornums = [int(c) for c in cornums]
    return b


def SinusoidBlock(number):
    """emp/block_repo.py:75-100: A cos(2 pi t / P + phase)."""
    suffix = f" {number}" if number else ""
    params = [_named(k, suffix) for k in ("dPeriod", "dAmplitude", "dPhase")]
    return Parameter_Block(params, {"name_": "SinusoidBlock", "type_": "Sinusoid", "class_": "Model",
                                    "number_": number})


def MagneticCycleBlock(number):
    """emp/block_repo.py:275-314: A1 cos(pi t / P + phase1) + A2 cos(2 pi t / P + phase2),
    parameters (Period S1, Amplitude S1, Amplitude S2, Phase S1, Phase S2)."""
    keys = ("dPeriod", "dAmplitude", "dAmplitude", "dPhase", "dPhase")
    params = [_named(k, f" S{s}") for k, s in zip(keys, (1, 1, 2, 1, 2))]
    return Parameter_Block(params, {"name_": "MagneticCycleBlock", "type_": "MagneticCycle",
                                    "class_": "Model", "number_": number})


class SmartSetter:
    """Data-driven default limits and priors (emp/block_repo.py:505-867).

    ``data`` is a dict with arrays ``BJD`` (already relative to common_t),
    ``RV`` and ``Flag``.
    """

    uni, norm = "Uniform", "Normal"
    angular = [0, 2 * np.pi]

    def __init__(self, data):
        self.data = data

    def _std(self):
        return float(np.std(self.data["RV"]))            # pandas std(ddof=0)

    def _span(self):
        return float(self.data["BJD"].max() - self.data["BJD"].min())

    # block_repo.py:517-630
    def set_Keplerian(self, b, **kw):
        per_lim = self._span()
        amp_lim = self._std() * np.sqrt(4)
        ecc_limits, ecc_prargs = kw.get("prargs", ([0, 1], [0, 0.1]))
        p = b.parameterisation
        U, N = self.uni, self.norm
        if p == 0:
            lims = [[1.5, per_lim], [1e-6, amp_lim], self.angular, ecc_limits, self.angular]
            pri, pra = [U, U, U, N, U], [None, None, None, ecc_prargs, None]
        elif p == 1:
            lims = [[1.5, per_lim], [1e-6, amp_lim], self.angular, [-1, 1], [-1, 1]]
            pri, pra = [U] * 5, [None] * 5
            self._ecc_additional(b)
        elif p == 2:
            kamp = np.sqrt(amp_lim / 2)
            lims = [[np.log(1.5), np.log(per_lim)], [-kamp, kamp], [-kamp, kamp], [-1, 1], [-1, 1]]
            pri, pra = [U] * 5, [None] * 5
            self._ecc_additional(b)
        elif p == 3:
            lims = [[1.5, per_lim], [1e-6, amp_lim], [-1000, 1000], ecc_limits, self.angular]
            pri, pra = [U, U, U, N, U], [None, None, None, ecc_prargs, None]
        elif p == 4:
            lims = [[1.5, per_lim], [1e-6, amp_lim], [-1000, 1000], [-1, 1], [-1, 1]]
            pri, pra = [U] * 5, [None] * 5
            self._ecc_additional(b)
        elif p == 6:
            lims = [[np.log(1.5), np.log(per_lim)], [1e-6, amp_lim], self.angular, ecc_limits,
                    self.angular]
            pri, pra = ["Jeffreys", U, U, N, U], [None, None, None, ecc_prargs, None]
        elif p == 7:
            lims = [[np.log(1.5), np.log(per_lim)], [1e-6, amp_lim], self.angular, [-1, 1], [-1, 1]]
            pri, pra = ["Jeffreys", U, U, U, U], [None] * 5
            self._ecc_additional(b, prargs=ecc_prargs)
        else:
            raise ValueError(p)
        return [list(map(float, l)) for l in lims], pri, pra

    def _ecc_additional(self, b, prargs=None):
        """ecc_w_add_additional, block_repo.py:662-683: a prior on e = S^2 + C^2."""
        e = Parameter(make_parameter("dEccentricity"))
        e.has_prior = True
        e.limits = [0, 1]
        if prargs is not None:
            e.prior = self.norm
            e.prargs = list(prargs)
        e.kind = "Ecc"
        b.add_additional_parameters(e)

    def _ins_absmax(self, nin):
        m = self.data["Flag"] == nin + 1
        return float(np.abs(self.data["RV"][m]).max())

    # block_repo.py:698-707
    def set_Offset(self, b, **kw):
        lims = [[-self._ins_absmax(n), self._ins_absmax(n)] for n in range(b.number_)]
        return lims, [self.uni] * b.number_, [None] * b.number_

    # block_repo.py:710-722
    def set_Jitter(self, b, jit_limits=None, jit_prargs=(5, 5), **kw):
        lims = [[1e-5, self._ins_absmax(n)] for n in range(b.number_)]
        return lims, [self.norm] * b.number_, [list(jit_prargs)] * b.number_

    # block_repo.py:725-739
    def set_MOAV(self, b, **kw):
        lims, pri, pra = [], [], []
        reps = len(b) // (2 * max(b.moav, 1)) if b.moav else 0
        for _ in range(reps):
            for _ in range(b.moav):
                lims += [[0.0, 0.3], [5, 25]]
                pri += [self.uni, self.uni]
                pra += [None, None]
        return lims, pri, pra

    # block_repo.py:815-826
    def set_Acceleration(self, b, **kw):
        return [[-1, 1]] * b.number_, [self.uni] * b.number_, [None] * b.number_

    # block_repo.py:742-749
    def set_StellarActivity(self, b, **kw):
        n = sum(b.cornums)
        return [[-1.0, 1.0]] * n, [self.uni] * n, [None] * n

    # block_repo.py:805-830
    def set_Sinusoid(self, b, **kw):
        lims = [[1.5, self._span()], [1e-6, self._std() * np.sqrt(4)], self.angular]
        return [list(map(float, l)) for l in lims], [self.uni] * 3, [None] * 3

    def set_MagneticCycle(self, b, **kw):
        amp = [1e-6, self._std() * np.sqrt(4)]
        lims = [[1.5, self._span()], amp, amp, self.angular, self.angular]
        return [list(map(float, l)) for l in lims], [self.uni] * 5, [None] * 5

    def __call__(self, block, *args, **kw):
        lims, pri, pra = getattr(self, f"set_{block.type_}")(block, *args, **kw)
        block.set_attr("limits", lims)
        block.set_attr("prior", pri)
        block.set_attr("prargs", [copy.deepcopy(x) for x in pra])
