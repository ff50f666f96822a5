"""The RV model: parameter layout, priors and the device likelihood handle.

Host mirror of ``ReddModel`` (emp/emp_model.py:200-433, 706-781) and of the
prior writer (emp/emp.py:182-254).  Instead of writing a Python script, the
model is flattened into an ``EmpModelDesc`` and handed to the HIP library,
which evaluates ``my_model``/``my_likelihood``/``my_prior`` on the device.

Block order follows the reference (emp/emp.py:1235, 2628-2650):
``[Keplerian 1..k, Acceleration, Offset, Jitter, MOAV, StellarActivity, Sinusoid,
MagneticCycle]``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .blocks import SC_PARAMETERISATIONS
from .data import RVData

_ORDER = {"Keplerian": 0, "Acceleration": 1, "Offset": 2, "Jitter": 3, "MOAV": 4, "StellarActivity": 5,
          "Sinusoid": 6, "MagneticCycle": 7}


def _as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class DeviceModel:
    """Owns one ``EmpModel*`` handle."""

    def __init__(self, desc, keep):
        self._keep = keep
        h = C.c_void_p()
        _abi.check(_abi.lib().emp_model_create(C.byref(desc), C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            _abi.lib().emp_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ReddModel:
    """Flattened RV model over ``blocks`` (list of Parameter_Block) and ``data``.

    semantics: 'reference' reproduces the reference's arithmetic exactly,
    including the per-instrument moving average being a no-op
    (support/models/moav00.model:18-19); 'intended' applies moav01's
    recurrence within each instrument.
    """

    def __init__(self, blocks, data: RVData, semantics="reference"):
        if semantics not in _abi.SEM:
            raise ValueError(f"semantics must be one of {list(_abi.SEM)}")
        self.blocks = sorted(blocks, key=lambda b: _ORDER[b.type_])
        for i in range(1, len(self.blocks)):
            if self.blocks[i].type_ == self.blocks[i - 1].type_ and self.blocks[i].type_ != "Keplerian":
                raise ValueError(f"duplicate {self.blocks[i].type_} block")
        self.data = data
        self.semantics = semantics
        self._device = None
        self.refresh__()

    # -- layout (emp_model.py:223-330) ------------------------------------
    def refresh__(self):
        blen = 0
        self.kplan__ = 0
        self.nins__ = self.data.nins
        self.parameterisation = None
        self.use_c = False
        for i, b in enumerate(self.blocks):
            b.refresh_block()
            b.check_prargs()
            b.bnumber_ = i
            b.slice = slice(blen, blen + len(b))
            blen += len(b)
            if b.type_ == "Keplerian":
                if self.parameterisation is None:
                    self.parameterisation = b.parameterisation
                    self.use_c = b.use_c
                elif b.parameterisation != self.parameterisation or b.use_c != self.use_c:
                    raise ValueError("all Keplerian blocks must share one parameterisation")
                self.kplan__ += 1
            elif b.type_ in ("Offset", "Jitter") and len(b) != self.nins__:
                raise ValueError(f"{b.type_} block has {len(b)} parameters for {self.nins__} instruments")
            elif b.type_ == "StellarActivity" and list(b.cornums) != list(self.data.nsai):
                raise ValueError(f"StellarActivity block columns {list(b.cornums)} differ from the data's "
                                 f"{list(self.data.nsai)}")
        self.ndim_full = blen
        params = [p for b in self.blocks for p in b]
        self.mod_fixed = [p.fixed for p in params]
        self.C_ = np.array([i for i, f in enumerate(self.mod_fixed) if f is None], dtype=int)
        self.A_ = np.array([i for i, f in enumerate(self.mod_fixed) if f is not None], dtype=int)
        self.ndim__ = len(self.C_)
        self._device = None

    def params(self):
        return [p for b in self.blocks for p in b]

    def block(self, typ):
        for b in self.blocks:
            if b.type_ == typ:
                return b
        return None

    # -- plain-dict description (consumed by tests and the oracle) ---------
    def spec(self):
        blocks = []
        for b in self.blocks:
            d = {"type": b.type_, "start": b.slice.start, "len": len(b)}
            if b.type_ == "Keplerian":
                d["parameterisation"] = b.parameterisation
                d["use_c"] = bool(b.use_c)
                d["additional"] = [{"kind": "Ecc", "prior": p.prior, "limits": list(p.limits),
                                    "prargs": p.prargs}
                                   for p in b.additional_parameters
                                   if getattr(p, "has_prior", False) and getattr(p, "kind", "") == "Ecc"]
            if b.type_ == "MOAV":
                d["order"] = b.moav
                d["global"] = bool(b.is_global)
            if b.type_ == "StellarActivity":
                d["cornums"] = list(b.cornums)
            blocks.append(d)
        params = [{"name": p.name, "prior": p.prior, "limits": list(p.limits), "prargs": p.prargs,
                   "fixed": p.fixed} for p in self.params()]
        return {"blocks": blocks, "params": params, "nins": self.nins__, "semantics": self.semantics}

    # -- flat descriptor for the HIP library ------------------------------
    @staticmethod
    def _prior_term(kind_name, limits, prargs, block):
        t = _abi.EmpPriorTerm()
        t.block = block
        if kind_name not in _abi.PRIOR_KIND:
            raise NotImplementedError(f"prior {kind_name!r} is not on the GPU path "
                                      f"(supported: {list(_abi.PRIOR_KIND)})")
        t.kind = _abi.PRIOR_KIND[kind_name]
        lo, hi = (np.nan, np.nan) if limits is None else limits
        t.lo, t.hi = float(lo), float(hi)
        if kind_name == "Uniform":
            t.a0 = float(prargs)
        elif kind_name == "Normal":
            mu, s, logZ = prargs
            t.a0, t.a1, t.a2 = float(mu), float(s), float(logZ)
            t.a3 = float(np.log(s * np.sqrt(2 * np.pi)))      # Normal.prior:8, evaluated as numpy does
        elif kind_name == "Jeffreys":
            t.a0 = float(np.log(1 / (hi - lo)))                 # Jeffreys.prior:5
        elif kind_name == "Beta":
            from scipy.special import betaln
            a, bb = prargs[0], prargs[1]
            t.a0, t.a1, t.a2 = float(a), float(bb), float(betaln(a, bb))
        return t

    def descriptor(self):
        d = self.data
        desc = _abi.EmpModelDesc()
        desc.abi_version = _abi.EMP_ABI_VERSION
        desc.n_epochs = d.ndat
        desc.n_ins = self.nins__
        desc.n_planets = self.kplan__
        desc.parameterisation = self.parameterisation if self.parameterisation is not None else 1
        desc.use_c = int(bool(self.use_c))
        acc = self.block("Acceleration")
        jit = self.block("Jitter")
        ma = self.block("MOAV")
        off = self.block("Offset")
        if off is None:
            raise ValueError("the RV model needs an Offset block")
        desc.acc_order = len(acc) if acc is not None else 0
        desc.has_jitter = int(jit is not None)
        desc.ma_order = ma.moav if ma is not None else 0
        desc.ma_global = int(bool(ma.is_global)) if ma is not None else 0
        desc.semantics = _abi.SEM[self.semantics]
        desc.ndim_full = self.ndim_full
        desc.ndim = self.ndim__
        desc.n_blocks = len(self.blocks)
        desc.off_kep = 0
        desc.off_acc = acc.slice.start if acc is not None else 0
        desc.off_offset = off.slice.start
        desc.off_jitter = jit.slice.start if jit is not None else 0
        desc.off_ma = ma.slice.start if ma is not None else 0
        sn, mc = self.block("Sinusoid"), self.block("MagneticCycle")
        desc.off_sinusoid = sn.slice.start if sn is not None else -1
        desc.off_magnetic = mc.slice.start if mc is not None else -1
        sa = self.block("StellarActivity")
        desc.off_sai = sa.slice.start if sa is not None and len(sa) else -1
        desc.n_sai = len(sa) if sa is not None else 0
        t, y, e = _as_f64(d.X), _as_f64(d.Y), _as_f64(d.YERR)
        ins = np.ascontiguousarray(d.FLAG - 1, dtype=np.int32)
        free = np.ascontiguousarray(self.C_, dtype=np.int32)
        fixed = _as_f64([np.nan if f is None else f for f in self.mod_fixed])
        terms = (_abi.EmpPriorTerm * self.ndim_full)()
        j = 0
        for bi, b in enumerate(self.blocks):
            for p in b:
                if p.fixed is not None:
                    terms[j] = self._prior_term("Fixed", None, None, bi)
                else:
                    terms[j] = self._prior_term(p.prior, p.limits, p.prargs, bi)
                j += 1
        ecc = (_abi.EmpPriorTerm * max(self.kplan__, 1))()
        for k in range(max(self.kplan__, 1)):
            ecc[k].kind = _abi.EMP_PRIOR_NONE
        for k, b in enumerate(self.blocks[:self.kplan__]):
            for p in b.additional_parameters:
                if getattr(p, "has_prior", False) and getattr(p, "kind", "") == "Ecc":
                    ecc[k] = self._prior_term(p.prior, p.limits, p.prargs, k)
        P = C.POINTER
        desc.t = t.ctypes.data_as(P(C.c_double))
        desc.y = y.ctypes.data_as(P(C.c_double))
        desc.yerr = e.ctypes.data_as(P(C.c_double))
        desc.ins = ins.ctypes.data_as(P(C.c_int32))
        desc.free_index = free.ctypes.data_as(P(C.c_int32))
        desc.fixed_value = fixed.ctypes.data_as(P(C.c_double))
        desc.prior = terms
        desc.ecc_prior = ecc
        sai = _as_f64(d.SA)
        counts = np.ascontiguousarray(d.nsai, dtype=np.int32)
        if desc.n_sai:
            desc.sai = sai.ctypes.data_as(P(C.c_double))
            desc.sai_counts = counts.ctypes.data_as(P(C.c_int32))
        return desc, (t, y, e, ins, free, fixed, terms, ecc, sai, counts)

    @property
    def device(self):
        if self._device is None:
            desc, keep = self.descriptor()
            self._device = DeviceModel(desc, keep)
        return self._device

    # -- batched evaluation ----------------------------------------------
    def _batch(self, fn, theta, precision=None, grad=False):
        """theta: numpy [B, D] (host) or a torch CUDA tensor [B, D] (device)."""
        L = _abi.lib()
        grad_flags = _abi.EMP_GRAD_MIXED if (grad and self._prec(precision) == _abi.EMP_MIXED) else 0
        h = self.device.handle
        is_t = type(theta).__module__.startswith("torch")
        if is_t:
            import torch
            if not theta.is_cuda or theta.dtype != torch.float64 or theta.dim() != 2:
                raise ValueError("device theta must be a 2-D float64 CUDA tensor")
            th = theta.contiguous()
            B, ld = th.shape
            out = torch.empty(B, dtype=torch.float64, device=th.device)
            stream = C.c_void_p(torch.cuda.current_stream(th.device).cuda_stream)
            args = [h, C.c_void_p(th.data_ptr()), B, ld, C.c_void_p(out.data_ptr())]
            if grad:
                g = torch.empty(B, self.ndim__, dtype=torch.float64, device=th.device)
                _abi.check(fn(*args, C.c_void_p(g.data_ptr()), grad_flags, stream))
                return out, g
            if precision is not None:
                args.append(precision)
            _abi.check(fn(*args, 0, stream))
            return out
        th = np.ascontiguousarray(np.atleast_2d(np.asarray(theta, dtype=np.float64)))
        B, ld = th.shape
        out = np.empty(B)
        args = [h, th.ctypes.data, B, ld, out.ctypes.data]
        if grad:
            g = np.empty((B, self.ndim__))
            _abi.check(fn(*args, g.ctypes.data, _abi.EMP_HOST_POINTERS | grad_flags, None))
            return out, g
        if precision is not None:
            args.append(precision)
        _abi.check(fn(*args, _abi.EMP_HOST_POINTERS, None))
        return out

    @staticmethod
    def _prec(precision):
        if precision in ("fp64", _abi.EMP_FP64, None):
            return _abi.EMP_FP64
        if precision in ("mixed", _abi.EMP_MIXED):
            return _abi.EMP_MIXED
        raise ValueError("precision must be 'fp64' or 'mixed'")

    def lnl_batch(self, theta, precision="fp64"):
        return self._batch(_abi.lib().emp_lnl_batch, theta, self._prec(precision))

    def lnprior_batch(self, theta):
        return self._batch(_abi.lib().emp_lnprior_batch, theta)

    def lnl_grad_batch(self, theta, precision="fp64"):
        """lnL and d lnL / d theta (adjoint kernel); precision "mixed" uses the
        mixed path's Kepler solve."""
        return self._batch(_abi.lib().emp_lnl_grad_batch, theta, precision=precision, grad=True)

    def set_waves_per_theta(self, w):
        _abi.check(_abi.lib().emp_model_set_waves_per_theta(self.device.handle, int(w)))

    def set_grad_kernel(self, which="auto"):
        """Adjoint kernel for lnl_grad_batch and the gradient move: 'auto'
        (single-pass for MA order <= 1 and >= 1024 epochs), 'two_pass' or
        'single_pass' (MA order <= 1 only)."""
        code = {"auto": _abi.EMP_GRAD_KERNEL_AUTO, "two_pass": _abi.EMP_GRAD_KERNEL_TWO_PASS,
                "single_pass": _abi.EMP_GRAD_KERNEL_SINGLE_PASS}[which]
        _abi.check(_abi.lib().emp_model_set_grad_kernel(self.device.handle, code))

    def nonfinite_count(self):
        v = C.c_int64()
        _abi.check(_abi.lib().emp_model_nonfinite_count(self.device.handle, C.byref(v)))
        return v.value

    # -- the generated script's per-walker callables (compat, B = 1) -------
    def my_model(self, theta):
        """emp/emp_model.py:706-781 -> (model0, err20) on the data's epochs."""
        th = _as_f64(theta)
        if th.shape != (self.ndim__,):
            raise ValueError(f"theta must have {self.ndim__} free parameters")
        m = np.empty(self.data.ndat)
        e = np.empty(self.data.ndat)
        _abi.check(_abi.lib().emp_model_eval(self.device.handle, th.ctypes.data, m.ctypes.data, e.ctypes.data))
        return m, e

    def my_likelihood(self, theta):
        """support/likelihoods/00.like:3-5 (non-finite values come back as -inf)."""
        return float(self.lnl_batch(np.asarray(theta, dtype=np.float64)[None, :])[0])

    def my_prior(self, theta):
        """emp/emp.py:182-254."""
        return float(self.lnprior_batch(np.asarray(theta, dtype=np.float64)[None, :])[0])

    def callables(self):
        """(my_model, my_likelihood, my_prior) tagged with this model, as the
        generated script would define them."""
        mm, ml, mp = self.my_model, self.my_likelihood, self.my_prior
        return mm, ml, mp

    # -- helpers ------------------------------------------------------------
    def limits_free(self):
        return np.array([p.limits for p in self.params() if p.fixed is None], dtype=np.float64)

    def names_free(self):
        return [p.name for p in self.params() if p.fixed is None]
