"""ctypes binding of include/emp_abi.h (libemp_mi355x.so).

The library is the only compute path: when it is missing or the device is
absent every compute call raises ``EmpError``.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemp_mi355x.so")

EMP_ABI_VERSION = 3
EMP_HOST_POINTERS = 1
EMP_FP64, EMP_MIXED = 0, 1
PRIOR_KIND = {"Uniform": 0, "Normal": 1, "Jeffreys": 2, "Beta": 3, "Fixed": 4}
EMP_PRIOR_NONE = -1
SEM = {"reference": 0, "intended": 1}
PT_WHAT = {"chain": 0, "lnl": 1, "lnp": 2, "betas": 3, "tsw": 4, "smd": 5, "accepted": 6,
           "x": 7, "state_lnl": 8, "state_lnp": 9, "state_betas": 10, "nsweeps": 11,
           "grad_steps": 12}
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EUNSUP", -5: "EIO", -6: "EFAULT_INJECTED"}


class EmpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class EmpPriorTerm(C.Structure):
    _fields_ = [("kind", C.c_int32), ("block", C.c_int32), ("lo", C.c_double), ("hi", C.c_double),
                ("a0", C.c_double), ("a1", C.c_double), ("a2", C.c_double), ("a3", C.c_double)]


class EmpModelDesc(C.Structure):
    _fields_ = [("abi_version", C.c_int32), ("n_epochs", C.c_int32), ("n_ins", C.c_int32),
                ("n_planets", C.c_int32), ("parameterisation", C.c_int32), ("use_c", C.c_int32),
                ("acc_order", C.c_int32), ("has_jitter", C.c_int32), ("ma_order", C.c_int32),
                ("ma_global", C.c_int32), ("semantics", C.c_int32), ("ndim_full", C.c_int32),
                ("ndim", C.c_int32), ("n_blocks", C.c_int32),
                ("off_kep", C.c_int32), ("off_acc", C.c_int32), ("off_offset", C.c_int32),
                ("off_jitter", C.c_int32), ("off_ma", C.c_int32),
                ("off_sinusoid", C.c_int32), ("off_magnetic", C.c_int32),
                ("off_sai", C.c_int32), ("n_sai", C.c_int32),
                ("t", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
                ("yerr", C.POINTER(C.c_double)), ("ins", C.POINTER(C.c_int32)),
                ("free_index", C.POINTER(C.c_int32)), ("fixed_value", C.POINTER(C.c_double)),
                ("prior", C.POINTER(EmpPriorTerm)), ("ecc_prior", C.POINTER(EmpPriorTerm)),
                ("sai", C.POINTER(C.c_double)), ("sai_counts", C.POINTER(C.c_int32))]


class EmpMoveDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("weight", C.c_double), ("scale", C.c_double)]


EMP_MOVE_STRETCH, EMP_MOVE_GRADIENT = 0, 1
EMP_GRAD_MIXED = 2
EMP_NOALLOC_SCRATCH_DOUBLES = 1 << 23     # doubles of the allocation-free entry point's static scratch
EMP_GRAD_KERNEL_AUTO, EMP_GRAD_KERNEL_TWO_PASS, EMP_GRAD_KERNEL_SINGLE_PASS = 0, 1, 2
EMP_SWAP_AUTO, EMP_SWAP_LDS, EMP_SWAP_STREAM, EMP_SWAP_GLOBAL = 0, 1, 2, 3


class EmpPTDesc(C.Structure):
    _fields_ = [("ntemps", C.c_int32), ("nwalkers", C.c_int32), ("nsteps", C.c_int32),
                ("precision", C.c_int32), ("a", C.c_double), ("seed", C.c_uint64),
                ("adapt_mode", C.c_int32), ("record", C.c_int32), ("adapt_tau", C.c_double),
                ("adapt_nu", C.c_double), ("max_sweeps", C.c_int64)]


class EmpSystemsDesc(C.Structure):
    _fields_ = [("n_systems", C.c_int32), ("ntemps", C.c_int32), ("nwalkers", C.c_int32),
                ("nsteps", C.c_int32), ("precision", C.c_int32), ("adapt_mode", C.c_int32),
                ("a", C.c_double), ("seed", C.c_uint64), ("adapt_tau", C.c_double),
                ("adapt_nu", C.c_double), ("record_every", C.c_int32), ("record_temps", C.c_int32),
                ("max_records", C.c_int64)]


class EmpSystemsTuning(C.Structure):
    _fields_ = [("waves_per_wg", C.c_int32), ("waves_per_theta", C.c_int32), ("wgs_per_cu", C.c_int32),
                ("order", C.c_int32), ("ms_per_sweep", C.c_double)]


_P = C.c_void_p
_SIGS = {
    "emp_model_create": (C.c_int, [C.POINTER(EmpModelDesc), C.POINTER(_P)]),
    "emp_model_destroy": (None, [_P]),
    "emp_lnl_batch": (C.c_int, [_P, _P, C.c_int64, C.c_int64, _P, C.c_int32, C.c_int32, _P]),
    "emp_lnprior_batch": (C.c_int, [_P, _P, C.c_int64, C.c_int64, _P, C.c_int32, _P]),
    "emp_model_eval": (C.c_int, [_P, _P, _P, _P]),
    "emp_lnl_grad_batch": (C.c_int, [_P, _P, C.c_int64, C.c_int64, _P, _P, C.c_int32, _P]),
    "emp_lnl_grad_batch_noalloc": (C.c_int, [_P, _P, C.c_int64, C.c_int64, _P, _P, C.c_int32, _P]),
    "emp_model_set_waves_per_theta": (C.c_int, [_P, C.c_int32]),
    "emp_model_set_grad_kernel": (C.c_int, [_P, C.c_int32]),
    "emp_model_get_waves_per_theta": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "emp_model_nonfinite_count": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "emp_pt_create": (C.c_int, [_P, C.POINTER(EmpPTDesc), _P, _P, C.POINTER(_P)]),
    "emp_pt_destroy": (None, [_P]),
    "emp_pt_run": (C.c_int, [_P, C.c_int64, C.c_int32]),
    "emp_pt_get": (C.c_int, [_P, C.c_int32, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "emp_pt_set_state": (C.c_int, [_P, _P, _P]),
    "emp_pt_set_adaptation": (C.c_int, [_P, C.c_int32, C.c_double, C.c_double]),
    "emp_pt_set_swap_path": (C.c_int, [_P, C.c_int32]),
    "emp_pt_set_moves": (C.c_int, [_P, _P, C.c_int32]),
    "emp_pt_set_gradient_steps": (C.c_int, [_P, _P]),
    "emp_pt_lnl_moments": (C.c_int, [_P, C.c_int64, _P, _P, _P]),
    "emp_pt_checkpoint": (C.c_int, [_P, C.c_char_p]),
    "emp_pt_restore": (C.c_int, [_P, C.c_char_p]),
    "emp_pt_last_timing": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "emp_pt_time_halfstep": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    "emp_pt_time_gradstep": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    "emp_pt_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "emp_systems_create": (C.c_int, [_P, C.POINTER(EmpSystemsDesc), _P, _P, C.POINTER(_P)]),
    "emp_systems_destroy": (None, [_P]),
    "emp_systems_run": (C.c_int, [_P, C.c_int64, C.c_int32]),
    "emp_systems_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(EmpSystemsTuning)]),
    "emp_systems_set_tuning": (C.c_int, [_P, C.POINTER(EmpSystemsTuning)]),
    "emp_systems_get_tuning": (C.c_int, [_P, C.POINTER(EmpSystemsTuning)]),
    "emp_systems_get": (C.c_int, [_P, C.c_int32, C.c_int32, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "emp_systems_last_timing": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "emp_philox4x32_10": (C.c_int, [_P, _P, _P]),
    "emp_last_error": (C.c_char_p, []),
    "emp_abi_version": (C.c_int, []),
    "emp_device_info": (C.c_int, [C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_char_p, C.c_int32]),
}

_lib = None


def _preload_torch_hip():
    """PyTorch-ROCm ships its own HIP runtime under the same soname
    (libamdhip64.so.7).  Load that copy first, so that this library binds to
    it: a process must hold one HIP runtime, and tensors and streams pass
    between torch and this library."""
    try:
        import importlib.util
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.submodule_search_locations:
        return
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(path):
        C.CDLL(path, mode=C.RTLD_GLOBAL)


def lib():
    """Load the HIP library (raises EmpError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EmpError(-4, f"{LIB_PATH} is not built; run astroemperor_amd.build.build()")
        _preload_torch_hip()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        if L.emp_abi_version() != EMP_ABI_VERSION:
            raise EmpError(-1, "library ABI version mismatch")
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().emp_last_error()
        raise EmpError(rc, msg.decode() if msg else "")
    return rc


def device_info():
    n = C.c_int32()
    clk = C.c_int32()
    name = C.create_string_buffer(64)
    check(lib().emp_device_info(C.byref(n), C.byref(clk), name, 64))
    return {"n_cu": n.value, "clock_khz": clk.value, "arch": name.value.decode()}
