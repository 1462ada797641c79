"""ctypes binding of the C ABI declared in include/dualres_hip.h.

The shared library is built in-tree (``dualres_amd/libdualres_hip.so``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdualres_hip.so")

DR_OK, DR_EDOMAIN, DR_ELOGIC, DR_ERUNTIME, DR_EHIP, DR_ENOMEM, DR_EBUSY, DR_ECANCELED = range(8)
DR_FP64, DR_FP32 = 0, 1
DR_STATE_COMPAT, DR_STATE_REAL = 0, 1
DR_COV_RBF, DR_COV_RQ = 0, 1
DR_BATCH_GRAPH, DR_BATCH_STREAMS, DR_BATCH_EAGER = 0, 1, 2


class DrError(RuntimeError):
    """Error raised by an engine call; ``code`` is the dr_status."""

    def __init__(self, code, msg):
        super().__init__(f"[dr_status {code}] {msg}")
        self.code = code


class DrDomainError(DrError, ValueError):
    pass


class dr_config(C.Structure):
    _fields_ = [
        ("n_datasets", C.c_int),
        ("precision", C.c_int),
        ("state_mode", C.c_int),
        ("covfun", C.c_int),
        ("embed", C.c_int * 3),
        ("theta", C.c_double * 3),
        ("qform", C.c_float * 16),
        ("seed", C.c_uint64),
        ("device", C.c_int),
    ]


class dr_update_result(C.Structure):
    _fields_ = [
        ("mh_rate", C.c_double), ("R", C.c_double), ("H0", C.c_double), ("H1", C.c_double),
        ("log_prior0", C.c_double), ("log_lik0", C.c_double), ("e0", C.c_double),
        ("log_prior1", C.c_double), ("log_lik1", C.c_double), ("k1", C.c_double),
        ("sse_h", C.c_double), ("sse_s", C.c_double), ("accepted", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_F = C.POINTER(C.c_float)

# name -> argtypes (restype is always int = dr_status)
SIGNATURES = {
    "dr_version": [],
    "dr_device_count": [C.POINTER(C.c_int)],
    "dr_create": [C.POINTER(dr_config), _I32, C.c_int64, C.POINTER(_P)],
    "dr_destroy": [_P],
    "dr_spectrum_report": [_P, C.POINTER(C.c_int), _I64, _I64, _D],
    "dr_get_lambda": [_P, _D],
    "dr_set_data": [_P, _D, C.c_int64, _D, C.c_int64, _I64, _I32, _D],
    "dr_add_data_to_mu": [_P],
    "dr_w_layout": [_P, C.c_int],
    "dr_w_stats": [_P, _I64],
    "dr_set_output_indices": [_P, _I32, C.c_int64],
    "dr_get_mu_output": [_P, _D, C.c_int64],
    "dr_estimate_bytes": [C.POINTER(dr_config), _I32, C.c_int64, C.c_int64, C.c_int64, _I64],
    "dr_summary_reset": [_P],
    "dr_summary_update": [_P],
    "dr_summary_read": [_P, _D, _D, _D, C.c_int64, _I64],
    "dr_get_mu_data": [_P, _D, C.c_int64],
    "dr_set_sigma_sq_inv": [_P, _D],
    "dr_get_sigma_sq_inv": [_P, _D],
    "dr_set_tau_sq_inv": [_P, C.c_double],
    "dr_get_tau_sq_inv": [_P, _D],
    "dr_mu_data_sumsq": [_P, _D],
    "dr_update": [_P, C.c_double, C.c_int, C.c_double, C.POINTER(dr_update_result)],
    "dr_update_with_noise": [_P, C.c_double, C.c_int, C.c_double, _D, _D, C.POINTER(dr_update_result)],
    "dr_autotune": [_P, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "dr_set_tuning": [_P, C.POINTER(C.c_int)],
    "dr_get_tuning": [_P, C.POINTER(C.c_int)],
    "dr_update_captured": [_P, C.c_double, C.c_int, C.c_double, C.POINTER(dr_update_result)],
    "dr_batch_create": [C.POINTER(_P), C.c_int, C.c_int, C.POINTER(_P)],
    "dr_batch_update": [_P, _D, C.POINTER(C.c_int), _D, C.POINTER(dr_update_result)],
    "dr_batch_destroy": [_P],
    "dr_log_prior": [_P, _D],
    "dr_log_likelihood": [_P, _D],
    "dr_log_posterior": [_P, _D],
    "dr_sse": [_P, _D, _D],
    "dr_set_state": [_P, _D, _D, _D, _D],
    "dr_get_state": [_P, _D, _D, _D, _D],
    "dr_gradient": [_P, _D, _D],
    "dr_leapfrog": [_P, C.c_double, C.c_int, C.POINTER(dr_update_result)],
    "dr_profile_enable": [_P, C.c_int],
    "dr_profile_read": [_P, _D, _I64, _D, _D],
    "dr_profile_kernels": [_P, _D, C.c_int, C.POINTER(C.c_int), _I64],
    "dr_kriging_build": [_D, C.POINTER(C.c_int), _F, _D, C.POINTER(C.c_int), _F, _D, C.c_double, C.c_int,
                         C.POINTER(_P)],
    "dr_csr_info": [_P, _I64, _I64, _I64, _I64],
    "dr_csr_solver_counts": [_P, _I64, _I64],
    "dr_csr_copy": [_P, _I64, _I32, _D],
    "dr_csr_free": [_P],
    "dr_mce_offsets": [_I64, _I32],
    "dr_mce_summary": [_D, C.POINTER(C.c_int), _F, C.c_int, _F, _F, _I32],
    "dr_mce_summary_noalloc": [C.c_void_p, C.POINTER(C.c_int), _F, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p],
    "dr_mce_binned": [_D, C.POINTER(C.c_int), _F, C.c_int, C.c_int, C.c_double, _D, _D, _I64],
}

_lib = None


def load(path=None):
    """Load the engine library (cached).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"dualres HIP engine not built: {p} missing (run __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = C.c_int
    lib.dr_last_error.argtypes = []
    lib.dr_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def check(status):
    if status != DR_OK:
        msg = load().dr_last_error().decode(errors="replace")
        if status == DR_EDOMAIN:
            raise DrDomainError(status, msg)
        raise DrError(status, msg)


def dptr(a):
    return a.ctypes.data_as(_D) if a is not None else None


def iptr32(a):
    return a.ctypes.data_as(_I32) if a is not None else None


def iptr64(a):
    return a.ctypes.data_as(_I64) if a is not None else None


def fptr(a):
    return a.ctypes.data_as(_F) if a is not None else None


def as_f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)
