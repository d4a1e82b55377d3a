"""ctypes binding of libpmcmc_amd.so (the C ABI in include/pmcmc_amd.h).

The library is the only compute path: if it is missing, importing the
compute entry points raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("PMCMC_AMD_LIB", os.path.join(_HERE, "libpmcmc_amd.so"))

PMCMC_OK = 0
PMCMC_ERR_ARGUMENT = 1
PMCMC_ERR_DIMENSION = 2
PMCMC_ERR_HIP = 3
PMCMC_ERR_ALLOC = 4
PMCMC_ERR_DEVICE = 5
PMCMC_ERR_UNSUPPORTED = 6

PREC_F64 = 0
PREC_F32 = 1
PREC_BF16_JAC = 2

KT_GLM_PASS = 0
KT_TILE_REDUCE = 1
KT_TILE_APPLY = 2
KT_GLM_REDUCE = 3
KT_SCAN_SPINE = 4
KT_ITERATION = 5
KT_LAMBDA_PASS = 6
KT_NAMES = ("glm_pass", "tile_reduce", "tile_apply", "glm_reduce", "scan_spine", "iteration", "lambda_pass")

JAC_STOCH_DIAG = 0
JAC_DIAG = 1
JAC_FULL = 2  # dense decision-frozen Jacobian; Gaussian targets with D <= 64
JAC_FULL_SIGMOID = 3  # dense surrogate (sigmoid) Jacobian


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


class PmcmcError(RuntimeError):
    pass


class DeerParams(C.Structure):
    _fields_ = [
        ("eps", C.c_double),
        ("tol_abs", C.c_double),
        ("tol_rel", C.c_double),
        ("damping", C.c_double),
        ("maxiter", C.c_int32),
        ("jacobian", C.c_int32),
        ("probes", C.c_int32),
        ("precision", C.c_int32),
        ("seed", C.c_uint64),
        ("fixed_probes", C.c_int32),
    ]


class DeerInfo(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iters", C.c_int32), ("metric", C.c_double)]


_P = C.c_void_p
_I64 = C.c_int64
class TileShape(C.Structure):
    _fields_ = [("tc", C.c_int32), ("parts", C.c_int32), ("single_pass", C.c_int32), ("pad", C.c_int32),
                ("us", C.c_double)]


TUNE_SCAN_DIAG = 0
TUNE_DEER = 1


class ChainDesc(C.Structure):
    """pmcmc_chain_desc (include/pmcmc_amd.h)."""

    _fields_ = [
        ("target", C.c_void_p),
        ("T", C.c_int64),
        ("eps", C.c_double),
        ("x0", C.c_void_p),
        ("Xi", C.c_void_p),
        ("U", C.c_void_p),
        ("S_out", C.c_void_p),
        ("accepted_out", C.c_void_p),
    ]


_SIGS = {
    "pmcmc_version": ([], C.c_char_p),
    "pmcmc_ctx_create": ([C.c_int, _P, C.POINTER(_P)], C.c_int),
    "pmcmc_ctx_destroy": ([_P], C.c_int),
    "pmcmc_last_error": ([_P], C.c_char_p),
    "pmcmc_ctx_synchronize": ([_P], C.c_int),
    "pmcmc_ctx_set_profiling": ([_P, C.c_int], C.c_int),
    "pmcmc_ctx_set_profiling_kinds": ([_P, C.c_uint], C.c_int),
    "pmcmc_ctx_kernel_time": ([_P, C.c_int, C.POINTER(_I64), C.POINTER(C.c_double)], C.c_int),
    "pmcmc_ctx_reset_kernel_times": ([_P], C.c_int),
    "pmcmc_buffer_alloc": ([_P, _I64, C.POINTER(_P)], C.c_int),
    "pmcmc_buffer_free": ([_P, _P], C.c_int),
    "pmcmc_buffer_upload": ([_P, _P, _P, _I64], C.c_int),
    "pmcmc_buffer_download": ([_P, _P, _P, _I64], C.c_int),
    "pmcmc_scan_diag_f32": ([_P, _P, _P, _P, _P, _I64, _I64], C.c_int),
    "pmcmc_scan_diag_f64": ([_P, _P, _P, _P, _P, _I64, _I64], C.c_int),
    "pmcmc_scan_diag_f32_noalloc": ([_P, _P, _P, _P, _P, _I64, _I64], C.c_int),
    "pmcmc_affine_seq_f32": ([_P, _P, _P, _P, _P, _I64, _I64], C.c_int),
    "pmcmc_affine_seq_f64": ([_P, _P, _P, _P, _P, _I64, _I64], C.c_int),
    "pmcmc_affine_residual_f32": ([_P, _P, _P, _P, _P, _I64, _I64, C.POINTER(C.c_double)], C.c_int),
    "pmcmc_affine_residual_f64": ([_P, _P, _P, _P, _P, _I64, _I64, C.POINTER(C.c_double)], C.c_int),
    "pmcmc_affine_compare_f32": ([_P, _P, _P, _I64, C.c_double, C.c_double, C.POINTER(C.c_double)], C.c_int),
    "pmcmc_affine_compare_f64": ([_P, _P, _P, _I64, C.c_double, C.c_double, C.POINTER(C.c_double)], C.c_int),
    "pmcmc_scan_full_f32": ([_P, _P, _P, _P, _P, _I64, _I64], C.c_int),
    "pmcmc_scan_full_f64": ([_P, _P, _P, _P, _P, _I64, _I64], C.c_int),
    "pmcmc_scan_tune": ([_P, C.c_int, _I64, _I64, C.c_int32, C.POINTER(TileShape)], C.c_int),
    "pmcmc_deer_tune": ([_P, _P, _P, C.POINTER(DeerParams), _P, _P, _P, C.c_int32, C.POINTER(TileShape)], C.c_int),
    "pmcmc_tune_get": ([C.c_int, C.c_int, _I64, _I64, C.POINTER(TileShape)], C.c_int),
    "pmcmc_tune_set": ([C.c_int, C.c_int, _I64, _I64, C.POINTER(TileShape)], C.c_int),
    "pmcmc_tune_clear": ([], C.c_int),
    "pmcmc_tune_count": ([], _I64),
    "pmcmc_tune_entry": ([_I64, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(_I64), C.POINTER(C.c_int32),
                          C.POINTER(TileShape)], C.c_int),
    "pmcmc_matmul": ([_P, C.c_int, C.c_int, C.c_int, _I64, _I64, _I64, _P, _I64, _P, _I64, _P, _I64], C.c_int),
    "pmcmc_dot": ([_P, C.c_int, _I64, _P, _P, _P], C.c_int),
    "pmcmc_dotsum": ([_P, C.c_int, _I64, _I64, _P, _I64, _P, _I64, _P], C.c_int),
    "pmcmc_target_diag_gaussian": ([_P, _I64, C.c_int, _P, _P, C.POINTER(_P)], C.c_int),
    "pmcmc_target_logistic": ([_P, _I64, _I64, C.c_int, _P, _P, C.c_double, C.POINTER(_P)], C.c_int),
    "pmcmc_target_gaussian": ([_P, _I64, C.c_int, _P, _P, C.POINTER(_P)], C.c_int),
    "pmcmc_target_logistic_set_prior": ([_P, _P, C.c_int], C.c_int),
    "pmcmc_target_destroy": ([_P], C.c_int),
    "pmcmc_target_dim": ([_P], _I64),
    "pmcmc_mala_fwd_and_jvp": ([_P, _P, C.c_int, C.c_double, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64], C.c_int),
    "pmcmc_mala_run_batched": ([_P, _P, C.c_int, C.c_double, _P, _P, _P, _P, _P, _I64, _I64, _I64], C.c_int),
    "pmcmc_deer_ws_create": ([_P, _I64, _I64, C.c_int, C.POINTER(_P)], C.c_int),
    "pmcmc_deer_ws_destroy": ([_P], C.c_int),
    "pmcmc_deer_solve": ([_P, _P, _P, C.POINTER(DeerParams), _P, _P, _P, _P, _P, _P, C.POINTER(DeerInfo)], C.c_int),
    "pmcmc_target_logdensity_batch": ([_P, _P, C.c_int, _P, _P, _I64], C.c_int),
    "pmcmc_deer_last_trajectory": ([_P, _P, _P], C.c_int),
    "pmcmc_tape_generate": ([_P, C.c_uint64, C.c_int, _P, _P, _I64, _I64], C.c_int),
    "pmcmc_batch_create": ([_P, C.c_int32, C.POINTER(ChainDesc), C.POINTER(DeerParams), _I64, C.POINTER(_P)],
                           C.c_int),
    "pmcmc_batch_set_graph": ([_P, C.c_int32], C.c_int),
    "pmcmc_batch_set_profiling": ([_P, C.c_int32], C.c_int),
    "pmcmc_batch_pass_stats": ([_P, _P, _P, _P], C.c_int),
    "pmcmc_batch_glm_pass_stats": ([_P, _P, _P, _P], C.c_int),
    "pmcmc_batch_solve": ([_P, C.POINTER(DeerInfo)], C.c_int),
    "pmcmc_batch_destroy": ([_P], C.c_int),
    "pmcmc_deer_adjoint": ([_P, _P, _P, C.POINTER(DeerParams), _P, _P, _P, _P, _P, C.POINTER(C.c_double), _P, _P,
                            C.POINTER(DeerInfo)], C.c_int),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise PmcmcError(
                    f"{LIB_PATH} not found: build it with `python parallelmcmc.jl_amd/build_lib.py`"
                )
            lib = C.CDLL(LIB_PATH)
            for name, (args, res) in _SIGS.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = res
            _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc, ctx=None, what=""):
    if rc == PMCMC_OK:
        return
    msg = ""
    if ctx is not None:
        m = load().pmcmc_last_error(ctx)
        msg = m.decode() if m else ""
    text = f"{what}: {msg}" if what else msg
    if rc == PMCMC_ERR_ARGUMENT:
        raise ArgumentError(text)
    if rc == PMCMC_ERR_DIMENSION:
        raise DimensionMismatch(text)
    if rc == PMCMC_ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    raise PmcmcError(f"status {rc}: {text}")
