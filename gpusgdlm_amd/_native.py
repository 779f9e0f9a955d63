"""ctypes binding of libsgdlm_hip.so (the C-ABI declared in include/sgdlm.h).

The library is built in-tree by ``python -m gpusgdlm_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsgdlm_hip.so")

SGDLM_OK = 0
SGDLM_ERR_ARG = 1
SGDLM_ERR_STATE = 2
SGDLM_ERR_DEVICE = 3
SGDLM_ERR_ALLOC = 4
SGDLM_ERR_NUMERIC = 5

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16X3 = 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}

BATCH_GROUP = 0
BATCH_STREAMS = 1

_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint)
_ctx = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "sgdlm_create": (C.c_int, [C.POINTER(_ctx), C.c_size_t]),
    "sgdlm_destroy": (C.c_int, [_ctx]),
    "sgdlm_last_error": (C.c_char_p, [_ctx]),
    "sgdlm_get_no_series": (C.c_int, [_ctx, C.POINTER(C.c_size_t)]),
    "sgdlm_get_max_p": (C.c_int, [_ctx, C.POINTER(C.c_size_t)]),
    "sgdlm_get_evolution_matrix_configuration": (C.c_int, [_ctx, C.POINTER(C.c_int)]),
    "sgdlm_init_memory": (C.c_int, [_ctx, C.c_size_t, C.c_size_t]),
    "sgdlm_clear_memory": (C.c_int, [_ctx]),
    "sgdlm_manage_evo_memory": (C.c_int, [_ctx, C.c_int]),
    "sgdlm_get_nsim": (C.c_int, [_ctx, C.POINTER(C.c_size_t)]),
    "sgdlm_get_nsim_batch": (C.c_int, [_ctx, C.POINTER(C.c_size_t)]),
    "sgdlm_init_sim_memory": (C.c_int, [_ctx, C.c_size_t, C.c_size_t]),
    "sgdlm_clear_sim_memory": (C.c_int, [_ctx]),
    "sgdlm_init_forecast_priors": (C.c_int, [_ctx]),
    "sgdlm_get_is_prior": (C.c_int, [_ctx, C.POINTER(C.c_int)]),
    "sgdlm_set_is_prior": (C.c_int, [_ctx, C.c_int]),
    "sgdlm_get_parameters": (C.c_int, [_ctx, _dp, _dp, _dp, _dp]),
    "sgdlm_set_parameters": (C.c_int, [_ctx, _dp, _dp, _dp, _dp]),
    "sgdlm_get_discount_factors": (C.c_int, [_ctx, _dp, _dp]),
    "sgdlm_set_discount_factors": (C.c_int, [_ctx, _dp, _dp]),
    "sgdlm_get_evolution_matrix": (C.c_int, [_ctx, _dp]),
    "sgdlm_set_evolution_matrix": (C.c_int, [_ctx, _dp]),
    "sgdlm_get_parental_sets": (C.c_int, [_ctx, _up, _up]),
    "sgdlm_set_parental_sets": (C.c_int, [_ctx, _up, _up]),
    "sgdlm_compute_prior": (C.c_int, [_ctx, _dp, C.c_int]),
    "sgdlm_evolve_forecast_samples": (C.c_int, [_ctx, _dp]),
    "sgdlm_compute_forecast": (C.c_int, [_ctx, _dp, _dp, _dp, _dp, C.c_int]),
    "sgdlm_compute_posterior": (C.c_int, [_ctx, _dp, _dp]),
    "sgdlm_compute_vb_posterior": (C.c_int, [_ctx, _dp, _dp, _dp, _dp, _dp, _dp]),
    "sgdlm_compute_vb_posterior_noalloc": (C.c_int, [_ctx, C.c_size_t, _dp, _dp, _dp, _dp, _dp, _dp,
                                                     C.POINTER(C.c_size_t)]),
    "sgdlm_set_precision": (C.c_int, [_ctx, C.c_int]),
    "sgdlm_get_precision": (C.c_int, [_ctx, C.POINTER(C.c_int)]),
    "sgdlm_set_rng_state": (C.c_int, [_ctx, C.c_uint64, C.c_uint32]),
    "sgdlm_get_rng_state": (C.c_int, [_ctx, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "sgdlm_get_vb_result": (C.c_int, [_ctx, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "sgdlm_set_timing": (C.c_int, [_ctx, C.c_int]),
    "sgdlm_get_timing": (C.c_int, [_ctx, C.c_char_p, _dp, C.POINTER(C.c_long)]),
    "sgdlm_get_lu_fallback_count": (C.c_int, [_ctx, C.POINTER(C.c_long)]),
    "sgdlm_get_lu_config": (C.c_int, [_ctx, C.POINTER(C.c_int), C.POINTER(C.c_size_t), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "sgdlm_get_lu_solve_config": (C.c_int, [_ctx, C.POINTER(C.c_int), C.POINTER(C.c_size_t), C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sgdlm_set_lu_tuning": (C.c_int, [_ctx, C.c_int, C.c_int]),
    "sgdlm_get_lu_kernel": (C.c_int, [_ctx, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sgdlm_set_lu_kernel": (C.c_int, [_ctx, C.c_int]),
    "sgdlm_set_lu_autotune": (C.c_int, [_ctx, C.c_int]),
    "sgdlm_set_graph_mode": (C.c_int, [_ctx, C.c_int]),
    "sgdlm_get_graph_stats": (C.c_int, [_ctx, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "sgdlm_get_lu_autotune": (C.c_int, [_ctx, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), _dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sgdlm_set_lu_max_front": (C.c_int, [_ctx, C.c_int]),
    "sgdlm_batch_compute_vb_posterior": (C.c_int, [C.POINTER(_ctx), C.c_size_t, C.POINTER(C.c_int)]),
    "sgdlm_set_batch_mode": (C.c_int, [_ctx, C.c_int]),
    "sgdlm_get_lu_solve_kernel": (C.c_int, [_ctx, C.POINTER(C.c_int)]),
    "sgdlm_set_lu_solve_kernel": (C.c_int, [_ctx, C.c_int]),
    "sgdlm_get_batch_stats": (C.c_int, [_ctx, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "sgdlm_get_lu_elimination": (C.c_int, [_ctx, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), C.POINTER(C.c_long)]),
}

_lib = None


def load(path: str | None = None) -> C.CDLL:
    """Load the HIP library (cached).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} not found: build the HIP extension with `python -m gpusgdlm_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def dptr(a):
    """double* of a contiguous float64 numpy array (or NULL for None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def uptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_up)
