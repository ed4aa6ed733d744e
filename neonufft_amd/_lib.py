"""Loads libneonufft_amd.so and declares the C ABI of include/neonufft_amd.h."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libneonufft_amd.so")

# Every symbol the C header declares (checked by the CPU test suite).
EXPORTED = (
    "nnf_options_default", "nnf_last_error", "nnf_version",
    "nnf_plan_create", "nnf_plan_set_points", "nnf_plan_set_points_noalloc",
    "nnf_plan_transform_type_1",
    "nnf_plan_transform_type_2", "nnf_plan_destroy",
    "nnf_plan_t3_grid_memory_size", "nnf_plan_t3_create", "nnf_plan_t3_create_bounds",
    "nnf_plan_t3_set_input_points", "nnf_plan_t3_set_output_points", "nnf_plan_t3_add_input",
    "nnf_plan_t3_transform", "nnf_plan_t3_reset", "nnf_plan_t3_destroy",
    "nnf_plan_get_info", "nnf_plan_t3_get_info", "nnf_plan_spread", "nnf_plan_interp",
    "nnf_plan_fine_grid", "nnf_plan_fseries", "nnf_plan_create_tuned",
    "nnf_batch_create", "nnf_batch_set_points", "nnf_batch_transform_type_1",
    "nnf_batch_transform_type_2", "nnf_batch_get_info", "nnf_batch_destroy",
    "nnf_autotune", "nnf_autotune_clear_cache", "nnf_plan_spread_tiles", "nnf_plan_reduce_tiles",
    "nnf_batch_create_tuned",
)


class NnfOptions(ctypes.Structure):
    _fields_ = [
        ("tol", ctypes.c_double),
        ("upsampfac", ctypes.c_double),
        ("recenter_threshold", ctypes.c_double),
        ("num_threads", ctypes.c_int),
        ("sort_input", ctypes.c_int),
        ("sort_output", ctypes.c_int),
        ("kernel_approximation", ctypes.c_int),
        ("order", ctypes.c_int),
        ("kernel_type", ctypes.c_int),
    ]


class NnfTuning(ctypes.Structure):
    _fields_ = [
        ("bin_size", ctypes.c_int32 * 3),
        ("max_points", ctypes.c_int32),
        ("spread_mode", ctypes.c_int32),
        ("band_rows", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
    ]


class NnfPlanInfo(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int),
        ("dim", ctypes.c_int),
        ("n_spread", ctypes.c_int),
        ("es_beta", ctypes.c_double),
        ("modes", ctypes.c_int64 * 3),
        ("grid", ctypes.c_int64 * 3),
        ("spread_grid", ctypes.c_int64 * 3),
        ("bin_size", ctypes.c_int64 * 3),
        ("num_bins", ctypes.c_int64),
        ("num_subproblems", ctypes.c_int64),
        ("num_points", ctypes.c_int64),
        ("num_out_points", ctypes.c_int64),
        ("horner_degree", ctypes.c_int),
        ("quad_nodes", ctypes.c_int),
        ("spread_mode", ctypes.c_int),
        ("band_rows", ctypes.c_int),
    ]


class NnfBatchInfo(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int),
        ("dim", ctypes.c_int),
        ("n_spread", ctypes.c_int),
        ("count", ctypes.c_int64),
        ("num_points", ctypes.c_int64),
        ("num_subproblems", ctypes.c_int64),
        ("num_shapes", ctypes.c_int64),
        ("grid_cells", ctypes.c_int64),
        ("graph", ctypes.c_int),
        ("spread_mode", ctypes.c_int),
    ]


_lib = None


def lib():
    """The loaded library. Raises if it has not been built: there is no
    fallback implementation."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m neonufft_amd.build` "
            "(the GPU plans have no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    vpp = ctypes.POINTER(ctypes.c_void_p)
    i64 = ctypes.c_int64
    i64p = ctypes.POINTER(ctypes.c_int64)
    dp = ctypes.POINTER(ctypes.c_double)
    c_int = ctypes.c_int
    optp = ctypes.POINTER(NnfOptions)
    L.nnf_options_default.argtypes = [optp]
    L.nnf_options_default.restype = None
    L.nnf_last_error.restype = ctypes.c_char_p
    L.nnf_version.restype = ctypes.c_char_p
    L.nnf_plan_create.argtypes = [vpp, c_int, c_int, optp, c_int, i64, vpp, i64p, vp]
    L.nnf_plan_create_tuned.argtypes = [vpp, c_int, c_int, optp, c_int, i64, vpp, i64p, vp,
                                        ctypes.POINTER(NnfTuning)]
    L.nnf_plan_set_points.argtypes = [vp, i64, vpp]
    L.nnf_plan_set_points_noalloc.argtypes = [vp, i64, vpp]
    L.nnf_plan_transform_type_1.argtypes = [vp, vp, vp, i64p]
    L.nnf_plan_transform_type_2.argtypes = [vp, vp, i64p, vp]
    L.nnf_plan_destroy.argtypes = [vp]
    L.nnf_plan_t3_grid_memory_size.argtypes = [c_int, c_int, optp, dp, dp, dp, dp,
                                               ctypes.POINTER(ctypes.c_uint64)]
    L.nnf_plan_t3_create.argtypes = [vpp, c_int, c_int, optp, c_int, i64, vpp, i64, vpp, vp, i64]
    L.nnf_plan_t3_create_bounds.argtypes = [vpp, c_int, c_int, optp, c_int, dp, dp, dp, dp, vp,
                                            i64]
    L.nnf_plan_t3_set_input_points.argtypes = [vp, i64, vpp]
    L.nnf_plan_t3_set_output_points.argtypes = [vp, i64, vpp]
    L.nnf_plan_t3_add_input.argtypes = [vp, vp, i64]
    L.nnf_plan_t3_transform.argtypes = [vp, vp, i64]
    L.nnf_plan_t3_reset.argtypes = [vp]
    L.nnf_plan_t3_destroy.argtypes = [vp]
    L.nnf_plan_get_info.argtypes = [vp, ctypes.POINTER(NnfPlanInfo)]
    L.nnf_plan_t3_get_info.argtypes = [vp, ctypes.POINTER(NnfPlanInfo)]
    L.nnf_plan_spread.argtypes = [vp, vp, c_int]
    L.nnf_plan_spread_tiles.argtypes = [vp, vp]
    L.nnf_plan_reduce_tiles.argtypes = [vp, c_int]
    L.nnf_plan_interp.argtypes = [vp, vp]
    L.nnf_plan_fine_grid.argtypes = [vp, vpp]
    L.nnf_plan_fseries.argtypes = [vp, c_int, vpp]
    L.nnf_batch_create.argtypes = [vpp, c_int, c_int, optp, c_int, i64, i64p, vpp, i64p, vp,
                                   c_int]
    L.nnf_batch_create_tuned.argtypes = [vpp, c_int, c_int, optp, c_int, i64, i64p, vpp, i64p,
                                         vp, c_int, vp]
    L.nnf_batch_set_points.argtypes = [vp, i64p, vpp]
    L.nnf_batch_transform_type_1.argtypes = [vp, vpp, vpp]
    L.nnf_batch_transform_type_2.argtypes = [vp, vpp, vpp]
    L.nnf_batch_get_info.argtypes = [vp, ctypes.POINTER(NnfBatchInfo)]
    L.nnf_batch_destroy.argtypes = [vp]
    L.nnf_autotune.argtypes = [c_int, c_int, optp, c_int, i64, vpp, i64p, vp, c_int, c_int, c_int,
                               ctypes.POINTER(NnfTuning), dp]
    L.nnf_autotune_clear_cache.argtypes = []
    for name in EXPORTED:
        if name not in ("nnf_last_error", "nnf_version", "nnf_options_default"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L
