"""ctypes binding of the C ABI in include/pbbfmm3d_amd.h.

The shared library is built in-tree (``__graft_entry__.build()`` or
``make -C pbbfmm3d_amd/csrc``) as ``pbbfmm3d_amd/libpbbfmm3d_amd.so``.  There
is no fallback: if the library is missing every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpbbfmm3d_amd.so")

# pbb_status
PBB_OK = 0
PBB_ERR_INVALID_ARG = 1
PBB_ERR_UNSUPPORTED = 2
PBB_ERR_HIP = 3
PBB_ERR_OUT_OF_MEMORY = 4
PBB_ERR_NO_DEVICE = 5
PBB_ERR_INTERNAL = 6
PBB_ERR_NOT_FOUND = 7

# pbb_kernel_id
KERNEL_NONE = 0
KERNEL_LAPLACIAN = 1
KERNEL_ONEOVERR2 = 2
KERNEL_GAUSSIAN = 3
KERNEL_EXPONENTIAL = 4
KERNEL_SOFT_LAPLACIAN = 5
KERNEL_QUADRIC = 6
KERNEL_INV_QUADRIC = 7
KERNEL_LOGARITHM = 8
KERNEL_THIN_PLATE = 9
KERNEL_R2 = 10
KERNEL_CONSTANT = 11
KERNEL_POLY_XY = 12
KERNEL_EXP_ANISO = 13

PREC_FP64 = 0
PREC_MIXED = 1          # BF16 x3 split M2L
PREC_MIXED_F16X2 = 2   # FP16 x2 split M2L
PREC_MIXED_BF16X2 = 3  # BF16 x2 split M2L
M2L_CHUNK_MAJOR = 1
M2L_ROW_SPLIT = 2
M2L_WAVE_ROWS = 4
M2L_SPLIT_ACC = 8


def M2L_DIR_GROUPS(g):
    """M2L flag: g groups of operators, one launch per group (mixed modes)."""
    return int(g) << 8


NUM_PHASES = 8
PLAN_HOST_TREE = 1
PLAN_REJECT_OUTSIDE = 2
PLAN_EXPANSION_ROUTE = 4
OPS_HOST_PRECOMPUTE = 1
BATCH_TRANSPOSE = 1
BATCH_OVERWRITE = 2
PHASE_NAMES = ("p2m", "m2m", "vt_proxy", "m2l", "u_expand", "l2l", "near", "total")


class KernelDesc(C.Structure):
    _fields_ = [("id", C.c_int32), ("symmetry", C.c_int32), ("homogen", C.c_double),
                ("param", C.c_double * 4)]


class OpsInfo(C.Structure):
    _fields_ = [("order", C.c_int32), ("n3", C.c_int32), ("rank_f", C.c_int32),
                ("rank_s", C.c_int32), ("nsets", C.c_int32), ("homogeneous", C.c_int32),
                ("tree_level", C.c_int32), ("reserved", C.c_int32), ("L", C.c_double),
                ("homogen", C.c_double), ("eps", C.c_double)]


class PlanInfo(C.Structure):
    _fields_ = [("npoints", C.c_int64), ("tree_level", C.c_int32),
                ("nonempty_levels", C.c_int32), ("boxes", C.c_int64 * 16),
                ("m2l_pairs", C.c_int64), ("near_pairs", C.c_int64),
                ("max_leaf_points", C.c_int64), ("center", C.c_double * 3),
                ("outside_points", C.c_int64), ("ntargets", C.c_int64),
                ("nsources", C.c_int64), ("m2l_pairs_level", C.c_int64 * 16)]


class BatchEntry(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("w_dev", C.c_void_p), ("out_dev", C.c_void_p),
                ("flags", C.c_int32), ("reserved", C.c_int32)]


HOST_KERNEL_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.POINTER(C.c_double),
                             C.POINTER(C.c_double))

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)

_SIGNATURES = {
    "pbb_last_error": (C.c_char_p, []),
    "pbb_abi_version": (C.c_int, []),
    "pbb_device_count": (C.c_int, []),
    "pbb_ops_create": (C.c_int, [C.POINTER(KernelDesc), HOST_KERNEL_FN, _vp, C.c_double,
                                 C.c_int32, C.c_int32, C.c_double, C.c_int32, C.POINTER(_vp)]),
    "pbb_ops_create_ex": (C.c_int, [C.POINTER(KernelDesc), HOST_KERNEL_FN, _vp, C.c_double,
                                    C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "pbb_ops_create_from_arrays": (C.c_int, [C.POINTER(KernelDesc), C.c_double, C.c_int32,
                                             C.c_int32, C.c_double, C.c_int32, C.c_int32,
                                             C.c_int32, _dp, _dp, _dp, C.POINTER(_vp)]),
    "pbb_ops_write_cache": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_char_p]),
    "pbb_ops_read_cache": (C.c_int, [C.POINTER(KernelDesc), C.c_double, C.c_int32, C.c_int32, C.c_double,
                                     C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "pbb_ops_query": (C.c_int, [_vp, C.POINTER(OpsInfo)]),
    "pbb_ops_export": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "pbb_ops_destroy": (None, [_vp]),
    "pbb_ops_set_mixed_tolerance": (C.c_int, [_vp, C.c_double]),
    "pbb_ops_set_fp64_recompressed": (C.c_int, [_vp, C.c_int32]),
    "pbb_ops_mixed_ranks": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _dp]),
    "pbb_plan_create": (C.c_int, [_vp, _dp, C.c_int64, C.POINTER(_vp)]),
    "pbb_plan_create_ex": (C.c_int, [_vp, _dp, C.c_int64, C.c_int32, C.POINTER(_vp)]),
    "pbb_plan_create_ts": (C.c_int, [_vp, _dp, C.c_int64, _dp, C.c_int64, C.c_int32, C.POINTER(_vp)]),
    "pbb_plan_query": (C.c_int, [_vp, C.POINTER(PlanInfo)]),
    "pbb_plan_update_points": (C.c_int, [_vp, _dp, C.c_int64, _vp]),
    "pbb_plan_matvec": (C.c_int, [_vp, _vp, C.c_int32, _vp, C.c_int32, _vp]),
    "pbb_plan_matvec_noalloc": (C.c_int, [_vp, _vp, C.c_int32, _vp, C.c_int32, _vp]),
    "pbb_plan_matvec_transpose": (C.c_int, [_vp, _vp, C.c_int32, _vp, C.c_int32, _vp]),
    "pbb_plan_grad_positions": (C.c_int, [_vp, _vp, _vp, C.c_int32, _vp, C.c_int32, _vp]),
    "pbb_batch_create": (C.c_int, [C.POINTER(_vp), C.c_int32, C.POINTER(_vp), C.POINTER(_vp), C.c_int32,
                                   C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "pbb_batch_create_entries": (C.c_int, [C.POINTER(BatchEntry), C.c_int32, C.c_int32, C.c_int32,
                                           C.POINTER(_vp)]),
    "pbb_batch_launch": (C.c_int, [_vp, _vp]),
    "pbb_batch_destroy": (None, [_vp]),
    "pbb_plan_timings": (C.c_int, [_vp, _dp]),
    "pbb_plan_set_workspace_limit": (C.c_int, [_vp, C.c_int64]),
    "pbb_plan_set_timing": (C.c_int, [_vp, C.c_int32]),
    "pbb_plan_set_m2l_tile": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "pbb_plan_set_m2l_tile_level": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "pbb_plan_get_m2l_tile": (C.c_int, [_vp, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "pbb_plan_tune_m2l": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, _vp, C.POINTER(C.c_int32)]),
    "pbb_plan_tune_m2l_within": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, C.c_double, _vp,
                                           C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "pbb_plan_destroy": (None, [_vp]),
    "pbb_matvec": (C.c_int, [_vp, _dp, _dp, C.c_int64, _dp, C.c_int32, _dp]),
    "pbb_matvec_ts": (C.c_int, [_vp, _dp, C.c_int64, _dp, C.c_int64, _dp, C.c_int32, _dp]),
    "pbb_matvec_prec": (C.c_int, [_vp, _dp, C.c_int64, _dp, C.c_int64, _dp, C.c_int32, _dp, C.c_int32]),
    "pbb_direct_sum": (C.c_int, [C.POINTER(KernelDesc), _vp, C.c_int64, _vp, C.c_int64, _vp,
                                 C.c_int32, _vp, _vp]),
}

_lib = None
_lock = threading.Lock()


class PBBError(RuntimeError):
    """A C-ABI call returned a non-zero pbb_status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"pbb status {status}: {message}")
        self.status = status


def lib() -> C.CDLL:
    """Loads the in-tree shared library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                    "or `make -C pbbfmm3d_amd/csrc`")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
        return _lib


def check(status: int) -> None:
    if status != PBB_OK:
        msg = lib().pbb_last_error()
        raise PBBError(status, msg.decode() if msg else "")


def exported_symbols() -> list:
    return list(_SIGNATURES)
