"""ctypes binding of libamgt_mbsr.so (the C-ABI declared in include/amgt_mbsr.h and
include/amgt_policy.h).

The library must be built in-tree (``python -m amgt_amd.build``).  There is no
CPU fallback: every entry point raises when the HIP library is missing.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamgt_mbsr.so")

AMGT_OK = 0
AMGT_ERR_ARG = 1
AMGT_ERR_DIM = 2
AMGT_ERR_ALLOC = 3
AMGT_ERR_HIP = 4
AMGT_ERR_INPUT = 5
AMGT_ERR_UNSUPPORTED = 6

AMGT_DIAG_VALUE = 0
AMGT_DIAG_INV_ABS = 1
AMGT_DIAG_INV_L1 = 2
AMGT_DIAG_INV = 3
AMGT_SPMV_DIRECT = 0
AMGT_SPMV_STAGED = 1
AMGT_SPMV_BLOCK = 2
AMGT_SPMV_BALANCED = 3
SPMV_KERNEL_NAMES = {0: "direct", 1: "staged", 2: "block", 3: "balanced"}

AMGT_DOUBLE = 0
AMGT_FLOAT = 1
AMGT_HALF = 2
AMGT_NOALLOC_SCRATCH_BYTES = 8 * 1024 * 1024

POLICY_FEATURES = 12
POLICY_OUTPUTS = 4
POLICY_PARAMS = 5252

# every symbol include/amgt_mbsr.h and include/amgt_policy.h declare
EXPORTS = (
    "amgt_status_string", "amgt_last_error", "amgt_set_allocator", "amgt_precision_for_level",
    "amgt_set_mixed_precision", "amgt_mbsr_create_from_csr", "amgt_mbsr_destroy",
    "amgt_mbsr_update_values", "amgt_mbsr_get_info", "amgt_mbsr_get_arrays",
    "amgt_mbsr_set_level", "amgt_mbsr_set_precision", "amgt_mbsr_set_spmv_lanes",
    "amgt_mbsr_set_spmv_kernel",
    "amgt_mbsr_autotune", "amgt_mbsr_spmv_sweep", "amgt_mbsr_tune_from_stats",
    "amgt_mbsr_spmv", "amgt_mbsr_spmv_noalloc", "amgt_mbsr_spmm",
    "amgt_restrict_plan_create", "amgt_restrict_plan_destroy", "amgt_mbsr_residual_restrict",
    "amgt_pcg_work_bytes", "amgt_dot", "amgt_pcg_update_xr", "amgt_pcg_update_p",
    "amgt_mbsr_spgemm",
    "amgt_mbsr_to_csr", "amgt_free", "amgt_csr_spgemm", "amgt_csr_matvec",
    "amgt_mbsr_diagonal", "amgt_mbsr_relax_jacobi", "amgt_mbsr_transpose",
    "amgt_read_ij", "amgt_read_mm", "amgt_host_free", "amgt_mbsr_spmm_interleaved",
    "amgt_amg_strength", "amgt_amg_pmis", "amgt_amg_extpi_interp", "amgt_amg_ruge_first_pass",
    "amgt_amg_hmis", "amgt_amg_rap_host_order",
    "amgt_policy_work_bytes", "amgt_policy_forward", "amgt_policy_loss_grad", "amgt_policy_adam",
    "amgt_level_stats_work_bytes", "amgt_level_stats", "amgt_interp_trilinear",
)


class AmgtError(RuntimeError):
    """A libamgt_mbsr entry point returned a non-zero status."""

    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: status {status}: {message}")
        self.status = status


class AmgtLibraryMissing(ImportError):
    pass


class MbsrInfo(ctypes.Structure):
    _fields_ = [
        ("nrows", ctypes.c_int32), ("ncols", ctypes.c_int32), ("brow", ctypes.c_int32),
        ("bcol", ctypes.c_int32), ("nblk", ctypes.c_int64), ("nnz", ctypes.c_int64),
        ("stand", ctypes.c_double), ("avg_nnz", ctypes.c_double), ("fill", ctypes.c_double),
        ("max_row_blocks", ctypes.c_int32), ("level", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("spmv_lanes", ctypes.c_int32),
        ("spmv_kernel", ctypes.c_int32),
    ]


class MbsrArrays(ctypes.Structure):
    _fields_ = [
        ("blk_ptr", ctypes.c_void_p), ("blk_col", ctypes.c_void_p), ("val_ptr", ctypes.c_void_p),
        ("blk_map", ctypes.c_void_p), ("val", ctypes.c_void_p),
    ]


_lib = None
_hip = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_pvp = ctypes.POINTER(ctypes.c_void_p)
_pi64 = ctypes.POINTER(ctypes.c_int64)


def load():
    """Loads the in-tree HIP library; raises AmgtLibraryMissing when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AmgtLibraryMissing(
            f"{LIB_PATH} is not built; run `python -m amgt_amd.build` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "amgt_status_string": (ctypes.c_char_p, [_int]),
        "amgt_last_error": (ctypes.c_char_p, []),
        "amgt_set_allocator": (_int, [_vp, _vp, _vp]),
        "amgt_precision_for_level": (_int, [_int]),
        "amgt_set_mixed_precision": (_int, [_int]),
        "amgt_mbsr_create_from_csr": (_int, [_i32, _i32, _i64, _vp, _vp, _vp, _vp, _pvp]),
        "amgt_mbsr_destroy": (_int, [_vp]),
        "amgt_mbsr_update_values": (_int, [_vp, _vp, _vp, _vp, _vp]),
        "amgt_mbsr_get_info": (_int, [_vp, ctypes.POINTER(MbsrInfo)]),
        "amgt_mbsr_get_arrays": (_int, [_vp, ctypes.POINTER(MbsrArrays)]),
        "amgt_mbsr_set_level": (_int, [_vp, _int]),
        "amgt_mbsr_set_precision": (_int, [_vp, _int]),
        "amgt_mbsr_set_spmv_lanes": (_int, [_vp, _int]),
        "amgt_mbsr_set_spmv_kernel": (_int, [_vp, _int]),
        "amgt_mbsr_autotune": (_int, [_vp, _vp]),
        "amgt_mbsr_spmv_sweep": (_int, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
        "amgt_mbsr_tune_from_stats": (_int, [_vp]),
        "amgt_mbsr_spmv": (_int, [_vp, _int, _dbl, _vp, _dbl, _vp, _vp, _vp]),
        "amgt_mbsr_spmv_noalloc": (_int, [_vp, _dbl, _vp, _dbl, _vp, _vp, _vp]),
        "amgt_restrict_plan_create": (_int, [_vp, _vp, _vp, _pvp]),
        "amgt_restrict_plan_destroy": (_int, [_vp]),
        "amgt_pcg_work_bytes": (ctypes.c_size_t, []),
        "amgt_dot": (_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
        "amgt_pcg_update_xr": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "amgt_pcg_update_p": (_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
        "amgt_mbsr_residual_restrict": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "amgt_mbsr_spmm": (_int, [_vp, _int, _int, _dbl, _vp, _i64, _dbl, _vp, _i64, _vp, _i64,
                                  _vp]),
        "amgt_mbsr_spgemm": (_int, [_vp, _vp, _vp, _pvp]),
        "amgt_mbsr_to_csr": (_int, [_vp, _vp, _pi64, _pvp, _pvp, _pvp]),
        "amgt_free": (_int, [_vp]),
        "amgt_csr_spgemm": (_int, [_i32, _i32, _i32, _i64, _vp, _vp, _vp, _pvp, _i64, _vp, _vp,
                                   _vp, _pvp, _pi64, _pvp, _pvp, _pvp, _pvp, _vp]),
        "amgt_csr_matvec": (_int, [_int, _dbl, _i32, _i32, _i64, _vp, _vp, _vp, _pvp, _vp, _dbl,
                                   _vp, _vp, _vp]),
        "amgt_mbsr_diagonal": (_int, [_vp, _int, _vp, _vp]),
        "amgt_mbsr_transpose": (_int, [_vp, _vp, _pvp]),
        "amgt_mbsr_spmm_interleaved": (_int, [_vp, _int, _int, _dbl, _vp, _dbl, _vp, _vp, _vp]),
        "amgt_read_ij": (_int, [ctypes.c_char_p, _int, ctypes.POINTER(_i32),
                                ctypes.POINTER(_i32), _pi64, _pvp, _pvp, _pvp]),
        "amgt_read_mm": (_int, [ctypes.c_char_p, ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                _pi64, _pvp, _pvp, _pvp]),
        "amgt_host_free": (None, [_vp]),
        "amgt_mbsr_relax_jacobi": (_int, [_vp, _dbl, _vp, _vp, _vp, _vp, _vp]),
        "amgt_amg_strength": (_int, [_i32, _vp, _vp, _vp, _dbl, _dbl, _pi64, _pvp, _pvp, _vp]),
        "amgt_amg_pmis": (_int, [_i32, _vp, _vp, _int, _vp, ctypes.POINTER(_i32), _vp]),
        "amgt_amg_ruge_first_pass": (_int, [_i32, _vp, _vp, _vp]),
        "amgt_amg_hmis": (_int, [_i32, _vp, _vp, _vp, ctypes.POINTER(_i32), _vp]),
        "amgt_amg_rap_host_order": (_int, [_i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _pi64, _pvp,
                                           _pvp, _pvp, _vp]),
        "amgt_amg_extpi_interp": (_int, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _int,
                                         ctypes.POINTER(_i32), _pi64, _pvp, _pvp, _pvp, _vp]),
        "amgt_policy_work_bytes": (ctypes.c_size_t, [_i64]),
        "amgt_policy_forward": (_int, [_vp, _vp, _i64, _vp, _vp, ctypes.c_size_t, _vp]),
        "amgt_policy_loss_grad": (_int, [_vp, _vp, _vp, _vp, _i64, ctypes.c_float, _vp, _vp, _vp,
                                         ctypes.c_size_t, _vp]),
        "amgt_level_stats_work_bytes": (ctypes.c_size_t, [_i32]),
        "amgt_interp_trilinear": (_int, [_i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                         ctypes.POINTER(_i32), _pi64, _pvp, _pvp, _pvp, _vp]),
        "amgt_level_stats": (_int, [_i32, _vp, _vp, _vp, ctypes.POINTER(_dbl), _vp, ctypes.c_size_t,
                                    _vp]),
        "amgt_policy_adam": (_int, [_vp, _vp, _vp, _vp, _i64, _i64, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status: int, where: str) -> None:
    if status != AMGT_OK:
        msg = load().amgt_last_error()
        raise AmgtError(status, where, msg.decode() if msg else "")


def hip():
    """The HIP runtime (the same instance torch uses), for raw device copies."""
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMemcpy.restype = ctypes.c_int
        h.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpyAsync.restype = ctypes.c_int
        h.hipMemcpyAsync.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp]
        _hip = h
    return _hip


HIP_MEMCPY_D2D = 3
