"""ctypes binding of libgsa_hip.so (declarations in include/gsa_hip.h).

The library is the only compute path: if it is missing, or no GPU is visible,
every operation raises — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsa_hip.so")

GSA_OK = 0
GSA_ERR_INVALID = -1
GSA_ERR_DIM = -2
GSA_ERR_HIP = -3
GSA_ERR_OOM = -4
GSA_ERR_UNSUPPORTED = -5

GSA_F64, GSA_F32, GSA_BF16 = 0, 1, 2
GSA_C64, GSA_C128, GSA_VI32, GSA_VI64 = 3, 4, 5, 6
GSA_I32, GSA_I64 = 0, 1
GSA_OP_N, GSA_OP_T, GSA_OP_C = 0, 1, 2
GSA_PLAN_WINDOWS, GSA_PLAN_XCD_ORDER, GSA_PLAN_ROW_VIEW, GSA_PLAN_ROW_BINS, GSA_PLAN_BATCH = 1, 2, 4, 8, 16
GSA_PLAN_TRANSPOSE_BINS, GSA_PLAN_GATHER_VIEW = 32, 64
GSA_FMT_CSR, GSA_FMT_CSC, GSA_FMT_COO = 0, 1, 2
GSA_BATCH_B_ROWMAJOR, GSA_BATCH_C_ROWMAJOR = 1, 2


class ArgumentError(ValueError):
    """Julia's ArgumentError (invalid argument, mixed backends, bad lengths)."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


class GSAError(RuntimeError):
    """A HIP runtime failure inside libgsa_hip.so."""


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_f64 = ctypes.c_double
_f32 = ctypes.c_float
_pi64 = ctypes.POINTER(ctypes.c_int64)


class BatchDesc(ctypes.Structure):
    """gsa_batch_desc (72 bytes)."""
    _fields_ = [
        ("rowptr", _vp), ("colval", _vp), ("nzval", _vp), ("row_scale", _vp),
        ("B", _vp), ("C", _vp),
        ("m", ctypes.c_int32), ("n", ctypes.c_int32), ("k", ctypes.c_int32),
        ("ldb", ctypes.c_int32), ("ldc", ctypes.c_int32), ("flags", ctypes.c_int32),
    ]


# name -> argtypes; every function returns int status unless listed in _RESTYPES
_SIGS = {
    "gsa_init": [],
    "gsa_version": [],
    "gsa_malloc": [ctypes.POINTER(_vp), _i64],
    "gsa_free": [_vp],
    "gsa_memcpy": [_vp, _vp, _i64, _i32, _vp],
    "gsa_memset": [_vp, _i32, _i64, _vp],
    "gsa_stream_create": [ctypes.POINTER(_vp)],
    "gsa_stream_destroy": [_vp],
    "gsa_stream_sync": [_vp],
    "gsa_plan_create": [_vp, _i32, _i64, _i64, _vp, ctypes.POINTER(_vp)],
    "gsa_plan_destroy": [_vp],
    "gsa_plan_set_columns": [_vp, _vp, _vp, _i64],
    "gsa_plan_describe": [_vp, ctypes.POINTER(ctypes.c_int)],
    "gsa_plan_add_transpose": [_vp, _vp, _i32, _i32, _i64, _i64, _i64, _vp, _vp],
    "gsa_plan_add_gather_view": [_vp, _vp, _i32, _i32, _i32, _i64, _i64, _i64, _vp, _vp],
    "gsa_spmm_csr": [_vp, _i32, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64,
                     _vp, _i64, _i32, _vp, _i64, _vp, _vp, _vp],
    "gsa_spmm_csc": [_vp, _i32, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64,
                     _vp, _i64, _i32, _vp, _i64, _vp, _vp, _vp],
    "gsa_spmm_coo": [_vp, _i32, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64,
                     _vp, _i64, _i32, _vp, _i64, _vp, _vp, _vp],
    "gsa_plan_create_coo": [_vp, _i32, _i64, _i64, _i64, _vp, _vp, ctypes.POINTER(_vp)],
    "gsa_plan_create_csc": [_vp, _i32, _i64, _i64, _i64, _vp, _vp, ctypes.POINTER(_vp)],
    "gsa_spmm_csr_bf16": [_vp, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64,
                          _vp, _i64, _i32, _vp, _i64, _i32, _f32, _f32, _vp],
    "gsa_spmm_csr_ex": [_vp, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64,
                        _vp, _i64, _i32, _vp, _i64, _i32, _f64, _f64, _vp],
    "gsa_index_shift": [_vp, _i32, _i32, _i64, _vp, _vp, _i64],
    "gsa_csr_to_coo": [_vp, _i32, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "gsa_coo_to_csr": [_vp, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "gsa_csr_to_csc": [_vp, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "gsa_csr_add_count": [_vp, _i32, _i32, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp,
                          _pi64],
    "gsa_csr_add_fill": [_vp, _i32, _i32, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp,
                         _vp, _vp],
    "gsa_coo_add_begin": [_vp, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                          ctypes.POINTER(_vp), _pi64],
    "gsa_coo_add_finish": [_vp, _vp, _vp, _vp],
    "gsa_csr_dropzeros_count": [_vp, _i32, _i32, _i64, _vp, _vp, _vp, _pi64],
    "gsa_csr_dropzeros_fill": [_vp, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "gsa_coo_dropzeros_count": [_vp, _i32, _i64, _vp, _pi64],
    "gsa_coo_dropzeros_fill": [_vp, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "gsa_kron_coo": [_vp, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                     _vp, _vp, _vp],
    "gsa_kron_csr": [_vp, _i32, _i32, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                     _i64, _vp, _vp, _vp],
    "gsa_spgemm_csr_begin": [_vp, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp,
                             _vp, _i64, ctypes.POINTER(_vp), _pi64],
    "gsa_spgemm_csr_finish": [_vp, _vp, _vp, _vp],
    "gsa_ctx_destroy": [_vp],
    "gsa_trace_ptr": [_vp, _i32, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "gsa_trace_coo": [_vp, _i32, _i32, _i64, _vp, _vp, _vp, _vp],
    "gsa_sum_values": [_vp, _i32, _i64, _vp, _vp],
    "gsa_dot3_ptr": [_vp, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _vp,
                     _vp],
    "gsa_dot3_coo": [_vp, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _i32, _i32,
                     _vp],
    "gsa_symmetric_ptr": [_vp, _i32, _i32, _i64, _i64, _vp, _vp, _vp, _i32, _vp,
                          ctypes.POINTER(_i32)],
    "gsa_add_diag_count": [_vp, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _pi64],
    "gsa_add_diag_fill": [_vp, _i32, _i32, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "gsa_diag_coo": [_vp, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp],
    "gsa_add_to_dense": [_vp, _i32, _i32, _i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64],
    "gsa_spvec_dot": [_vp, _i32, _i32, _i64, _i64, _vp, _vp, _vp, _vp],
    "gsa_norm_values": [_vp, _i32, _i64, _vp, _f64, ctypes.POINTER(_f64)],
    "gsa_batch_plan_create": [_vp, _i32, _i32, _i64, ctypes.POINTER(BatchDesc), _i32,
                              ctypes.POINTER(_vp)],
    "gsa_batch_spmm": [_vp, _vp, _f32, _f32],
    "gsa_batch_spmm_tables": [_vp, _i32, _i32, _vp, _vp, _i64, _vp, _f32, _f32, _i32],
    "gsa_batch_graph_create": [_vp, _vp, _f32, _f32, ctypes.POINTER(_vp)],
    "gsa_graph_launch": [_vp, _vp],
    "gsa_graph_destroy": [_vp],
    "gsa_csr_to_bf16": [_vp, _i32, _i32, _i64, _vp, _vp, _vp, _vp],
    "gsa_spmv_csr_noalloc": [_vp, _i32, _i32, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp,
                             _f64, _f64],
}
_RESTYPES = {"gsa_last_error": ctypes.c_char_p}
# functions returning something other than an int status: name -> (argtypes, restype)
_OTHER = {"gsa_scratch_cached": ([_vp], ctypes.c_int64),
          "gsa_scratch_graph_blocks": ([], ctypes.c_int64)}
GSA_NOALLOC_SCRATCH_BYTES = 4 << 20
GSA_NOALLOC_MAX_NNZ = 100_000_000

_lib = None
_lock = threading.Lock()


def exported_symbols():
    """Names this binding expects the library to export."""
    return sorted(list(_SIGS) + list(_RESTYPES) + list(_OTHER))


def load(path: str | None = None):
    """Load libgsa_hip.so (no GPU needed to load).  Raises if it is not built."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise GSAError(
                f"libgsa_hip.so not found at {p}: build it with "
                "`make -C genericsparsearrays.jl_amd` (or __graft_entry__.build())")
        lib = ctypes.CDLL(p)
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        for name, res in _RESTYPES.items():
            fn = getattr(lib, name)
            fn.argtypes = []
            fn.restype = res
        for name, (args, res) in _OTHER.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        if path is None:
            _lib = lib
        return lib


def lib():
    return load()


def check(status: int, what: str = ""):
    """Map a C status to the reference's exception types."""
    if status == GSA_OK:
        return
    msg = (load().gsa_last_error() or b"").decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if status == GSA_ERR_INVALID:
        raise ArgumentError(text)
    if status == GSA_ERR_DIM:
        raise DimensionMismatch(text)
    if status == GSA_ERR_OOM:
        raise MemoryError(text)
    raise GSAError(text)


def call(name: str, *args):
    fn = getattr(load(), name)
    check(fn(*args), name)
