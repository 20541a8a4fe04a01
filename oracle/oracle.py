"""Python face of the CPU oracle (oracle/gsa_oracle.c).

TEST INFRASTRUCTURE ONLY — the parity checker.  Imported by tests/, by the
smoke check and by bench.py's CPU-baseline leg; never by the product package.

All functions take and return numpy arrays with ONE-BASED int64 indices (the
reference's storage), and follow the reference's semantics as documented in
gsa_oracle.c (each C function cites the reference file:line it restates).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libgsa_oracle.so")
_lib = None

_I = np.int64
_P = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ctypes.CDLL(_SO)
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


_SFX = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32",
        np.dtype(np.complex128): "c128", np.dtype(np.complex64): "c64",
        np.dtype(np.int64): "i64", np.dtype(np.int32): "i32"}


def _sfx(dtype):
    """Suffix of the oracle's instantiation for a value type (integers run as
    unsigned wrapping arithmetic of their width: the same bits)."""
    return _SFX[np.dtype(dtype)]


def _ctv(dtype):
    return ctypes.c_double if np.dtype(dtype) == np.float64 else ctypes.c_float


def _sref(value, dtype):
    """A one-element numpy array holding a scalar of the value type, passed by
    pointer (complex scalars cannot be passed by value through ctypes)."""
    return np.array([value], dtype=dtype)


def _i(a):
    return np.ascontiguousarray(a, dtype=_I)


def max_threads() -> int:
    return int(lib().ora_max_threads())


# ---------------------------------------------------------------- products --
def spmm_csr_N(m, rowptr, colval, nzval, B, C=None, alpha=1.0, beta=0.0, threads=1):
    """mul!(C, A::CSR, B, α, β) — bit-exact restatement (matrix_csr_kernels.jl:1-22).
    B, C are numpy arrays (vector or 2-D, any layout); returns C."""
    T = nzval.dtype
    B2 = np.asarray(B, dtype=T)
    vec = B2.ndim == 1
    Bf = np.asfortranarray(B2.reshape(-1, 1) if vec else B2)
    k = Bf.shape[1]
    Cf = np.zeros((m, k), dtype=T, order="F") if C is None else np.array(
        np.asarray(C, dtype=T).reshape(m, k), order="F", copy=True)
    fn = getattr(lib(), "ora_spmm_csr_N_" + _sfx(T))
    fn.argtypes = [ctypes.c_int64, ctypes.c_int64, _P, _P, _P, _P, ctypes.c_int64, ctypes.c_int,
                   _P, ctypes.c_int64, _P, _P, ctypes.c_int]
    rp, cv, nz = _i(rowptr), _i(colval), np.ascontiguousarray(nzval)
    a_ref, b_ref = _sref(alpha, T), _sref(beta, T)  # kept alive across the call
    fn(m, k, _p(rp), _p(cv), _p(nz), _p(Bf), max(Bf.shape[0], 1), 0, _p(Cf), max(m, 1),
       _p(a_ref), _p(b_ref), int(threads))
    return Cf[:, 0].copy() if vec else Cf


def _conj_a(flag):
    lib().ora_set_conj_a(int(bool(flag)))


def spmm_csr_T(m, n, rowptr, colval, nzval, B, C=None, alpha=1.0, beta=0.0, conj_a=False):
    """mul!(C, transpose(A::CSR), B, α, β): scatter in storage order
    (matrix_csr_kernels.jl:24-45); also CSC-N on (colptr, rowval).  conj_a:
    adjoint(A), the stored values conjugated (CONJA)."""
    T = nzval.dtype
    B2 = np.asarray(B, dtype=T)
    vec = B2.ndim == 1
    Bf = np.asfortranarray(B2.reshape(-1, 1) if vec else B2)
    k = Bf.shape[1]
    Cf = np.zeros((n, k), dtype=T, order="F") if C is None else np.array(
        np.asarray(C, dtype=T).reshape(n, k), order="F", copy=True)
    fn = getattr(lib(), "ora_spmm_csr_T_" + _sfx(T))
    fn.argtypes = [ctypes.c_int64] * 3 + [_P] * 4 + [ctypes.c_int64, ctypes.c_int, _P,
                                                     ctypes.c_int64, _P, _P]
    rp, cv, nz = _i(rowptr), _i(colval), np.ascontiguousarray(nzval)
    a_ref, b_ref = _sref(alpha, T), _sref(beta, T)
    _conj_a(conj_a)
    try:
        fn(m, n, k, _p(rp), _p(cv), _p(nz), _p(Bf), max(Bf.shape[0], 1), 0, _p(Cf), max(n, 1),
           _p(a_ref), _p(b_ref))
    finally:
        _conj_a(False)
    return Cf[:, 0].copy() if vec else Cf


def spmm_coo(trans, rows_out, rowind, colind, nzval, B, C=None, alpha=1.0, beta=0.0,
             conj_a=False):
    """mul! for COO (matrix_coo_kernels.jl:1-43), entries in storage order;
    conj_a: adjoint(A) (CONJA)."""
    T = nzval.dtype
    B2 = np.asarray(B, dtype=T)
    vec = B2.ndim == 1
    Bf = np.asfortranarray(B2.reshape(-1, 1) if vec else B2)
    k = Bf.shape[1]
    Cf = np.zeros((rows_out, k), dtype=T, order="F") if C is None else np.array(
        np.asarray(C, dtype=T).reshape(rows_out, k), order="F", copy=True)
    fn = getattr(lib(), "ora_spmm_coo_" + _sfx(T))
    fn.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _P, _P, _P, _P,
                   ctypes.c_int64, ctypes.c_int, _P, ctypes.c_int64, _P, _P]
    r, c, v = _i(rowind), _i(colind), np.ascontiguousarray(nzval)
    a_ref, b_ref = _sref(alpha, T), _sref(beta, T)
    _conj_a(conj_a)
    try:
        fn(int(trans), rows_out, k, v.size, _p(r), _p(c), _p(v), _p(Bf), max(Bf.shape[0], 1), 0,
           _p(Cf), max(rows_out, 1), _p(a_ref), _p(b_ref))
    finally:
        _conj_a(False)
    return Cf[:, 0].copy() if vec else Cf


# ---------------------------------------------------------------- conversions
def csr_to_coo(m, rowptr, colval, nzval):
    nnz = colval.size
    r = np.empty(nnz, _I); c = np.empty(nnz, _I); v = np.empty(nnz, nzval.dtype)
    fn = getattr(lib(), "ora_csr_to_coo_" + _sfx(nzval.dtype))
    fn.argtypes = [ctypes.c_int64] + [_P] * 6
    rp, cv, nz = _i(rowptr), _i(colval), np.ascontiguousarray(nzval)
    fn(m, _p(rp), _p(cv), _p(nz), _p(r), _p(c), _p(v))
    return r, c, v


def coo_to_csr(m, n, rowind, colind, nzval):
    """GenericSparseMatrixCSR(::COO) (conversions.jl:282-323)."""
    nnz = nzval.size
    rp = np.empty(m + 1, _I); cv = np.empty(nnz, _I); v = np.empty(nnz, nzval.dtype)
    fn = getattr(lib(), "ora_coo_to_csr_" + _sfx(nzval.dtype))
    fn.argtypes = [ctypes.c_int64] * 3 + [_P] * 6
    r, c, nz = _i(rowind), _i(colind), np.ascontiguousarray(nzval)
    fn(m, n, nnz, _p(r), _p(c), _p(nz), _p(rp), _p(cv), _p(v))
    return rp, cv, v


def coo_to_csc(m, n, rowind, colind, nzval):
    """GenericSparseMatrixCSC(::COO) (conversions.jl:171-212): keys colind*m + rowind."""
    return coo_to_csr(n, m, colind, rowind, nzval)


def csr_to_csc(m, n, rowptr, colval, nzval):
    """GenericSparseMatrixCSC(::CSR) = CSC(COO(A)) (conversions.jl:85-86)."""
    r, c, v = csr_to_coo(m, rowptr, colval, nzval)
    return coo_to_csc(m, n, r, c, v)


def csc_to_csr(m, n, colptr, rowval, nzval):
    """GenericSparseMatrixCSR(::CSC) = CSR(COO(A)) (conversions.jl:104-105); CSC -> COO
    is column-major (conversions.jl:153-169)."""
    c, r, v = csr_to_coo(n, colptr, rowval, nzval)
    return coo_to_csr(m, n, r, c, v)


# ---------------------------------------------------------------- structural
def csr_add(m, A, B):
    """A + B for CSR tuples (rowptr, colval, nzval) (matrix_csr.jl:290-355)."""
    T = A[2].dtype
    fn = getattr(lib(), "ora_csr_add_" + _sfx(T))
    fn.argtypes = [ctypes.c_int64] + [_P] * 9
    fn.restype = ctypes.c_int64
    a = [_i(A[0]), _i(A[1]), np.ascontiguousarray(A[2], T)]
    b = [_i(B[0]), _i(B[1]), np.ascontiguousarray(B[2], T)]
    cnt = fn(m, *map(_p, a), *map(_p, b), _p(None), _p(None), _p(None))
    rp = np.empty(m + 1, _I); cv = np.empty(cnt, _I); v = np.empty(cnt, T)
    fn(m, *map(_p, a), *map(_p, b), _p(rp), _p(cv), _p(v))
    return rp, cv, v


def csr_dropzeros(m, rowptr, colval, nzval):
    T = nzval.dtype
    fn = getattr(lib(), "ora_csr_dropzeros_" + _sfx(T))
    fn.argtypes = [ctypes.c_int64] + [_P] * 6
    fn.restype = ctypes.c_int64
    a = [_i(rowptr), _i(colval), np.ascontiguousarray(nzval)]
    cnt = fn(m, *map(_p, a), _p(None), _p(None), _p(None))
    rp = np.empty(m + 1, _I); cv = np.empty(cnt, _I); v = np.empty(cnt, T)
    fn(m, *map(_p, a), _p(rp), _p(cv), _p(v))
    return rp, cv, v


def coo_add(m, A, B):
    """A + B for COO tuples (rowind, colind, nzval) (matrix_coo.jl:300-384)."""
    T = A[2].dtype
    fn = getattr(lib(), "ora_coo_add_" + _sfx(T))
    fn.argtypes = [ctypes.c_int64, ctypes.c_int64, _P, _P, _P, ctypes.c_int64, _P, _P, _P,
                   _P, _P, _P]
    fn.restype = ctypes.c_int64
    a = [_i(A[0]), _i(A[1]), np.ascontiguousarray(A[2], T)]
    b = [_i(B[0]), _i(B[1]), np.ascontiguousarray(B[2], T)]
    cnt = fn(m, a[2].size, *map(_p, a), b[2].size, *map(_p, b), _p(None), _p(None), _p(None))
    r = np.empty(cnt, _I); c = np.empty(cnt, _I); v = np.empty(cnt, T)
    fn(m, a[2].size, *map(_p, a), b[2].size, *map(_p, b), _p(r), _p(c), _p(v))
    return r, c, v


def coo_dropzeros(rowind, colind, nzval):
    keep = nzval != 0
    return _i(rowind)[keep], _i(colind)[keep], nzval[keep]


def kron_coo(mB, nB, A, B):
    """kron(COO, COO) (matrix_coo_kernels.jl:144-183): A-major / B-minor."""
    T = A[2].dtype
    nA_, nB_ = A[2].size, B[2].size
    r = np.empty(nA_ * nB_, _I); c = np.empty(nA_ * nB_, _I); v = np.empty(nA_ * nB_, T)
    fn = getattr(lib(), "ora_kron_coo_" + _sfx(T))
    fn.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _P, _P, _P, ctypes.c_int64,
                   _P, _P, _P, _P, _P, _P]
    a = [_i(A[0]), _i(A[1]), np.ascontiguousarray(A[2], T)]
    b = [_i(B[0]), _i(B[1]), np.ascontiguousarray(B[2], T)]
    fn(mB, nB, nA_, *map(_p, a), nB_, *map(_p, b), _p(r), _p(c), _p(v))
    return r, c, v


def kron_csr(mA, nA, mB, nB, A, B):
    """kron(CSR, CSR) = CSR(kron(COO(A), COO(B))) (matrix_csr.jl:442-458)."""
    ca = csr_to_coo(mA, *A)
    cb = csr_to_coo(mB, *B)
    r, c, v = kron_coo(mB, nB, ca, cb)
    return coo_to_csr(mA * mB, nA * nB, r, c, v)


def spgemm_csr(m, n, A, B):
    """A * B for CSR tuples (matrix_csr.jl:498-572, kernels :254-343)."""
    T = A[2].dtype
    fn = getattr(lib(), "ora_spgemm_csr_" + _sfx(T))
    fn.argtypes = [ctypes.c_int64, ctypes.c_int64] + [_P] * 9
    fn.restype = ctypes.c_int64
    a = [_i(A[0]), _i(A[1]), np.ascontiguousarray(A[2], T)]
    b = [_i(B[0]), _i(B[1]), np.ascontiguousarray(B[2], T)]
    cnt = fn(m, n, *map(_p, a), *map(_p, b), _p(None), _p(None), _p(None))
    rp = np.empty(m + 1, _I); cv = np.empty(cnt, _I); v = np.empty(cnt, T)
    fn(m, n, *map(_p, a), *map(_p, b), _p(rp), _p(cv), _p(v))
    return rp, cv, v


def sortperm(keys):
    k = _i(keys)
    p = np.empty(k.size, _I)
    fn = lib().ora_sortperm
    fn.argtypes = [ctypes.c_int64, _P, _P]
    fn(k.size, _p(k), _p(p))
    return p + 1


def searchsortedfirst(v, x):
    fn = lib().ora_searchsortedfirst
    fn.argtypes = [_P, ctypes.c_int64, ctypes.c_int64]
    fn.restype = ctypes.c_int64
    a = _i(v)
    return int(fn(_p(a), a.size, int(x)))


# ---------------------------------------------------------------- reductions
def _red(name, T, argtypes, *args):
    fn = getattr(lib(), name + "_" + _sfx(T))
    fn.argtypes = argtypes + [ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_double
    a = ctypes.c_double()
    s = fn(*args, ctypes.byref(a))
    return float(s), float(a.value)


def trace_ptr(n, ptr, idx, val):
    """tr for CSR (rows) or CSC (columns) (matrix_csr.jl:118-142).  Returns
    (sum, sum of |term|)."""
    p, i, v = _i(ptr), _i(idx), np.ascontiguousarray(val)
    return _red("ora_trace_ptr", v.dtype, [ctypes.c_int64, _P, _P, _P], n, _p(p), _p(i), _p(v))


def trace_coo(rowind, colind, val):
    """tr(::COO) (matrix_coo.jl:131-150)."""
    r, c, v = _i(rowind), _i(colind), np.ascontiguousarray(val)
    return _red("ora_trace_coo", v.dtype, [ctypes.c_int64, _P, _P, _P], v.size, _p(r), _p(c),
                _p(v))


def dot3_ptr(nmaj, ptr, idx, val, x, y, major_left):
    """dot(x, A, y) in ptr form (matrix_csr_kernels.jl:47-131).  Returns (sum, Σ|term|)."""
    p, i, v = _i(ptr), _i(idx), np.ascontiguousarray(val)
    xs = np.ascontiguousarray(x, v.dtype)
    ys = np.ascontiguousarray(y, v.dtype)
    return _red("ora_dot3_ptr", v.dtype, [ctypes.c_int64, _P, _P, _P, _P, _P, ctypes.c_int],
                nmaj, _p(p), _p(i), _p(v), _p(xs), _p(ys), int(major_left))


def dot3_coo(rowind, colind, val, x, y, trans):
    """dot(x, op(A::COO), y) (matrix_coo_kernels.jl:45-131)."""
    r, c, v = _i(rowind), _i(colind), np.ascontiguousarray(val)
    xs = np.ascontiguousarray(x, v.dtype)
    ys = np.ascontiguousarray(y, v.dtype)
    return _red("ora_dot3_coo", v.dtype, [ctypes.c_int64, _P, _P, _P, _P, _P, ctypes.c_int],
                v.size, _p(r), _p(c), _p(v), _p(xs), _p(ys), int(trans))


def symmetric_ptr(n, ptr, idx, val) -> bool:
    """issymmetric for a square matrix in ptr form (matrix_csr_kernels.jl:348-407)."""
    p, i, v = _i(ptr), _i(idx), np.ascontiguousarray(val)
    fn = getattr(lib(), "ora_symmetric_ptr_" + _sfx(v.dtype))
    fn.argtypes = [ctypes.c_int64, _P, _P, _P]
    fn.restype = ctypes.c_int
    return bool(fn(n, _p(p), _p(i), _p(v)))


def add_diag(n, ptr, idx, val, lam):
    """CSR/CSC + λI, λ != 0 (matrix_csr.jl:357-406).  Returns (ptr, idx, val)."""
    p, i, v = _i(ptr), _i(idx), np.ascontiguousarray(val)
    T = v.dtype
    fn = getattr(lib(), "ora_add_diag_" + _sfx(T))
    fn.argtypes = [ctypes.c_int64, _P, _P, _P, _P, _P, _P, _P]
    fn.restype = ctypes.c_int64
    lam_a = _sref(lam, T)
    cnt = fn(n, _p(p), _p(i), _p(v), _p(lam_a), _p(None), _p(None), _p(None))
    po = np.empty(n + 1, _I); io = np.empty(cnt, _I); vo = np.empty(cnt, T)
    fn(n, _p(p), _p(i), _p(v), _p(lam_a), _p(po), _p(io), _p(vo))
    return po, io, vo


def diag_coo(n, dvals):
    """The n entries (i, i, d_i) (kernel_fill_diagonal_coo!, matrix_coo_kernels.jl:312-325)."""
    k = np.arange(1, n + 1, dtype=_I)
    return k, k.copy(), np.ascontiguousarray(dvals)


def coo_add_diag(n, A, lam):
    """COO + λI, λ != 0 (matrix_coo.jl:386-457): COO + the diagonal entries."""
    T = A[2].dtype
    return coo_add(n, A, diag_coo(n, np.full(n, lam, dtype=T)))


def kron_diag_coo(mB, nB, d, B):
    """kron(Diagonal(d), B::COO) (kernel_kron_diag_coo!, matrix_coo_kernels.jl:185-221)."""
    return kron_coo(mB, nB, diag_coo(len(d), d), B)


def kron_coo_diag(A, d):
    """kron(A::COO, Diagonal(d)) (kernel_kron_coo_diag!, matrix_coo_kernels.jl:223-258)."""
    p = len(d)
    return kron_coo(p, p, A, diag_coo(p, d))
