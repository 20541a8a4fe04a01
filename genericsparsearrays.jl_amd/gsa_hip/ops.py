"""Operators of the sparse kernel layer: mul!, *, +, -, kron, dropzeros, tr, sum,
dot(x, A, y), issymmetric / ishermitian, A ± λI, kron with a Diagonal.

Each function keeps the reference's argument checks, messages and result
types, then calls libgsa_hip.so.  References are to GenericSparseArrays.jl
v0.1.2.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L
from ._lib import ArgumentError, DimensionMismatch
from .matrix import (GenericSparseMatrixCSR, GenericSparseMatrixCSC, GenericSparseMatrixCOO,
                     SPARSE_TYPES, Transpose, Adjoint, UniformScaling, Diagonal, itype_of,
                     vtype_of, ptr, stream_of, require_device, similar_dense, promote_dtype,
                     promote_types, scalar_ref, is_complex_dtype, VTYPES)
from . import convert



def _unwrap(A):
    """(parent, opA) for identity / Transpose / Adjoint (src/core.jl:102-113)."""
    if isinstance(A, Adjoint):
        return A.parent, L.GSA_OP_C
    if isinstance(A, Transpose):
        return A.parent, L.GSA_OP_T
    return A, L.GSA_OP_N


def _dense_operand(B: torch.Tensor):
    """(tensor, k, ld, opB) for a dense vector or matrix in any stride order.  A
    torch conjugate view read in row-major order is Julia's Adjoint of the
    column-major matrix (opB = C, CONJB); other conjugate views are resolved."""
    if B.is_conj():
        if B.dim() == 2 and B.stride(1) == 1 and B.shape[0] > 1 and B.stride(0) >= B.shape[1]:
            base = B.conj()  # the stored values; the kernels conjugate them
            _, k, ld, _ = _dense_operand(base)
            return base, k, ld, L.GSA_OP_C
        B = B.resolve_conj()
    if B.dim() == 1:
        if B.stride(0) != 1:
            B = B.contiguous()
        return B, 1, max(B.shape[0], 1), L.GSA_OP_N
    if B.dim() != 2:
        raise ArgumentError("dense operand must be a vector or a matrix")
    rows, k = B.shape
    if B.stride(0) == 1 and (k == 1 or B.stride(1) >= max(rows, 1)):
        return B, k, max(B.stride(1), rows, 1), L.GSA_OP_N
    if B.stride(1) == 1 and (rows == 1 or B.stride(0) >= max(k, 1)):
        return B, k, max(B.stride(0), k, 1), L.GSA_OP_T
    Bc = B.contiguous()
    return Bc, k, max(k, 1), L.GSA_OP_T


def mul_(C: torch.Tensor, A, B: torch.Tensor, alpha=1.0, beta=0.0):
    """LinearAlgebra.mul!(C, A, B, α, β) for a sparse A (maybe wrapped) and dense B, C.

    Reference: src/matrix_csr/matrix_csr.jl:145-211, src/matrix_csc/matrix_csc.jl:144-213,
    src/matrix_coo/matrix_coo.jl:155-221 (same checks and messages)."""
    Ap, opA = _unwrap(A)
    if not isinstance(Ap, SPARSE_TYPES):
        raise ArgumentError(f"mul!: unsupported sparse operand {type(A).__name__}")
    if not isinstance(B, torch.Tensor) or not isinstance(C, torch.Tensor):
        raise ArgumentError("mul!: B and C must be device tensors")
    mA, nA = (Ap.m, Ap.n) if opA == L.GSA_OP_N else (Ap.n, Ap.m)
    b_rows = B.shape[0]
    b_cols = 1 if B.dim() == 1 else B.shape[1]
    c_rows = C.shape[0]
    c_cols = 1 if C.dim() == 1 else C.shape[1]
    if nA != b_rows:
        raise DimensionMismatch(f"second dimension of A, {nA}, does not match the first "
                                f"dimension of B, {b_rows}")
    if mA != c_rows:
        raise DimensionMismatch(f"first dimension of A, {mA}, does not match the first "
                                f"dimension of C, {c_rows}")
    if b_cols != c_cols:
        raise DimensionMismatch(f"second dimension of B, {b_cols}, does not match the second "
                                f"dimension of C, {c_cols}")
    T3 = C.dtype
    # promote_type(T1, T2, eltype(α), eltype(β)) <: T3 (matrix_csr.jl:176-180)
    bad = T3 not in VTYPES or any(d not in VTYPES or promote_types(d, T3) != T3
                                  for d in (Ap.nzval.dtype, B.dtype))
    try:
        a_ref = scalar_ref(alpha, T3) if not bad else None
        b_ref = scalar_ref(beta, T3) if not bad else None
    except ArgumentError:
        bad = True
    if bad:
        raise ArgumentError("element types of A, B, α, and β must be promotable to the "
                            "element type of C")
    dev = require_device(C, Ap.nzval, B)
    vals = Ap.nzval if Ap.nzval.dtype == T3 else Ap.nzval.to(T3)
    Bt = B if B.dtype == T3 else B.to(T3)
    Bt, k, ldb, opB = _dense_operand(Bt)
    # C must be column-major for the kernels; compute into a temp otherwise
    Cw = C
    if C.is_conj() or (C.dim() == 2 and not (C.stride(0) == 1 and
                                             (c_cols == 1 or C.stride(1) >= max(c_rows, 1)))):
        Cw = similar_dense(c_rows, c_cols, T3, dev)
        if beta != 0:
            Cw.copy_(C)
    ldc = max(c_rows, 1) if Cw.dim() == 1 else max(Cw.stride(1), c_rows, 1)
    s = stream_of(dev)
    vt = vtype_of(Cw)
    # the row bins accumulate real products; other value types take the gather
    # views for every scatter-shaped product (no atomics there either)
    real = T3 in (torch.float32, torch.float64)
    if isinstance(Ap, GenericSparseMatrixCSR):
        it = itype_of(Ap.rowptr)
        plan = None
        if Ap.nnz() > 0:  # transpose(A) * B: the plan's bins (x) or gather view (B) replace
            if opA == L.GSA_OP_N:  # the atomic scatter
                plan = Ap.plan()
            else:
                plan = Ap.plan_transposed() if (k == 1 and real) else Ap.plan_gather()
        L.call("gsa_spmm_csr", s, vt, it, opA, Ap.m, Ap.n, k, ptr(Ap.rowptr), ptr(Ap.colval),
               ptr(vals), Ap.nnz(), ptr(Bt), ldb, opB, ptr(Cw), ldc, ctypes.byref(a_ref),
               ctypes.byref(b_ref), plan)
    elif isinstance(Ap, GenericSparseMatrixCSC):
        it = itype_of(Ap.colptr)
        plan = None
        if Ap.nnz() > 0:  # A * B with several columns: the gather view replaces the scatter
            plan = Ap.plan_gather() if (opA == L.GSA_OP_N and (k > 1 or not real)) else Ap.plan()
        L.call("gsa_spmm_csc", s, vt, it, opA, Ap.m, Ap.n, k, ptr(Ap.colptr), ptr(Ap.rowval),
               ptr(vals), Ap.nnz(), ptr(Bt), ldb, opB, ptr(Cw), ldc, ctypes.byref(a_ref),
               ctypes.byref(b_ref), plan)
    else:
        it = itype_of(Ap.rowind)
        # SpMV (k = 1): the row view or bins for A * x, the transposed bins for
        # transpose(A) * x; other products scatter
        plan = None
        if Ap.nnz() > 0:
            if k == 1 and real:
                plan = Ap.plan() if opA == L.GSA_OP_N else Ap.plan_transposed()
            else:
                plan = Ap.plan_gather(trans=opA != L.GSA_OP_N)
        L.call("gsa_spmm_coo", s, vt, it, opA, Ap.m, Ap.n, k, ptr(Ap.rowind), ptr(Ap.colind),
               ptr(vals), Ap.nnz(), ptr(Bt), ldb, opB, ptr(Cw), ldc, ctypes.byref(a_ref),
               ctypes.byref(b_ref), plan)
    if Cw is not C:
        C.copy_(Cw)
    return C


def _result_dtype(*dts):
    return promote_dtype(*dts)


def times(A, B):
    """Base.:*(A, B): sparse x dense allocates and calls mul!; sparse x sparse is SpGEMM."""
    if isinstance(B, torch.Tensor):
        Ap, opA = _unwrap(A)
        rows = Ap.m if opA == L.GSA_OP_N else Ap.n
        T = _result_dtype(Ap.nzval.dtype, B.dtype)
        C = similar_dense(rows, None if B.dim() == 1 else B.shape[1], T, B.device)
        return mul_(C, A, B, 1.0, 0.0)
    if isinstance(B, (int, float)):
        return scale(A, B)
    if isinstance(B, UniformScaling):  # A * J = A * J.λ (src/core.jl:46)
        return scale(A, B.value)
    return spgemm(A, B)


# ------------------------------------------------------------ scalar ops ----
def scale(A, alpha):
    """α*A / A*α / A/α on the stored values (src/core.jl:26-47)."""
    if isinstance(A, Transpose):
        return type(A)(scale(A.parent, alpha))
    B = A.copy()
    B.nzval.mul_(alpha)
    return B


def negate(A):
    """Base.:-(A) (src/matrix_csr/matrix_csr.jl:88-90); a dense operand (the B
    of A - B, src/core.jl:91-93) is negated as the broadcast it is in Julia."""
    if isinstance(A, UniformScaling):
        return -A
    if isinstance(A, torch.Tensor):
        return -A
    if isinstance(A, Transpose):
        return type(A)(negate(A.parent))
    B = A.copy()
    B.nzval.neg_()
    return B


# ------------------------------------------------------------ dropzeros -----
def dropzeros(A):
    """SparseArrays.dropzeros (src/matrix_csr/matrix_csr.jl:723-745;
    src/matrix_csc/matrix_csc.jl:728-750; src/matrix_coo/matrix_coo.jl:909-935)."""
    if isinstance(A, GenericSparseMatrixCOO):
        dev = require_device(A.rowind, A.colind, A.nzval)
        s = stream_of(dev)
        cnt = ctypes.c_int64()
        L.call("gsa_coo_dropzeros_count", s, vtype_of(A.nzval), A.nnz(), ptr(A.nzval),
               ctypes.byref(cnt))
        if cnt.value == A.nnz():
            return A.copy()
        r = torch.empty(cnt.value, dtype=A.rowind.dtype, device=dev)
        c = torch.empty(cnt.value, dtype=A.rowind.dtype, device=dev)
        v = torch.empty(cnt.value, dtype=A.nzval.dtype, device=dev)
        L.call("gsa_coo_dropzeros_fill", s, vtype_of(A.nzval), itype_of(A.rowind), A.nnz(),
               ptr(A.rowind), ptr(A.colind), ptr(A.nzval), ptr(r), ptr(c), ptr(v))
        return GenericSparseMatrixCOO(A.m, A.n, r, c, v, copy=False)
    if isinstance(A, GenericSparseMatrixCSR):
        major, pt, ix = A.m, A.rowptr, A.colval
    elif isinstance(A, GenericSparseMatrixCSC):
        major, pt, ix = A.n, A.colptr, A.rowval
    else:
        raise ArgumentError(f"dropzeros: unsupported {type(A).__name__}")
    dev = require_device(pt, ix, A.nzval)
    s = stream_of(dev)
    newptr = torch.empty(major + 1, dtype=pt.dtype, device=dev)
    cnt = ctypes.c_int64()
    L.call("gsa_csr_dropzeros_count", s, vtype_of(A.nzval), itype_of(pt), major, ptr(pt),
           ptr(A.nzval), ptr(newptr), ctypes.byref(cnt))
    newix = torch.empty(cnt.value, dtype=pt.dtype, device=dev)
    newv = torch.empty(cnt.value, dtype=A.nzval.dtype, device=dev)
    L.call("gsa_csr_dropzeros_fill", s, vtype_of(A.nzval), itype_of(pt), major, ptr(pt), ptr(ix),
           ptr(A.nzval), ptr(newptr), ptr(newix), ptr(newv))
    return type(A)(A.m, A.n, newptr, newix, newv, copy=False)


def dropzeros_(A):
    """SparseArrays.dropzeros! — keeps stored zeros when every entry is zero
    (src/matrix_csr/matrix_csr.jl:685-721)."""
    B = dropzeros(A)
    if B.nnz() == 0 and A.nnz() > 0:
        return A
    if isinstance(A, GenericSparseMatrixCSR):
        A.rowptr, A.colval = B.rowptr, B.colval
    elif isinstance(A, GenericSparseMatrixCSC):
        A.colptr, A.rowval = B.colptr, B.rowval
    else:
        A.rowind, A.colind = B.rowind, B.colind
    A.nzval = B.nzval
    return A


# ------------------------------------------------------------ sparse + ------
def _csr_like_add(A, B, cls, major, ptrA, ixA, ptrB, ixB):
    dev = require_device(ptrA, ptrB, A.nzval, B.nzval)
    T = _result_dtype(A.nzval.dtype, B.nzval.dtype)
    if ptrA.dtype != ptrB.dtype:
        raise ArgumentError("operands must share the index type")
    vA = A.nzval.to(T)
    vB = B.nzval.to(T)
    s = stream_of(dev)
    vt, it = vtype_of(vA), itype_of(ptrA)
    ptrC = torch.empty(major + 1, dtype=ptrA.dtype, device=dev)
    cnt = ctypes.c_int64()
    nA, nB = min(ixA.numel(), vA.numel()), min(ixB.numel(), vB.numel())
    L.call("gsa_csr_add_count", s, vt, it, major, nA, ptr(ptrA), ptr(ixA), ptr(vA), nB, ptr(ptrB),
           ptr(ixB), ptr(vB), ptr(ptrC), ctypes.byref(cnt))
    ixC = torch.empty(cnt.value, dtype=ptrA.dtype, device=dev)
    vC = torch.empty(cnt.value, dtype=T, device=dev)
    L.call("gsa_csr_add_fill", s, vt, it, major, nA, ptr(ptrA), ptr(ixA), ptr(vA), nB, ptr(ptrB),
           ptr(ixB), ptr(vB), ptr(ptrC), ptr(ixC), ptr(vC))
    return cls(A.m, A.n, ptrC, ixC, vC, copy=False)


def _coo_add(A, B):
    dev = require_device(A.rowind, B.rowind, A.nzval, B.nzval)
    T = _result_dtype(A.nzval.dtype, B.nzval.dtype)
    if A.rowind.dtype != B.rowind.dtype:
        raise ArgumentError("operands must share the index type")
    vA, vB = A.nzval.to(T), B.nzval.to(T)
    s = stream_of(dev)
    ctx = ctypes.c_void_p()
    cnt = ctypes.c_int64()
    L.call("gsa_coo_add_begin", s, vtype_of(vA), itype_of(A.rowind), A.m, A.n, A.nnz(),
           ptr(A.rowind), ptr(A.colind), ptr(vA), B.nnz(), ptr(B.rowind), ptr(B.colind), ptr(vB),
           ctypes.byref(ctx), ctypes.byref(cnt))
    try:
        r = torch.empty(cnt.value, dtype=A.rowind.dtype, device=dev)
        c = torch.empty(cnt.value, dtype=A.rowind.dtype, device=dev)
        v = torch.empty(cnt.value, dtype=T, device=dev)
    except BaseException:
        L.lib().gsa_ctx_destroy(ctx)
        raise
    L.call("gsa_coo_add_finish", ctx, ptr(r), ptr(c), ptr(v))
    return GenericSparseMatrixCOO(A.m, A.n, r, c, v, copy=False)


def add(A, B):
    """Base.:+(A, B) for sparse operands (src/matrix_csr/matrix_csr.jl:290-355,409-440;
    src/matrix_csc/matrix_csc.jl:292-357,411-442; src/matrix_coo/matrix_coo.jl:300-384,461-582).
    A UniformScaling on either side is A + λI (src/core.jl:89-94)."""
    if isinstance(A, UniformScaling):
        A, B = B, A
    if isinstance(B, UniformScaling):
        return add_uniform_scaling(A, B)
    if isinstance(A, torch.Tensor):
        A, B = B, A  # B + A = A + B (src/core.jl:89)
    if isinstance(B, torch.Tensor):
        return add_dense(A, B)
    if tuple(A.shape) != tuple(B.shape):
        raise DimensionMismatch(f"dimensions must match: A has dims {tuple(A.shape)}, "
                                f"B has dims {tuple(B.shape)}")
    Ap, opA = _unwrap(A)
    Bp, opB = _unwrap(B)
    if type(Ap) is not type(Bp):
        raise ArgumentError(f"+: mixed formats {type(Ap).__name__} and {type(Bp).__name__}")
    if isinstance(Ap, GenericSparseMatrixCOO):
        if opA == L.GSA_OP_N and opB == L.GSA_OP_N:
            return _coo_add(Ap, Bp)
        return _coo_add(convert.to_coo(A), convert.to_coo(B))
    if isinstance(Ap, GenericSparseMatrixCSR):
        if opA == L.GSA_OP_N and opB == L.GSA_OP_N:
            return _csr_like_add(Ap, Bp, GenericSparseMatrixCSR, Ap.m, Ap.rowptr, Ap.colval,
                                 Bp.rowptr, Bp.colval)
        # transposed operands: through CSC and back (matrix_csr.jl:409-440)
        return convert.to_csr(add(convert.to_csc(A), convert.to_csc(B)))
    if isinstance(Ap, GenericSparseMatrixCSC):
        if opA == L.GSA_OP_N and opB == L.GSA_OP_N:
            return _csr_like_add(Ap, Bp, GenericSparseMatrixCSC, Ap.n, Ap.colptr, Ap.rowval,
                                 Bp.colptr, Bp.rowval)
        return convert.to_csc(add(convert.to_csr(A), convert.to_csr(B)))
    raise ArgumentError(f"+: unsupported operand {type(A).__name__}")


def add_dense(A, B: torch.Tensor):
    """Base.:+(A, B::DenseMatrix) and B + A (src/core.jl:67-89): a copy of the
    dense B, then every stored entry of the sparse A added into it
    (gsa_add_to_dense).  The result keeps B's element type and layout (column-
    or row-major); A's values are converted to it when that is exact under
    promote_type (a wider A than B is an ArgumentError here).  Wrappers go
    through their parent's arrays: transpose(CSR) reads as a CSC of the
    transpose, and so on; adjoint conjugates the values."""
    if B.dim() != 2 or tuple(A.shape) != tuple(B.shape):
        raise DimensionMismatch(f"dimensions must match: A has dims {tuple(A.shape)}, "
                                f"B has dims {tuple(B.shape)}")
    Ap, opA = _unwrap(A)
    vals = Ap.nzval
    if vals.device != B.device:
        raise ArgumentError("Both arrays must have the same backend")
    dev = require_device(vals, B)
    if vals.dtype != B.dtype:
        if promote_dtype(vals.dtype, B.dtype) != B.dtype:
            raise ArgumentError(f"+: the sparse values ({vals.dtype}) do not convert exactly "
                                f"to the dense element type ({B.dtype})")
        vals = vals.to(B.dtype)
    C = B.resolve_conj() if B.is_conj() else B.clone()  # copy(B), same strides
    m, n = B.shape
    if m == 0 or n == 0 or Ap.nnz() == 0:
        return C
    if isinstance(Ap, GenericSparseMatrixCOO):
        fmt, a, b = L.GSA_FMT_COO, Ap.rowind, Ap.colind
        if opA != L.GSA_OP_N:
            a, b = b, a
    elif isinstance(Ap, GenericSparseMatrixCSR):
        fmt, a, b = (L.GSA_FMT_CSR if opA == L.GSA_OP_N else L.GSA_FMT_CSC), Ap.rowptr, Ap.colval
    else:
        fmt, a, b = (L.GSA_FMT_CSC if opA == L.GSA_OP_N else L.GSA_FMT_CSR), Ap.colptr, Ap.rowval
    if C.stride(0) == 1 or m == 1:      # column-major: element (i, j) at i + j*ld
        ld = C.stride(1) if n > 1 else max(m, 1)
    elif C.stride(1) == 1:              # row-major: the column-major transpose
        ld = C.stride(0)
        m, n = n, m
        if fmt == L.GSA_FMT_COO:
            a, b = b, a
        else:
            fmt = L.GSA_FMT_CSC if fmt == L.GSA_FMT_CSR else L.GSA_FMT_CSR
    else:
        C = B.t().contiguous().t()  # a strided B: work on a column-major copy
        ld = C.stride(1)
    L.call("gsa_add_to_dense", stream_of(dev), vtype_of(vals), itype_of(a), fmt,
           1 if opA == L.GSA_OP_C else 0, m, n, Ap.nnz(), ptr(a), ptr(b), ptr(vals), ptr(C),
           max(ld, m, 1))
    return C


# ------------------------------------------------------------ kron ---------
def kron(A, B):
    """LinearAlgebra.kron for sparse operands (src/matrix_coo/matrix_coo.jl:641-699;
    CSR/CSC route through COO, src/matrix_csr/matrix_csr.jl:442-458), and with a
    Diagonal on either side (src/matrix_coo/matrix_coo.jl:704-809,
    src/matrix_csr/matrix_csr.jl:460-496, src/matrix_csc/matrix_csc.jl:462-498)."""
    if isinstance(A, Diagonal) or isinstance(B, Diagonal):
        return _kron_diagonal(A, B)
    Ap, opA = _unwrap(A)
    Bp, opB = _unwrap(B)
    if type(Ap) is not type(Bp):
        raise ArgumentError(f"kron: mixed formats {type(Ap).__name__} and {type(Bp).__name__}")
    if isinstance(Ap, GenericSparseMatrixCOO):
        if opA != L.GSA_OP_N or opB != L.GSA_OP_N:
            return kron(convert.to_coo(A), convert.to_coo(B))
        return _kron_coo(Ap, Bp)
    if isinstance(Ap, GenericSparseMatrixCSR):
        if opA != L.GSA_OP_N or opB != L.GSA_OP_N:
            return convert.coo_to_csr(_kron_coo(convert.to_coo(A), convert.to_coo(B)))
        return _kron_ptr(Ap, Bp, GenericSparseMatrixCSR)
    if isinstance(Ap, GenericSparseMatrixCSC):
        if opA != L.GSA_OP_N or opB != L.GSA_OP_N:
            return convert.coo_to_csc(_kron_coo(convert.to_coo(A), convert.to_coo(B)))
        return _kron_ptr(Ap, Bp, GenericSparseMatrixCSC)
    raise ArgumentError(f"kron: unsupported operand {type(A).__name__}")


def _kron_coo(A, B):
    dev = require_device(A.rowind, B.rowind, A.nzval, B.nzval)
    T = _result_dtype(A.nzval.dtype, B.nzval.dtype)
    it = torch.int64 if torch.int64 in (A.rowind.dtype, B.rowind.dtype) else torch.int32
    mC, nC = A.m * B.m, A.n * B.n
    if it == torch.int32 and (mC >= 2 ** 31 or nC >= 2 ** 31):
        raise ArgumentError("kron result dimensions do not fit Int32 indices")
    rA, cA, rB, cB = (x.to(it) for x in (A.rowind, A.colind, B.rowind, B.colind))
    vA, vB = A.nzval.to(T), B.nzval.to(T)
    nnzC = A.nnz() * B.nnz()
    r = torch.empty(nnzC, dtype=it, device=dev)
    c = torch.empty(nnzC, dtype=it, device=dev)
    v = torch.empty(nnzC, dtype=T, device=dev)
    L.call("gsa_kron_coo", stream_of(dev), vtype_of(vA), itype_of(r), B.m, B.n, A.nnz(), ptr(rA),
           ptr(cA), ptr(vA), B.nnz(), ptr(rB), ptr(cB), ptr(vB), ptr(r), ptr(c), ptr(v))
    return GenericSparseMatrixCOO(mC, nC, r, c, v, copy=False)


def _kron_ptr(A, B, cls):
    csr = cls is GenericSparseMatrixCSR
    pA, iA = (A.rowptr, A.colval) if csr else (A.colptr, A.rowval)
    pB, iB = (B.rowptr, B.colval) if csr else (B.colptr, B.rowval)
    dev = require_device(pA, pB, A.nzval, B.nzval)
    T = _result_dtype(A.nzval.dtype, B.nzval.dtype)
    it = torch.int64 if torch.int64 in (pA.dtype, pB.dtype) else torch.int32
    mC, nC = A.m * B.m, A.n * B.n
    if it == torch.int32 and (mC >= 2 ** 31 or nC >= 2 ** 31 or A.nnz() * B.nnz() >= 2 ** 31):
        raise ArgumentError("kron result does not fit Int32 indices")
    pA, iA, pB, iB = (x.to(it) for x in (pA, iA, pB, iB))
    vA, vB = A.nzval.to(T), B.nzval.to(T)
    major = mC if csr else nC
    nnzC = A.nnz() * B.nnz()
    pC = torch.empty(major + 1, dtype=it, device=dev)
    iC = torch.empty(nnzC, dtype=it, device=dev)
    vC = torch.empty(nnzC, dtype=T, device=dev)
    if csr:
        dims = (A.m, A.n, B.m, B.n)
    else:  # CSC arrays are the CSR arrays of the transpose
        dims = (A.n, A.m, B.n, B.m)
    L.call("gsa_kron_csr", stream_of(dev), vtype_of(vA), itype_of(pC), *dims, ptr(pA), ptr(iA),
           ptr(vA), A.nnz(), ptr(pB), ptr(iB), ptr(vB), B.nnz(), ptr(pC), ptr(iC), ptr(vC))
    return cls(mC, nC, pC, iC, vC, copy=False)


# ------------------------------------------------------------ SpGEMM -------
def _spgemm_arrays(dev, T, m, kdim, n, pA, iA, vA, nnzA, pB, iB, vB, nnzB):
    s = stream_of(dev)
    ctx = ctypes.c_void_p()
    cnt = ctypes.c_int64()
    L.call("gsa_spgemm_csr_begin", s, vtype_of(vA), itype_of(pA), m, kdim, n, ptr(pA), ptr(iA),
           ptr(vA), nnzA, ptr(pB), ptr(iB), ptr(vB), nnzB, ctypes.byref(ctx), ctypes.byref(cnt))
    try:
        pC = torch.empty(m + 1, dtype=pA.dtype, device=dev)
        iC = torch.empty(cnt.value, dtype=pA.dtype, device=dev)
        vC = torch.empty(cnt.value, dtype=T, device=dev)
    except BaseException:
        L.lib().gsa_ctx_destroy(ctx)
        raise
    L.call("gsa_spgemm_csr_finish", ctx, ptr(pC), ptr(iC), ptr(vC))
    return pC, iC, vC


def spgemm(A, B):
    """Base.:*(A, B) for sparse operands (src/matrix_csr/matrix_csr.jl:498-609;
    src/matrix_csc/matrix_csc.jl:500-611; src/matrix_coo/matrix_coo.jl:584-639)."""
    if A.size(2) != B.size(1):
        raise DimensionMismatch(f"second dimension of A, {A.size(2)}, does not match first "
                                f"dimension of B, {B.size(1)}")
    Ap, opA = _unwrap(A)
    Bp, opB = _unwrap(B)
    if type(Ap) is not type(Bp):
        raise ArgumentError(f"*: mixed formats {type(Ap).__name__} and {type(Bp).__name__}")
    if isinstance(Ap, GenericSparseMatrixCOO):
        return convert.to_coo(spgemm(convert.to_csc(A), convert.to_csc(B)))
    if isinstance(Ap, GenericSparseMatrixCSR):
        if opA != L.GSA_OP_N or opB != L.GSA_OP_N:
            return convert.to_csr(spgemm(convert.to_csc(A), convert.to_csc(B)))
        dev = require_device(Ap.rowptr, Bp.rowptr, Ap.nzval, Bp.nzval)
        if Ap.rowptr.dtype != Bp.rowptr.dtype:
            raise ArgumentError("operands must share the index type")
        T = _result_dtype(Ap.nzval.dtype, Bp.nzval.dtype)
        pC, iC, vC = _spgemm_arrays(dev, T, Ap.m, Ap.n, Bp.n, Ap.rowptr, Ap.colval,
                                    Ap.nzval.to(T), Ap.nnz(), Bp.rowptr, Bp.colval,
                                    Bp.nzval.to(T), Bp.nnz())
        return GenericSparseMatrixCSR(Ap.m, Bp.n, pC, iC, vC, copy=False)
    if isinstance(Ap, GenericSparseMatrixCSC):
        if opA != L.GSA_OP_N or opB != L.GSA_OP_N:
            return convert.to_csc(spgemm(convert.to_csr(A), convert.to_csr(B)))
        dev = require_device(Ap.colptr, Bp.colptr, Ap.nzval, Bp.nzval)
        if Ap.colptr.dtype != Bp.colptr.dtype:
            raise ArgumentError("operands must share the index type")
        T = _result_dtype(Ap.nzval.dtype, Bp.nzval.dtype)
        # Cᵀ = Bᵀ Aᵀ: the CSC arrays of B and A are CSR arrays of Bᵀ and Aᵀ.
        pC, iC, vC = _spgemm_arrays(dev, T, Bp.n, Bp.m, Ap.m, Bp.colptr, Bp.rowval,
                                    Bp.nzval.to(T), Bp.nnz(), Ap.colptr, Ap.rowval,
                                    Ap.nzval.to(T), Ap.nnz())
        return GenericSparseMatrixCSC(Ap.m, Bp.n, pC, iC, vC, copy=False)
    raise ArgumentError(f"*: unsupported operand {type(A).__name__}")


# ------------------------------------------------------------ reductions ----
def _acc_buffer():
    """Room for the accumulator a reduction returns: a Float64, a ComplexF64 or
    an Int64 (16 bytes covers each)."""
    return (ctypes.c_byte * 16)()


def _acc_value(buf, dtype, widen_int=False):
    """The accumulator in `buf` as a host scalar of the element type `dtype`
    (Float32 rounded once; integers wrap to their width, or stay Int64 when
    `widen_int`, as Julia's sum of small integers does)."""
    raw = bytes(buf)
    if dtype in (torch.complex64, torch.complex128):
        re, im = np.frombuffer(raw, dtype=np.float64, count=2)
        if dtype == torch.complex64:
            return complex(np.complex64(complex(float(re), float(im))))
        return complex(float(re), float(im))
    if dtype in (torch.int32, torch.int64):
        v = int(np.frombuffer(raw, dtype=np.int64, count=1)[0])
        if dtype == torch.int32 and not widen_int:
            v = int(np.int64(v).astype(np.int32))
        return v
    v = float(np.frombuffer(raw, dtype=np.float64, count=1)[0])
    return float(np.float32(v)) if dtype == torch.float32 else v


def _as_scalar(v: float, dtype):
    """A host Float64 in the element type (a Float32 result is rounded once)."""
    return float(np.float32(v)) if dtype == torch.float32 else float(v)


def _ptr_form(A):
    """(major count, index count, ptr, idx) of a CSR (rows) or CSC (columns)."""
    if isinstance(A, GenericSparseMatrixCSR):
        return A.m, A.n, A.rowptr, A.colval
    return A.n, A.m, A.colptr, A.rowval


def tr(A):
    """LinearAlgebra.tr (src/matrix_csr/matrix_csr.jl:118-142,
    src/matrix_csc/matrix_csc.jl:117-140, src/matrix_coo/matrix_coo.jl:131-150);
    tr(transpose(A)) = tr(A), tr(adjoint(A)) = conj(tr(A))."""
    Ap, opA = _unwrap(A)
    if not isinstance(Ap, SPARSE_TYPES):
        raise ArgumentError(f"tr: unsupported operand {type(A).__name__}")
    if Ap.m != Ap.n:
        raise DimensionMismatch("Matrix must be square to compute the trace.")
    dev = require_device(Ap._idx(), Ap.nzval)
    out = _acc_buffer()
    if isinstance(Ap, GenericSparseMatrixCOO):
        L.call("gsa_trace_coo", stream_of(dev), vtype_of(Ap.nzval), itype_of(Ap.rowind), Ap.nnz(),
               ptr(Ap.rowind), ptr(Ap.colind), ptr(Ap.nzval), out)
    else:
        major, _, pt, ix = _ptr_form(Ap)
        L.call("gsa_trace_ptr", stream_of(dev), vtype_of(Ap.nzval), itype_of(pt), major, Ap.nnz(),
               ptr(pt), ptr(ix), ptr(Ap.nzval), Ap.cached_plan(), out)
    v = _acc_value(out, Ap.nzval.dtype)
    return v.conjugate() if opA == L.GSA_OP_C and isinstance(v, complex) else v


def sum_(A):
    """Base.sum(A) = sum(nonzeros(A)) (src/core.jl:24); Int32 values sum to an
    Int64, as Julia's sum of small integers does; sum(adjoint(A)) is conjugated."""
    Ap, opA = _unwrap(A)
    if not isinstance(Ap, SPARSE_TYPES):
        raise ArgumentError(f"sum: unsupported operand {type(A).__name__}")
    dev = require_device(Ap.nzval)
    out = _acc_buffer()
    L.call("gsa_sum_values", stream_of(dev), vtype_of(Ap.nzval), Ap.nnz(), ptr(Ap.nzval), out)
    v = _acc_value(out, Ap.nzval.dtype, widen_int=True)
    return v.conjugate() if opA == L.GSA_OP_C and isinstance(v, complex) else v


def dot(x: torch.Tensor, A, y: torch.Tensor):
    """LinearAlgebra.dot(x, A, y) = x' A y over the stored entries
    (src/matrix_csr/matrix_csr.jl:214-277, src/matrix_csc/matrix_csc.jl:216-279,
    src/matrix_coo/matrix_coo.jl:224-285: same checks and messages)."""
    Ap, opA = _unwrap(A)
    if not isinstance(Ap, SPARSE_TYPES):
        raise ArgumentError(f"dot: unsupported operand {type(A).__name__}")
    if not isinstance(x, torch.Tensor) or not isinstance(y, torch.Tensor) or \
            x.dim() != 1 or y.dim() != 1:
        raise ArgumentError("dot(x, A, y): x and y must be device vectors")
    mA, nA = (Ap.m, Ap.n) if opA == L.GSA_OP_N else (Ap.n, Ap.m)
    if mA != x.shape[0]:
        raise DimensionMismatch(f"first dimension of A, {mA}, does not match the length of x, "
                                f"{x.shape[0]}")
    if nA != y.shape[0]:
        raise DimensionMismatch(f"second dimension of A, {nA}, does not match the length of y, "
                                f"{y.shape[0]}")
    dev = require_device(x, Ap.nzval, y)
    if any(d not in VTYPES for d in (Ap.nzval.dtype, x.dtype, y.dtype)):
        raise ArgumentError("dot(x, A, y): unsupported element type")
    T = promote_dtype(Ap.nzval.dtype, x.dtype, y.dtype)  # promote_type(T1, T2, T3)
    vals = Ap.nzval if Ap.nzval.dtype == T else Ap.nzval.to(T)
    xt = x.resolve_conj().to(T).contiguous()
    yt = y.resolve_conj().to(T).contiguous()
    out = _acc_buffer()
    s = stream_of(dev)
    conj_a = int(opA == L.GSA_OP_C)
    if isinstance(Ap, GenericSparseMatrixCOO):
        L.call("gsa_dot3_coo", s, vtype_of(vals), itype_of(Ap.rowind), Ap.m, Ap.n, Ap.nnz(),
               ptr(Ap.rowind), ptr(Ap.colind), ptr(vals), ptr(xt), ptr(yt),
               int(opA != L.GSA_OP_N), conj_a, out)
    else:
        major, minor, pt, ix = _ptr_form(Ap)
        # x pairs with the row: the major of CSR, the index of CSC (swapped by a wrapper)
        major_left = isinstance(Ap, GenericSparseMatrixCSR) == (opA == L.GSA_OP_N)
        L.call("gsa_dot3_ptr", s, vtype_of(vals), itype_of(pt), major, minor, Ap.nnz(), ptr(pt),
               ptr(ix), ptr(vals), ptr(xt), ptr(yt), int(major_left), conj_a, Ap.cached_plan(),
               out)
    return _acc_value(out, T)


def _symmetric(A, hermitian: bool) -> bool:
    Ap, _ = _unwrap(A)
    if not isinstance(Ap, SPARSE_TYPES):
        raise ArgumentError(f"issymmetric: unsupported operand {type(A).__name__}")
    if Ap.m != Ap.n:
        return False
    if isinstance(Ap, GenericSparseMatrixCOO):
        Ap = convert.coo_to_csc(Ap)
    if Ap.nnz() == 0:
        return True
    major, _, pt, ix = _ptr_form(Ap)
    dev = require_device(pt, ix, Ap.nzval)
    out = ctypes.c_int()
    L.call("gsa_symmetric_ptr", stream_of(dev), vtype_of(Ap.nzval), itype_of(pt), major, Ap.nnz(),
           ptr(pt), ptr(ix), ptr(Ap.nzval), int(hermitian), Ap.cached_plan(), ctypes.byref(out))
    return bool(out.value)


def issymmetric(A) -> bool:
    """LinearAlgebra.issymmetric (src/matrix_csr/matrix_csr.jl:611-628,
    src/matrix_csc/matrix_csc.jl:616-633; COO through CSC, src/matrix_coo/matrix_coo.jl:811-815;
    wrappers test the parent, src/core.jl:96-99)."""
    return _symmetric(A, False)


def ishermitian(A) -> bool:
    """LinearAlgebra.ishermitian (src/matrix_csr/matrix_csr.jl:630-647): every
    off-diagonal value the conjugate of its mirror, every diagonal value its own
    conjugate (kernel_check_symmetry_csr! with HERMITIAN, matrix_csr_kernels.jl:348-407)."""
    return _symmetric(A, True)


# ------------------------------------------------------------ A + λI --------
def _materialise(A):
    """A wrapped matrix as a plain one of its parent's format."""
    if not isinstance(A, Transpose):
        return A
    p = A.parent
    if isinstance(p, GenericSparseMatrixCSR):
        return convert.to_csr(A)
    if isinstance(p, GenericSparseMatrixCSC):
        return convert.to_csc(A)
    return convert.to_coo(A)


def _diag_coo(n: int, dev, vdtype, idtype, lam: float = 0.0, dvals=None):
    """The n diagonal entries (i, i, d_i) as a COO matrix (kernel_fill_diagonal_coo!,
    src/matrix_coo/matrix_coo_kernels.jl:312-325)."""
    r = torch.empty(n, dtype=idtype, device=dev)
    c = torch.empty(n, dtype=idtype, device=dev)
    v = torch.empty(n, dtype=vdtype, device=dev)
    src = None
    if dvals is not None:
        src = dvals if dvals.dtype == vdtype else dvals.to(vdtype)
        src = src.contiguous()
    lam_ref = scalar_ref(lam, vdtype)
    L.call("gsa_diag_coo", stream_of(dev), vtype_of(v), itype_of(r), n, ptr(src),
           ctypes.byref(lam_ref), ptr(r), ptr(c), ptr(v))
    return GenericSparseMatrixCOO(n, n, r, c, v, copy=False)


def add_uniform_scaling(A, J: UniformScaling):
    """Base.:+(A, λI) (src/matrix_csr/matrix_csr.jl:357-406, src/matrix_csc/matrix_csc.jl:359-408,
    src/matrix_coo/matrix_coo.jl:386-457): λ == 0 returns a copy; otherwise the
    element type is promote_type(Tv, typeof(λ))."""
    A = _materialise(A)
    if not isinstance(A, SPARSE_TYPES):
        raise ArgumentError(f"+: unsupported operand {type(A).__name__}")
    if A.m != A.n:
        raise DimensionMismatch("Matrix must be square to add UniformScaling.")
    lam = J.value
    if lam == 0:
        return A.copy()
    T = promote_dtype(A.nzval.dtype, J.dtype)
    dev = require_device(A._idx(), A.nzval)
    if isinstance(A, GenericSparseMatrixCOO):
        # concatenate the n entries λ, sum duplicates in column-major order, drop zeros
        D = _diag_coo(A.n, dev, T, A.rowind.dtype, lam=lam)
        return _coo_add(A, D)
    major, _, pt, ix = _ptr_form(A)
    vals = A.nzval if A.nzval.dtype == T else A.nzval.to(T)
    s = stream_of(dev)
    newptr = torch.empty(major + 1, dtype=pt.dtype, device=dev)
    cnt = ctypes.c_int64()
    plan = A.cached_plan()
    L.call("gsa_add_diag_count", s, itype_of(pt), major, A.nnz(), ptr(pt), ptr(ix), plan,
           ptr(newptr), ctypes.byref(cnt))
    newix = torch.empty(cnt.value, dtype=pt.dtype, device=dev)
    newv = torch.empty(cnt.value, dtype=T, device=dev)
    lam_ref = scalar_ref(lam, T)
    L.call("gsa_add_diag_fill", s, vtype_of(vals), itype_of(pt), major, A.nnz(), ptr(pt), ptr(ix),
           ptr(vals), ctypes.byref(lam_ref), plan, ptr(newptr), ptr(newix), ptr(newv))
    return type(A)(A.m, A.n, newptr, newix, newv, copy=False)


# ------------------------------------------------------------ kron(Diagonal) -
def _kron_diagonal(A, B):
    """kron(D, B) and kron(A, D): kron of COO operands with D's n entries (i, i, d_i),
    the entry order of kernel_kron_diag_coo! / kernel_kron_coo_diag!
    (src/matrix_coo/matrix_coo_kernels.jl:185-258; for Diagonal(Fill(v, n)),
    kernel_kron_scaled_identity_coo! / kernel_kron_coo_scaled_identity!,
    ext/GenericSparseArraysFillArraysExt/kernels.jl:4-61, which list the same
    entries); a CSR or CSC operand gives a result of its format by conversion
    from that COO (src/matrix_csr/matrix_csr.jl:460-470, ext/.../kron_csr.jl:5-46)."""
    if isinstance(A, Diagonal) and isinstance(B, Diagonal):
        raise ArgumentError("kron of two Diagonals is not a sparse operation here")
    D, S = (A, B) if isinstance(A, Diagonal) else (B, A)
    Sp, _ = _unwrap(S)
    if not isinstance(Sp, SPARSE_TYPES):
        raise ArgumentError(f"kron: unsupported operand {type(S).__name__}")
    Scoo = convert.to_coo(S)
    dev = require_device(Scoo.rowind, Scoo.nzval, D.diag)
    if D.dtype not in VTYPES:
        raise ArgumentError(f"kron: unsupported Diagonal element type {D.dtype}")
    if D.diag is None:  # Diagonal(Fill(v, n)): kernel_kron_scaled_identity_coo!, same entries
        if D.device != dev:
            raise ArgumentError("All arrays must have the same backend")
        Dcoo = _diag_coo(D.shape[0], dev, D.dtype, Scoo.rowind.dtype, lam=D.fill_value)
    else:
        Dcoo = _diag_coo(D.shape[0], dev, D.dtype, Scoo.rowind.dtype, dvals=D.diag)
    C = _kron_coo(Dcoo, Scoo) if D is A else _kron_coo(Scoo, Dcoo)
    if isinstance(Sp, GenericSparseMatrixCSR):
        return convert.coo_to_csr(C)
    if isinstance(Sp, GenericSparseMatrixCSC):
        return convert.coo_to_csc(C)
    return C
