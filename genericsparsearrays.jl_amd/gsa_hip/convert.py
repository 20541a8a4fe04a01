"""Format conversions (src/conversions/conversions.jl) over libgsa_hip.so.

Device-to-device conversions are HIP kernels.  Host ingest from a
scipy.sparse matrix (the SparseMatrixCSC interop of conversions.jl:7-79)
copies a CSC or CSR matrix's arrays to the device as they are (one-based) and
transposes there when the other format is asked for; other scipy formats are
made canonical on the host first.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from ._lib import ArgumentError
from .matrix import (GenericSparseMatrixCSR, GenericSparseMatrixCSC, GenericSparseMatrixCOO,
                     Transpose, Adjoint, itype_of, vtype_of, ptr, stream_of, require_device)


def _is_scipy(A) -> bool:
    try:
        import scipy.sparse as sp
        return sp.issparse(A)
    except Exception:  # pragma: no cover
        return False


def _dev(device):
    return torch.device(device if device is not None else "cuda")


def _conj(v: torch.Tensor, conj: bool) -> torch.Tensor:
    return v.conj().resolve_conj() if conj and v.is_complex() else v


def _index_up(a, index_dtype, dev):
    """A zero-based host index array as a one-based device vector of index_dtype:
    copied as it is (Int32 or Int64), then shifted and widened on the device
    (gsa_index_shift), so the host makes no shifted copy."""
    a = np.ascontiguousarray(a)
    if a.dtype not in (np.int32, np.int64):
        a = a.astype(np.int64)
    raw = torch.from_numpy(a).to(dev)
    out = raw if raw.dtype == index_dtype else torch.empty(a.size, dtype=index_dtype, device=dev)
    L.call("gsa_index_shift", stream_of(dev), itype_of(raw), itype_of(out), a.size, ptr(raw),
           ptr(out), 1)
    return out


def _down(t: torch.Tensor) -> torch.Tensor:
    """A device tensor copied into pinned host memory (torch's caching host
    allocator, so repeated egress of like sizes reuses it) on the current stream;
    the caller synchronises before reading it."""
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t, non_blocking=True)
    return h


def _index_down(t: torch.Tensor, small: bool) -> torch.Tensor:
    """A one-based device index vector as a zero-based host vector, Int32 when
    `small` (the width scipy keeps for matrices whose sizes fit it, so it makes no
    narrowing copy of its own): shifted and narrowed on the device, then copied."""
    out = torch.empty(t.numel(), dtype=torch.int32 if small else torch.int64, device=t.device)
    L.call("gsa_index_shift", stream_of(t.device), itype_of(t), itype_of(out), t.numel(), ptr(t),
           ptr(out), -1)
    return _down(out)


def _host(dev, *ts):
    """The host tensors' numpy arrays, once the copies into them are done."""
    torch.cuda.current_stream(dev).synchronize()
    return tuple(t.numpy() for t in ts)


def _small(A) -> bool:
    return max(A.m, A.n, A.nnz() + 1) < 2**31


def _values_up(v, dtype, dev):
    vals = v if dtype is None else v.astype(dtype)
    return torch.from_numpy(np.ascontiguousarray(vals)).to(dev)


def _upload(A, dev, index_dtype, dtype):
    """The compressed arrays of a scipy CSR or CSC matrix as one-based device vectors."""
    return (_index_up(A.indptr, index_dtype, dev), _index_up(A.indices, index_dtype, dev),
            _values_up(A.data, dtype, dev))


# ------------------------------------------------------------------ to CSR --
def csr_from_scipy(A, device=None, index_dtype=torch.int64, dtype=None):
    """GenericSparseMatrixCSR(::SparseMatrixCSC) (conversions.jl:25-30).

    The reference transposes the CSC arrays twice on the host.  A scipy CSC
    input is copied to the device as it is and transposed there (the stable
    CSC -> CSR sort: rows in column order, as the reference's transposition
    leaves them); other inputs are made canonical CSR on the host."""
    dev = _dev(device)
    if A.format == "csc":
        m, n = A.shape
        Ac = GenericSparseMatrixCSC(m, n, *_upload(A, dev, index_dtype, dtype), copy=False)
        return csc_to_csr(Ac)
    B = A.tocsr(copy=True)
    B.sum_duplicates()
    B.sort_indices()
    return GenericSparseMatrixCSR(B.shape[0], B.shape[1], *_upload(B, dev, index_dtype, dtype),
                                  copy=False)


def coo_to_csr(A: GenericSparseMatrixCOO) -> GenericSparseMatrixCSR:
    """GenericSparseMatrixCSR(::COO) (conversions.jl:282-323)."""
    dev = require_device(A.rowind, A.colind, A.nzval)
    nnz = A.nnz()
    rowptr = torch.empty(A.m + 1, dtype=A.rowind.dtype, device=dev)
    colval = torch.empty(nnz, dtype=A.rowind.dtype, device=dev)
    nzval = torch.empty(nnz, dtype=A.nzval.dtype, device=dev)
    L.call("gsa_coo_to_csr", stream_of(dev), vtype_of(A.nzval), itype_of(A.rowind), A.m, A.n,
           nnz, ptr(A.rowind), ptr(A.colind), ptr(A.nzval), ptr(rowptr), ptr(colval), ptr(nzval))
    return GenericSparseMatrixCSR(A.m, A.n, rowptr, colval, nzval, copy=False)


def csr_to_csc(A: GenericSparseMatrixCSR) -> GenericSparseMatrixCSC:
    """GenericSparseMatrixCSC(::CSR) (conversions.jl:85-86) by a stable column sort."""
    dev = require_device(A.rowptr, A.colval, A.nzval)
    nnz = A.nnz()
    colptr = torch.empty(A.n + 1, dtype=A.rowptr.dtype, device=dev)
    rowval = torch.empty(nnz, dtype=A.rowptr.dtype, device=dev)
    nzval = torch.empty(nnz, dtype=A.nzval.dtype, device=dev)
    L.call("gsa_csr_to_csc", stream_of(dev), vtype_of(A.nzval), itype_of(A.rowptr), A.m, A.n,
           nnz, ptr(A.rowptr), ptr(A.colval), ptr(A.nzval), ptr(colptr), ptr(rowval), ptr(nzval))
    return GenericSparseMatrixCSC(A.m, A.n, colptr, rowval, nzval, copy=False)


def csc_to_csr(A: GenericSparseMatrixCSC) -> GenericSparseMatrixCSR:
    """GenericSparseMatrixCSR(::CSC) (conversions.jl:104-105): the same sort on Aᵀ's arrays."""
    dev = require_device(A.colptr, A.rowval, A.nzval)
    nnz = A.nnz()
    rowptr = torch.empty(A.m + 1, dtype=A.colptr.dtype, device=dev)
    colval = torch.empty(nnz, dtype=A.colptr.dtype, device=dev)
    nzval = torch.empty(nnz, dtype=A.nzval.dtype, device=dev)
    L.call("gsa_csr_to_csc", stream_of(dev), vtype_of(A.nzval), itype_of(A.colptr), A.n, A.m,
           nnz, ptr(A.colptr), ptr(A.rowval), ptr(A.nzval), ptr(rowptr), ptr(colval), ptr(nzval))
    return GenericSparseMatrixCSR(A.m, A.n, rowptr, colval, nzval, copy=False)


def to_csr(src, device=None, index_dtype=torch.int64) -> GenericSparseMatrixCSR:
    if isinstance(src, GenericSparseMatrixCSR):
        return src.copy()
    if isinstance(src, GenericSparseMatrixCOO):
        return coo_to_csr(src)
    if isinstance(src, GenericSparseMatrixCSC):
        return csc_to_csr(src)
    if isinstance(src, Transpose):
        p = src.parent
        if isinstance(p, GenericSparseMatrixCSC):  # reinterpretation (conversions.jl:106-123)
            return GenericSparseMatrixCSR(p.n, p.m, p.colptr, p.rowval, _conj(p.nzval, src.conj))
        if isinstance(p, GenericSparseMatrixCOO):
            return coo_to_csr(to_coo(src))
        if isinstance(p, GenericSparseMatrixCSR):
            return csc_to_csr(to_csc(src))
    if _is_scipy(src):
        return csr_from_scipy(src, device=device, index_dtype=index_dtype)
    raise ArgumentError(f"cannot convert {type(src).__name__} to GenericSparseMatrixCSR")


# ------------------------------------------------------------------ to CSC --
def csc_from_scipy(A, device=None, index_dtype=torch.int64, dtype=None):
    """GenericSparseMatrixCSC(::SparseMatrixCSC) (conversions.jl:7-8): the CSC arrays
    as they are.  A scipy CSR input is copied to the device and transposed there."""
    if A.format == "csc":
        return GenericSparseMatrixCSC(A.shape[0], A.shape[1],
                                      *_upload(A, _dev(device), index_dtype, dtype), copy=False)
    if A.format == "csr":
        m, n = A.shape
        Ar = GenericSparseMatrixCSR(m, n, *_upload(A, _dev(device), index_dtype, dtype),
                                    copy=False)
        return csr_to_csc(Ar)
    B = A.tocsc(copy=True)
    B.sum_duplicates()
    B.sort_indices()
    return GenericSparseMatrixCSC(B.shape[0], B.shape[1],
                                  *_upload(B, _dev(device), index_dtype, dtype), copy=False)


def coo_to_csc(A: GenericSparseMatrixCOO) -> GenericSparseMatrixCSC:
    """GenericSparseMatrixCSC(::COO) (conversions.jl:171-212): keys colind*m + rowind."""
    dev = require_device(A.rowind, A.colind, A.nzval)
    nnz = A.nnz()
    colptr = torch.empty(A.n + 1, dtype=A.rowind.dtype, device=dev)
    rowval = torch.empty(nnz, dtype=A.rowind.dtype, device=dev)
    nzval = torch.empty(nnz, dtype=A.nzval.dtype, device=dev)
    L.call("gsa_coo_to_csr", stream_of(dev), vtype_of(A.nzval), itype_of(A.rowind), A.n, A.m,
           nnz, ptr(A.colind), ptr(A.rowind), ptr(A.nzval), ptr(colptr), ptr(rowval), ptr(nzval))
    return GenericSparseMatrixCSC(A.m, A.n, colptr, rowval, nzval, copy=False)


def to_csc(src, device=None, index_dtype=torch.int64) -> GenericSparseMatrixCSC:
    if isinstance(src, GenericSparseMatrixCSC):
        return src.copy()
    if isinstance(src, GenericSparseMatrixCSR):
        return csr_to_csc(src)
    if isinstance(src, GenericSparseMatrixCOO):
        return coo_to_csc(src)
    if isinstance(src, Transpose):
        p = src.parent
        if isinstance(p, GenericSparseMatrixCSR):  # reinterpretation (conversions.jl:87-103)
            return GenericSparseMatrixCSC(p.n, p.m, p.rowptr, p.colval, _conj(p.nzval, src.conj))
        if isinstance(p, GenericSparseMatrixCOO):  # conversions.jl:215-236
            return coo_to_csc(to_coo(src))
        if isinstance(p, GenericSparseMatrixCSC):
            return csr_to_csc(to_csr(src))
    if _is_scipy(src):
        return csc_from_scipy(src, device=device, index_dtype=index_dtype)
    raise ArgumentError(f"cannot convert {type(src).__name__} to GenericSparseMatrixCSC")


# ------------------------------------------------------------------ to COO --
def csr_to_coo(A: GenericSparseMatrixCSR) -> GenericSparseMatrixCOO:
    """GenericSparseMatrixCOO(::CSR) (conversions.jl:264-280): row-major order."""
    dev = require_device(A.rowptr, A.colval, A.nzval)
    nnz = A.nnz()
    r = torch.empty(nnz, dtype=A.rowptr.dtype, device=dev)
    c = torch.empty(nnz, dtype=A.rowptr.dtype, device=dev)
    v = torch.empty(nnz, dtype=A.nzval.dtype, device=dev)
    L.call("gsa_csr_to_coo", stream_of(dev), vtype_of(A.nzval), itype_of(A.rowptr), A.m, nnz,
           ptr(A.rowptr), ptr(A.colval), ptr(A.nzval), ptr(r), ptr(c), ptr(v))
    return GenericSparseMatrixCOO(A.m, A.n, r, c, v, copy=False)


def csc_to_coo(A: GenericSparseMatrixCSC) -> GenericSparseMatrixCOO:
    """GenericSparseMatrixCOO(::CSC) (conversions.jl:153-169): column-major order."""
    dev = require_device(A.colptr, A.rowval, A.nzval)
    nnz = A.nnz()
    r = torch.empty(nnz, dtype=A.colptr.dtype, device=dev)
    c = torch.empty(nnz, dtype=A.colptr.dtype, device=dev)
    v = torch.empty(nnz, dtype=A.nzval.dtype, device=dev)
    L.call("gsa_csr_to_coo", stream_of(dev), vtype_of(A.nzval), itype_of(A.colptr), A.n, nnz,
           ptr(A.colptr), ptr(A.rowval), ptr(A.nzval), ptr(c), ptr(r), ptr(v))
    return GenericSparseMatrixCOO(A.m, A.n, r, c, v, copy=False)


def coo_from_scipy(A, device=None, index_dtype=torch.int64, dtype=None):
    """GenericSparseMatrixCOO(::SparseMatrixCSC) (conversions.jl:52-56): findnz order,
    i.e. column-major."""
    B = A.tocsc(copy=True)
    B.sum_duplicates()
    B.sort_indices()
    C = B.tocoo()
    order = np.lexsort((C.row, C.col))
    dev = _dev(device)
    return GenericSparseMatrixCOO(B.shape[0], B.shape[1], _index_up(C.row[order], index_dtype, dev),
                                  _index_up(C.col[order], index_dtype, dev),
                                  _values_up(C.data[order], dtype, dev), copy=False)


def to_coo(src, device=None, index_dtype=torch.int64) -> GenericSparseMatrixCOO:
    if isinstance(src, GenericSparseMatrixCOO):
        return src  # GenericSparseMatrixCOO(A::COO) = A (conversions.jl:128)
    if isinstance(src, GenericSparseMatrixCSR):
        return csr_to_coo(src)
    if isinstance(src, GenericSparseMatrixCSC):
        return csc_to_coo(src)
    if isinstance(src, Transpose):
        p = src.parent
        base = p if isinstance(p, GenericSparseMatrixCOO) else to_coo(p)
        # conversions.jl:130-147, 238-258, 326-346: swap the index vectors
        return GenericSparseMatrixCOO(p.n, p.m, base.colind, base.rowind,
                                      _conj(base.nzval, src.conj))
    if _is_scipy(src):
        return coo_from_scipy(src, device=device, index_dtype=index_dtype)
    raise ArgumentError(f"cannot convert {type(src).__name__} to GenericSparseMatrixCOO")


# ------------------------------------------------------------ host egress --
def to_scipy(A):
    """The host form of a device matrix as it is stored: CSR -> scipy csr_matrix,
    CSC -> csc_matrix, COO -> coo_matrix (entries in storage order, duplicates
    kept), index arrays zero-based (shifted on the device, then copied), values
    copied.  The collect() of the reference's conversions.jl:10-79 before any
    host-side reordering."""
    import scipy.sparse as sp
    if isinstance(A, GenericSparseMatrixCSR):
        k = _small(A)
        v, i, p = _host(A.device, _down(A.nzval), _index_down(A.colval, k),
                        _index_down(A.rowptr, k))
        return sp.csr_matrix((v, i, p), shape=(A.m, A.n), copy=False)
    if isinstance(A, GenericSparseMatrixCSC):
        k = _small(A)
        v, i, p = _host(A.device, _down(A.nzval), _index_down(A.rowval, k),
                        _index_down(A.colptr, k))
        return sp.csc_matrix((v, i, p), shape=(A.m, A.n), copy=False)
    if isinstance(A, GenericSparseMatrixCOO):
        k = _small(A)
        v, r, c = _host(A.device, _down(A.nzval), _index_down(A.rowind, k),
                        _index_down(A.colind, k))
        return sp.coo_matrix((v, (r, c)), shape=(A.m, A.n), copy=False)
    raise ArgumentError(f"cannot convert {type(A).__name__} to a host sparse matrix")


def sparse_matrix_csc(src):
    """SparseMatrixCSC(A) (conversions.jl:10-79) as a scipy csc_matrix.

    CSC: its arrays.  CSR: transposed on the device (the stable column sort of
    GenericSparseMatrixCSC(::CSR)), then copied; the reference collects the CSR
    arrays and transposes them on the host, with the same result.
    transpose(CSR) / adjoint(CSR): the CSR arrays are the CSC arrays of the
    wrapper (values conjugated for the adjoint), as in the reference.  COO (and
    its wrappers): sparse(rowind, colind, nzval, m, n), duplicates summed
    (scipy's tocsc, which adds them in storage order as Julia's sparse does)."""
    import scipy.sparse as sp
    if isinstance(src, GenericSparseMatrixCSC):
        return to_scipy(src)
    if isinstance(src, GenericSparseMatrixCSR):
        return to_scipy(csr_to_csc(src))
    if isinstance(src, GenericSparseMatrixCOO):
        c = to_scipy(src).tocsc()
        c.sort_indices()
        return c
    if isinstance(src, Transpose):
        p = src.parent
        if isinstance(p, GenericSparseMatrixCSR):
            k = _small(p)
            v, i, q = _host(p.device, _down(_conj(p.nzval, src.conj)), _index_down(p.colval, k),
                            _index_down(p.rowptr, k))
            return sp.csc_matrix((v, i, q), shape=(p.n, p.m), copy=False)
        if isinstance(p, GenericSparseMatrixCSC):  # conversions.jl:17-22: through a CSR
            return sparse_matrix_csc(to_csr(src))
        if isinstance(p, GenericSparseMatrixCOO):
            return sparse_matrix_csc(to_coo(src))
    raise ArgumentError(f"cannot convert {type(src).__name__} to SparseMatrixCSC")

