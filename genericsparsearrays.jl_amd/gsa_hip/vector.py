"""GenericSparseVector on HIP device storage (src/vector/vector.jl).

A sparse vector of logical length n with one-based stored indices `nzind` and
values `nzval` (torch tensors on one device).  The constructor keeps the
reference's checks and copies its buffers (vector.jl:16-37).  The reductions
(dot, sum, norm) run in libgsa_hip.so in a fixed order, without atomics; the
element-wise value maps (scaling, negation, conj, real, imag) are maps over
`nzval`, as the reference's broadcasts are.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L
from ._lib import ArgumentError, DimensionMismatch
from .matrix import (VTYPES, ITYPES, itype_of, vtype_of, ptr, stream_of, require_device,
                     promote_dtype, scalar_dtype, is_complex_dtype)


class GenericSparseVector:
    """GenericSparseVector(n, nzind, nzval) (src/vector/vector.jl:16-37)."""

    def __init__(self, n: int, nzind: torch.Tensor, nzval: torch.Tensor, copy: bool = True):
        if nzind.device != nzval.device:
            raise ArgumentError("Index and value vectors must be on the same device/backend.")
        if n < 0:
            raise ArgumentError("The number of elements must be non-negative.")
        if nzind.numel() != nzval.numel():
            raise ArgumentError("index and value vectors must be the same length")
        if nzind.dtype not in ITYPES:
            raise ArgumentError(f"index type {nzind.dtype} is not provided (Int32 and Int64 are)")
        self.n = int(n)
        self.nzind = nzind.clone() if copy else nzind
        self.nzval = nzval.clone() if copy else nzval

    # ---- conversions (vector.jl:39-48) -------------------------------------------
    @classmethod
    def from_arrays(cls, n, nzind, nzval, device="cuda", index_dtype=torch.int64):
        """From host arrays of one-based indices and values (SparseVector's fields)."""
        ind = torch.as_tensor(np.asarray(nzind)).to(device=device, dtype=index_dtype)
        val = torch.as_tensor(np.asarray(nzval)).to(device)
        return cls(n, ind, val, copy=False)

    @classmethod
    def from_scipy(cls, v, device="cuda", index_dtype=torch.int64):
        """GenericSparseVector(::SparseVector): a 1-D scipy sparse array (or a
        single-row / single-column sparse matrix), its entries in index order."""
        import scipy.sparse as sp
        c = sp.coo_array(v)
        if c.ndim == 2:
            if 1 not in c.shape:
                raise ArgumentError("a sparse vector needs one dimension of length 1")
            n = max(c.shape)
            idx = c.coords[1] if c.shape[0] == 1 else c.coords[0]
        else:
            n = c.shape[0]
            idx = c.coords[0]
        order = np.argsort(idx, kind="stable")
        return cls.from_arrays(n, idx[order].astype(np.int64) + 1, c.data[order], device,
                               index_dtype)

    def to_scipy(self):
        """SparseVector(::GenericSparseVector): a 1-D scipy coo_array."""
        import scipy.sparse as sp
        idx = self.nzind.cpu().numpy().astype(np.int64) - 1
        return sp.coo_array((self.nzval.cpu().numpy(), (idx,)), shape=(self.n,))

    def adapt(self, device):
        """Adapt.adapt_structure (vector.jl:44-48): the same vector on `device`."""
        return GenericSparseVector(self.n, self.nzind.to(device), self.nzval.to(device))

    def collect(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=self.nzval.cpu().numpy().dtype)
        idx = self.nzind.cpu().numpy().astype(np.int64) - 1
        np.add.at(out, idx, self.nzval.cpu().numpy())
        return out

    # ---- basic methods (vector.jl:50-62) ---------------------------------------------
    def __len__(self):
        return self.n

    @property
    def shape(self):
        return (self.n,)

    def size(self, dim=None):
        return self.shape if dim is None else self.shape[dim - 1]

    def nnz(self) -> int:
        return int(self.nzval.numel())

    def nonzeros(self) -> torch.Tensor:
        return self.nzval

    def nonzeroinds(self) -> torch.Tensor:
        return self.nzind

    @property
    def dtype(self):
        return self.nzval.dtype

    @property
    def device(self):
        return self.nzval.device

    def copy(self):
        return GenericSparseVector(self.n, self.nzind, self.nzval)

    def copyto_(self, src: "GenericSparseVector"):
        """copyto!(dest, src) (vector.jl:56-61): same length, storage resized."""
        if self.n < src.n:
            raise IndexError("BoundsError: destination shorter than the source")
        if self.n != src.n:
            raise ArgumentError("Cannot copy to a sparse vector of different length")
        self.nzind = src.nzind.to(self.nzind.dtype).clone()
        self.nzval = src.nzval.to(self.nzval.dtype).clone()
        return self

    def similar(self, dtype=None):
        """similar(V, Tv2) (vector.jl:63-64): same indices, uninitialised values."""
        return GenericSparseVector(self.n, self.nzind.clone(),
                                   torch.empty_like(self.nzval, dtype=dtype or self.dtype),
                                   copy=False)

    # ---- scalar operations (vector.jl:66-74) -----------------------------------------
    def _scaled(self, alpha):
        T = promote_dtype(self.dtype, scalar_dtype(alpha)) or self.dtype
        a = alpha.item() if isinstance(alpha, torch.Tensor) else alpha
        vals = self.nzval.to(T) * a
        if vals.dtype != T:  # torch keeps integer tensors integral only for integral scalars
            vals = vals.to(T)
        return GenericSparseVector(self.n, self.nzind.clone(), vals, copy=False)

    def __mul__(self, alpha):
        return self._scaled(alpha)

    __rmul__ = __mul__

    def __truediv__(self, alpha):
        a = alpha.item() if isinstance(alpha, torch.Tensor) else alpha
        if self.dtype in (torch.int32, torch.int64) and isinstance(a, (int, np.integer)):
            return GenericSparseVector(self.n, self.nzind.clone(),
                                       self.nzval.to(torch.float64) / a, copy=False)
        return self._scaled(1 / a if not isinstance(a, np.floating) else a.dtype.type(1) / a)

    def __neg__(self):
        return GenericSparseVector(self.n, self.nzind.clone(), -self.nzval, copy=False)

    def __pos__(self):
        return self.copy()

    def conj(self):
        """conj(V) (vector.jl:76-80): the same object for real values."""
        if not is_complex_dtype(self.dtype):
            return self
        return GenericSparseVector(self.n, self.nzind.clone(),
                                   self.nzval.conj().resolve_conj(), copy=False)

    def real(self):
        """real(V) (vector.jl:81-85): the same object for real values."""
        if not is_complex_dtype(self.dtype):
            return self
        return GenericSparseVector(self.n, self.nzind.clone(), self.nzval.real.clone(),
                                   copy=False)

    def imag(self):
        """imag(V) (vector.jl:86-90): zero(V) (no stored entries) for real values."""
        if not is_complex_dtype(self.dtype):
            return GenericSparseVector(self.n, self.nzind[:0].clone(), self.nzval[:0].clone(),
                                       copy=False)
        return GenericSparseVector(self.n, self.nzind.clone(), self.nzval.imag.clone(),
                                   copy=False)

    def __repr__(self):
        return (f"{self.n}-element GenericSparseVector{{{self.dtype}, {self.nzind.dtype}}} "
                f"with {self.nnz()} stored entries on {self.device}")


def _acc_scalar(buf, dtype, widen_int=False):
    from .ops import _acc_value
    return _acc_value(buf, dtype, widen_int)


def norm(V: GenericSparseVector, p=2):
    """LinearAlgebra.norm(V, p) = norm(nonzeros(V), p) (src/vector/vector.jl:91-93)."""
    dev = require_device(V.nzval)
    out = ctypes.c_double()
    L.call("gsa_norm_values", stream_of(dev), vtype_of(V.nzval), V.nnz(), ptr(V.nzval),
           float(p), ctypes.byref(out))
    real32 = V.dtype in (torch.float32, torch.complex64)
    return float(np.float32(out.value)) if real32 else out.value


def normalize(V: GenericSparseVector, p=2):
    """LinearAlgebra.normalize(V, p) (src/vector/vector.jl:95-100): a copy in
    float(promote_eltype(V, norm)) divided by the norm."""
    nrm = norm(V, p)
    T = V.dtype if V.dtype not in (torch.int32, torch.int64) else torch.float64
    W = GenericSparseVector(V.n, V.nzind, V.nzval.to(T))
    return normalize_(W, p, nrm)


def normalize_(V: GenericSparseVector, p=2, nrm=None):
    """LinearAlgebra.normalize!(V, p): V's values divided by norm(V, p) in place."""
    if V.dtype in (torch.int32, torch.int64):
        raise ArgumentError("normalize! needs floating-point values")
    nrm = norm(V, p) if nrm is None else nrm
    V.nzval.div_(nrm)
    return V


def spvec_sum(V: GenericSparseVector):
    """sum(V) = sum(nonzeros(V)) (src/core.jl:24)."""
    from .ops import _acc_buffer
    dev = require_device(V.nzval)
    out = _acc_buffer()
    L.call("gsa_sum_values", stream_of(dev), vtype_of(V.nzval), V.nnz(), ptr(V.nzval), out)
    return _acc_scalar(out, V.dtype, widen_int=True)


def dot(x, y):
    """LinearAlgebra.dot(x::GenericSparseVector, y::DenseVector) (vector.jl:102-139),
    and dot(y::DenseVector, x::GenericSparseVector) = dot(x, y) for real values,
    conj(dot(x, y)) for complex ones (vector.jl:140-145)."""
    from .ops import _acc_buffer
    if isinstance(x, torch.Tensor) and isinstance(y, GenericSparseVector):
        cx, cy = is_complex_dtype(x.dtype), is_complex_dtype(y.dtype)
        if cx != cy:
            raise ArgumentError("dot(::DenseVector, ::GenericSparseVector) needs both real "
                                "or both complex element types")
        r = dot(y, x)
        return r.conjugate() if isinstance(r, complex) else r
    if not isinstance(x, GenericSparseVector) or not isinstance(y, torch.Tensor) or y.dim() != 1:
        raise ArgumentError("dot: a GenericSparseVector and a dense device vector are needed")
    if len(x) != y.shape[0]:
        raise DimensionMismatch(f"Vector x has a length {len(x)} but y has a length {y.shape[0]}")
    if y.device != x.device:
        raise ArgumentError("Vectors x and y must be on the same device/backend.")
    require_device(x.nzval, y)
    if x.dtype not in VTYPES or y.dtype not in VTYPES:
        raise ArgumentError("dot: unsupported element type")
    T = promote_dtype(x.dtype, y.dtype)
    vals = x.nzval if x.dtype == T else x.nzval.to(T)
    yt = y.resolve_conj().to(T).contiguous()
    out = _acc_buffer()
    L.call("gsa_spvec_dot", stream_of(x.device), vtype_of(vals), itype_of(x.nzind), x.n, x.nnz(),
           ptr(x.nzind), ptr(vals), ptr(yt), out)
    return _acc_scalar(out, T)
