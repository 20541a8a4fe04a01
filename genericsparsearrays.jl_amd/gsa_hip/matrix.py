"""Sparse matrix types: GenericSparseMatrixCSR / CSC / COO on HIP device storage.

Mirrors the reference's type layer (src/matrix_csr/matrix_csr.jl:17-63,
src/matrix_csc/matrix_csc.jl:17-63, src/matrix_coo/matrix_coo.jl:17-61):
dimensions m, n and three storage vectors with ONE-BASED indices, the
constructor's checks and its copy of the buffers.  Storage is torch tensors
(used only as device memory); every operation runs in libgsa_hip.so.

Dense operands follow Julia's column-major layout: a dense m x k matrix is a
torch tensor of shape (m, k) whose stride(0) == 1 (see `dense`), and a
row-major tensor is read as the Transpose of a column-major one.
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np
import torch

from . import _lib as L
from ._lib import ArgumentError, DimensionMismatch

# value types: the reference's test matrix (test/runtests.jl:34-68) -- Int32,
# Int64, Float32, Float64, ComplexF32, ComplexF64
VTYPES = {torch.float64: L.GSA_F64, torch.float32: L.GSA_F32,
          torch.complex64: L.GSA_C64, torch.complex128: L.GSA_C128,
          torch.int32: L.GSA_VI32, torch.int64: L.GSA_VI64}
ITYPES = {torch.int32: L.GSA_I32, torch.int64: L.GSA_I64}
_SCALAR_CT = {torch.float64: ctypes.c_double, torch.float32: ctypes.c_float,
              torch.complex128: ctypes.c_double * 2, torch.complex64: ctypes.c_float * 2,
              torch.int64: ctypes.c_int64, torch.int32: ctypes.c_int32}


def vtype_of(t: torch.Tensor) -> int:
    try:
        return VTYPES[t.dtype]
    except KeyError:
        raise ArgumentError(f"element type {t.dtype} is not provided by the HIP path "
                            "(Int32, Int64, Float32, Float64, ComplexF32 and ComplexF64 "
                            "are)") from None


def is_complex_dtype(dt) -> bool:
    return dt in (torch.complex64, torch.complex128)


def scalar_ref(value, dtype):
    """A host value of the element type `dtype`, passed to the C ABI by pointer
    (alpha, beta, lambda).  Integers wrap to their width, as Julia's do."""
    ct = _SCALAR_CT.get(dtype)
    if ct is None:
        raise ArgumentError(f"no scalar of element type {dtype}")
    if dtype in (torch.complex64, torch.complex128):
        z = complex(value)
        return ct(z.real, z.imag)
    if dtype in (torch.int64, torch.int32):
        if isinstance(value, complex) or (isinstance(value, float) and not float(value).is_integer()):
            raise ArgumentError(f"InexactError: {value} is not an integer of {dtype}")
        bits = 64 if dtype == torch.int64 else 32
        v = int(value) & ((1 << bits) - 1)
        if v >= 1 << (bits - 1):
            v -= 1 << bits
        return ct(v)
    if isinstance(value, complex):
        if value.imag != 0:
            raise ArgumentError(f"InexactError: {value} is not real")
        value = value.real
    return ct(float(value))


def itype_of(t: torch.Tensor) -> int:
    try:
        return ITYPES[t.dtype]
    except KeyError:
        raise ArgumentError(f"index type {t.dtype} is not provided (Int32 and Int64 are)") from None


def ptr(t: torch.Tensor | None):
    return ctypes.c_void_p(t.data_ptr() if t is not None and t.numel() > 0 else 0)


def stream_of(device: torch.device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_device(*tensors: torch.Tensor) -> torch.device:
    """All storage on one HIP device (the reference's backend-equality check,
    src/matrix_csr/matrix_csr.jl:185-190)."""
    devs = {t.device for t in tensors if t is not None}
    if len(devs) != 1:
        raise ArgumentError("All arrays must have the same backend")
    dev = devs.pop()
    if dev.type != "cuda":
        raise ArgumentError(
            f"storage is on {dev}; the HIP kernel layer needs device storage "
            "(use adapt('cuda', A)) — there is no CPU path")
    return dev


def dense(a, device="cuda", dtype=None) -> torch.Tensor:
    """A column-major (Julia layout) device copy of a host/numpy/torch 1-D or 2-D array."""
    t = torch.as_tensor(np.asarray(a) if not isinstance(a, torch.Tensor) else a)
    if dtype is not None:
        t = t.to(dtype)
    if t.dim() == 1:
        return t.to(device).contiguous()
    out = torch.empty((t.shape[1], t.shape[0]), dtype=t.dtype, device=device).t()
    out.copy_(t)
    return out


def similar_dense(rows: int, k: int | None, dtype, device) -> torch.Tensor:
    if k is None:
        return torch.empty(rows, dtype=dtype, device=device)
    return torch.empty((k, rows), dtype=dtype, device=device).t()


class AbstractGenericSparseMatrix:
    """Common interface (src/core.jl:11-54)."""

    m: int
    n: int

    @property
    def shape(self):
        return (self.m, self.n)

    def size(self, dim=None):
        return self.shape if dim is None else self.shape[dim - 1]

    def __len__(self):
        return self.m * self.n

    def nnz(self) -> int:
        return int(self.nzval.numel())

    def nonzeros(self) -> torch.Tensor:
        return self.nzval

    @property
    def dtype(self):
        return self.nzval.dtype

    @property
    def index_dtype(self):
        return self._idx().dtype

    @property
    def device(self):
        return self.nzval.device

    def _idx(self):
        raise NotImplementedError

    # ---- wrappers / arithmetic sugar (Julia operators) ----
    @property
    def T(self):
        return Transpose(self)

    def transpose(self):
        return Transpose(self)

    def adjoint(self):
        return Adjoint(self)

    def __matmul__(self, other):
        from . import ops
        return ops.times(self, other)

    def __mul__(self, other):
        from . import ops
        if isinstance(other, (int, float)):
            return ops.scale(self, other)
        return ops.times(self, other)

    def __rmul__(self, other):
        from . import ops
        if isinstance(other, (int, float)):
            return ops.scale(self, other)
        return NotImplemented

    def __truediv__(self, other):
        from . import ops
        return ops.scale(self, 1.0 / other)

    def __add__(self, other):
        from . import ops
        return ops.add(self, other)

    def __sub__(self, other):
        from . import ops
        return ops.add(self, ops.negate(other))

    def __radd__(self, other):  # dense + sparse (src/core.jl:89)
        from . import ops
        return ops.add(other, self)

    def __rsub__(self, other):  # dense - sparse (src/core.jl:93)
        from . import ops
        return ops.add(other, ops.negate(self))

    def __neg__(self):
        from . import ops
        return ops.negate(self)

    def __pos__(self):
        return self.copy()

    # ---- maps over the stored values (src/matrix_csr/matrix_csr.jl:80-106,
    # src/core.jl:26-40; CSC and COO alike): the structure is copied, the values
    # mapped on the device as the reference's broadcasts are ----
    def _index_arrays(self):
        raise NotImplementedError

    def _with_values(self, vals: torch.Tensor):
        return type(self)(self.m, self.n, *(t.clone() for t in self._index_arrays()), vals,
                          copy=False)

    def zero(self):
        """Base.zero(A): the same size, no stored entries (a pointer of ones)."""
        a, b = self._index_arrays()
        empty_i = torch.empty(0, dtype=b.dtype, device=b.device)
        empty_v = torch.empty(0, dtype=self.nzval.dtype, device=self.nzval.device)
        if isinstance(self, GenericSparseMatrixCOO):
            return type(self)(self.m, self.n, empty_i, empty_i.clone(), empty_v, copy=False)
        ones = torch.ones(a.numel(), dtype=a.dtype, device=a.device)
        return type(self)(self.m, self.n, ones, empty_i, empty_v, copy=False)

    def conj(self):
        """Base.conj(A): A itself for real values, else conjugated values."""
        if not self.nzval.is_complex():
            return self
        return self._with_values(self.nzval.conj().resolve_conj())

    def real(self):
        """Base.real(A): A itself for real values, else the real parts."""
        if not self.nzval.is_complex():
            return self
        return self._with_values(self.nzval.real.contiguous())

    def imag(self):
        """Base.imag(A): zero(A) for real values, else the imaginary parts (the
        structure kept, zeros included)."""
        if not self.nzval.is_complex():
            return self.zero()
        return self._with_values(self.nzval.imag.contiguous())

    def rmul_(self, x):
        """LinearAlgebra.rmul!(A, x): the stored values times x, in place."""
        self.nzval.mul_(x)
        return self

    def lmul_(self, x):
        """LinearAlgebra.lmul!(x, A)."""
        self.nzval.mul_(x)
        return self

    def rdiv_(self, x):
        """LinearAlgebra.rdiv!(A, x) = rmul!(A, inv(Tv(x)))."""
        if self.nzval.is_floating_point() or self.nzval.is_complex():
            inv = 1.0 / torch.tensor(x, dtype=self.nzval.dtype).item()
            self.nzval.mul_(inv)
            return self
        raise ArgumentError("rdiv!: integer values have no inverse")

    def to_scipy(self):
        raise NotImplementedError

    def collect(self) -> np.ndarray:
        """Dense host copy (Base.collect)."""
        return np.asarray(self.to_scipy().toarray())


def _check_common(m, n, tensors, itype_tensor):
    if m < 0:
        raise ArgumentError("m must be non-negative")
    if n < 0:
        raise ArgumentError("n must be non-negative")
    devs = {t.device for t in tensors}
    if len(devs) != 1:
        raise ArgumentError("All storage vectors must be on the same device/backend.")
    # SparseArrays.sparse_check_Ti: the dimensions must fit the index type
    if itype_tensor.dtype == torch.int32 and (m >= 2 ** 31 or n >= 2 ** 31):
        raise ArgumentError("number of rows/columns does not fit the index type Int32")
    if itype_tensor.dtype not in ITYPES:
        raise ArgumentError(f"index type {itype_tensor.dtype} is not supported")


class _PlanCache:
    """Per-matrix cached analysis (gsa_plan): CSR row blocks and SpMV windows;
    for CSC also the row view, for COO the row view alone (the sorted-key form
    of mul!, without atomics).  It is rebuilt when an array it was built from
    changes (keyed on each array's address, type and torch version counter).
    A plan is read-only once built, so products on several streams at the same
    time may share it."""

    def __init__(self):
        self.handle = None
        self.key = None
        self.transposed = False  # the handle holds the transposed product's bins
        self.gathers = set()     # gather views it holds: 0 = A*B, 1 = transpose(A)*B

    def get(self, ptr_t: torch.Tensor, m: int, nnz: int, col_t: torch.Tensor | None = None,
            n: int = 0):
        """CSR: the analysis of row pointer `ptr_t` (m rows); with the column array
        `col_t` (n columns) also the XCD launch order of the SpMV windows."""
        def create(h):
            L.call("gsa_plan_create", stream_of(ptr_t.device), itype_of(ptr_t), m, nnz,
                   ptr(ptr_t), ctypes.byref(h))
            if col_t is not None and nnz > 0:
                L.call("gsa_plan_set_columns", stream_of(ptr_t.device), h, ptr(col_t), n)
        arrays = (ptr_t,) if col_t is None else (ptr_t, col_t)
        return self._get(("csr",), create, arrays, m, nnz, n)

    def get_csc(self, colptr: torch.Tensor, rowval: torch.Tensor, m: int, n: int, nnz: int):
        return self._get(("csc",), lambda h: L.call(
            "gsa_plan_create_csc", stream_of(colptr.device), itype_of(colptr), m, n, nnz,
            ptr(colptr), ptr(rowval), ctypes.byref(h)), (colptr, rowval), m, n, nnz)

    def get_coo(self, rowind: torch.Tensor, colind: torch.Tensor, m: int, n: int, nnz: int):
        return self._get(("coo",), lambda h: L.call(
            "gsa_plan_create_coo", stream_of(rowind.device), itype_of(rowind), m, n, nnz,
            ptr(rowind), ptr(colind), ctypes.byref(h)), (rowind, colind), m, n, nnz)

    @staticmethod
    def _key(kind, arrays, dims):
        return (kind[0],) + tuple((t.data_ptr(), t.dtype, t._version) for t in arrays) + dims

    def peek(self, kind, arrays, *dims):
        """The cached plan if it was built on these arrays, else None (builds nothing)."""
        if self.handle is not None and self.key == self._key(kind, arrays, dims):
            return self.handle
        return None

    def _get(self, kind, create, arrays, *dims):
        key = self._key(kind, arrays, dims)
        if self.handle is not None and self.key == key:
            return self.handle
        self.release()
        h = ctypes.c_void_p()
        try:
            create(h)
        except Exception:
            if h.value:
                L.lib().gsa_plan_destroy(h)
            raise
        self.handle, self.key = h, key
        self.transposed = False
        self.gathers = set()
        return h

    def add_gather(self, h, fmt, trans, itype, m, n, nnz, a, b):
        """Give plan `h` the gather view of A*B (trans 0) or transpose(A)*B (trans 1)
        for several right-hand sides (built once per plan)."""
        if trans not in self.gathers:
            L.call("gsa_plan_add_gather_view", stream_of(a.device), h, fmt, int(trans), itype, m,
                   n, nnz, ptr(a), ptr(b))
            self.gathers.add(trans)
        return h

    def add_transpose(self, h, fmt, itype, m, n, nnz, a, b):
        """Give plan `h` the bins of transpose(A) * x (built once per plan)."""
        if not self.transposed:
            L.call("gsa_plan_add_transpose", stream_of(a.device), h, fmt, itype, m, n, nnz,
                   ptr(a), ptr(b))
            self.transposed = True
        return h

    def release(self):
        if self.handle is not None:
            try:
                L.lib().gsa_plan_destroy(self.handle)
            except Exception:
                pass
            self.handle = None

    def __del__(self):
        self.release()


class GenericSparseMatrixCSR(AbstractGenericSparseMatrix):
    """CSR matrix: rowptr (m+1), colval (nnz), nzval (nnz), one-based."""

    def __init__(self, *args, copy: bool = True, device=None, index_dtype=torch.int64):
        if len(args) == 5:
            m, n, rowptr, colval, nzval = args
            m, n = int(m), int(n)
            _check_common(m, n, (rowptr, colval, nzval), rowptr)
            if rowptr.numel() != m + 1:
                raise ArgumentError("rowptr length must be m+1")
            if colval.numel() != nzval.numel():
                raise ArgumentError("colval and nzval must have same length")
            if colval.dtype != rowptr.dtype:
                raise ArgumentError("rowptr and colval must share the index type")
            self.m, self.n = m, n
            self.rowptr = rowptr.clone() if copy else rowptr
            self.colval = colval.clone() if copy else colval
            self.nzval = nzval.clone() if copy else nzval
            self._plan = _PlanCache()
            return
        if len(args) == 1:
            from . import convert
            src = args[0]
            out = convert.to_csr(src, device=device, index_dtype=index_dtype)
            self.m, self.n = out.m, out.n
            self.rowptr, self.colval, self.nzval = out.rowptr, out.colval, out.nzval
            self._plan = _PlanCache()
            return
        raise ArgumentError("GenericSparseMatrixCSR(m, n, rowptr, colval, nzval) or (A)")

    def _idx(self):
        return self.colval

    def getrowptr(self):
        return self.rowptr

    def colvals(self):
        return self.colval

    def plan(self):
        return self._plan.get(self.rowptr, self.m, self.nnz(), self.colval, self.n)

    def cached_plan(self):
        """plan() if it is already built for the current arrays, else None."""
        return self._plan.peek(("csr",), (self.rowptr, self.colval), self.m, self.nnz(), self.n)

    def plan_transposed(self):
        """plan() with the row bins of transpose(A) * x added (no atomics)."""
        return self._plan.add_transpose(self.plan(), L.GSA_FMT_CSR, itype_of(self.rowptr), self.m,
                                        self.n, self.nnz(), self.rowptr, self.colval)

    def plan_gather(self, trans: bool = True):
        """plan() with the gather view of transpose(A) * B (several right-hand sides)."""
        return self._plan.add_gather(self.plan(), L.GSA_FMT_CSR, 1, itype_of(self.rowptr), self.m,
                                     self.n, self.nnz(), self.rowptr, self.colval)

    def copy(self):
        return GenericSparseMatrixCSR(self.m, self.n, self.rowptr, self.colval, self.nzval)

    def _index_arrays(self):
        return self.rowptr, self.colval

    def to_scipy(self):
        import scipy.sparse as sp
        rp = self.rowptr.cpu().numpy().astype(np.int64) - 1
        cv = self.colval.cpu().numpy().astype(np.int64) - 1
        return sp.csr_matrix((self.nzval.cpu().numpy(), cv, rp), shape=(self.m, self.n))

    def __repr__(self):
        return (f"{self.m}×{self.n} GenericSparseMatrixCSR{{{self.dtype}, {self.index_dtype}}} "
                f"with {self.nnz()} stored entries on {self.device}")


class GenericSparseMatrixCSC(AbstractGenericSparseMatrix):
    """CSC matrix: colptr (n+1), rowval (nnz), nzval (nnz), one-based."""

    def __init__(self, *args, copy: bool = True, device=None, index_dtype=torch.int64):
        if len(args) == 5:
            m, n, colptr, rowval, nzval = args
            m, n = int(m), int(n)
            _check_common(m, n, (colptr, rowval, nzval), colptr)
            if colptr.numel() != n + 1:
                raise ArgumentError("colptr length must be n+1")
            if rowval.numel() != nzval.numel():
                raise ArgumentError("rowval and nzval must have same length")
            if rowval.dtype != colptr.dtype:
                raise ArgumentError("colptr and rowval must share the index type")
            self.m, self.n = m, n
            self.colptr = colptr.clone() if copy else colptr
            self.rowval = rowval.clone() if copy else rowval
            self.nzval = nzval.clone() if copy else nzval
            self._plan = _PlanCache()
            return
        if len(args) == 1:
            from . import convert
            out = convert.to_csc(args[0], device=device, index_dtype=index_dtype)
            self.m, self.n = out.m, out.n
            self.colptr, self.rowval, self.nzval = out.colptr, out.rowval, out.nzval
            self._plan = _PlanCache()
            return
        raise ArgumentError("GenericSparseMatrixCSC(m, n, colptr, rowval, nzval) or (A)")

    def _idx(self):
        return self.rowval

    def getcolptr(self):
        return self.colptr

    def rowvals(self):
        return self.rowval

    def plan(self):
        return self._plan.get_csc(self.colptr, self.rowval, self.m, self.n, self.nnz())

    def cached_plan(self):
        """plan() if it is already built for the current arrays, else None."""
        return self._plan.peek(("csc",), (self.colptr, self.rowval), self.m, self.n, self.nnz())

    def plan_gather(self, trans: bool = False):
        """plan() with the gather view of A * B (several right-hand sides)."""
        return self._plan.add_gather(self.plan(), L.GSA_FMT_CSC, 0, itype_of(self.colptr), self.m,
                                     self.n, self.nnz(), self.colptr, self.rowval)

    def copy(self):
        return GenericSparseMatrixCSC(self.m, self.n, self.colptr, self.rowval, self.nzval)

    def _index_arrays(self):
        return self.colptr, self.rowval

    def to_scipy(self):
        import scipy.sparse as sp
        cp = self.colptr.cpu().numpy().astype(np.int64) - 1
        rv = self.rowval.cpu().numpy().astype(np.int64) - 1
        return sp.csc_matrix((self.nzval.cpu().numpy(), rv, cp), shape=(self.m, self.n))

    def __repr__(self):
        return (f"{self.m}×{self.n} GenericSparseMatrixCSC{{{self.dtype}, {self.index_dtype}}} "
                f"with {self.nnz()} stored entries on {self.device}")


class GenericSparseMatrixCOO(AbstractGenericSparseMatrix):
    """COO matrix: rowind, colind, nzval (nnz each), one-based, any order."""

    def __init__(self, *args, copy: bool = True, device=None, index_dtype=torch.int64):
        if len(args) == 5:
            m, n, rowind, colind, nzval = args
            m, n = int(m), int(n)
            _check_common(m, n, (rowind, colind, nzval), rowind)
            if not (rowind.numel() == colind.numel() == nzval.numel()):
                raise ArgumentError("rowind, colind, and nzval must have same length")
            if rowind.dtype != colind.dtype:
                raise ArgumentError("rowind and colind must share the index type")
            self.m, self.n = m, n
            self.rowind = rowind.clone() if copy else rowind
            self.colind = colind.clone() if copy else colind
            self.nzval = nzval.clone() if copy else nzval
            return
        if len(args) == 1:
            from . import convert
            out = convert.to_coo(args[0], device=device, index_dtype=index_dtype)
            self.m, self.n = out.m, out.n
            self.rowind, self.colind, self.nzval = out.rowind, out.colind, out.nzval
            return
        raise ArgumentError("GenericSparseMatrixCOO(m, n, rowind, colind, nzval) or (A)")

    def _idx(self):
        return self.rowind

    def plan(self):
        """The row view of the entries (a stable sort by row, once; none when
        the storage is already row-ordered): mul! then reduces rows without
        atomics."""
        if getattr(self, "_plan", None) is None:
            self._plan = _PlanCache()
        return self._plan.get_coo(self.rowind, self.colind, self.m, self.n, self.nnz())

    def plan_transposed(self):
        """plan() with the row bins of transpose(A) * x added (no atomics)."""
        h = self.plan()  # creates the cache on first use
        return self._plan.add_transpose(h, L.GSA_FMT_COO, itype_of(self.rowind), self.m, self.n,
                                        self.nnz(), self.rowind, self.colind)

    def plan_gather(self, trans: bool = False):
        """plan() with the gather view of A * B or transpose(A) * B (several right-hand sides)."""
        h = self.plan()
        return self._plan.add_gather(h, L.GSA_FMT_COO, 1 if trans else 0, itype_of(self.rowind),
                                     self.m, self.n, self.nnz(), self.rowind, self.colind)

    def getrowind(self):
        return self.rowind

    def getcolind(self):
        return self.colind

    def copy(self):
        return GenericSparseMatrixCOO(self.m, self.n, self.rowind, self.colind, self.nzval)

    def _index_arrays(self):
        return self.rowind, self.colind

    def to_scipy(self):
        """SparseMatrixCSC(::COO) sums duplicates (conversions.jl:59-66)."""
        import scipy.sparse as sp
        r = self.rowind.cpu().numpy().astype(np.int64) - 1
        c = self.colind.cpu().numpy().astype(np.int64) - 1
        return sp.coo_matrix((self.nzval.cpu().numpy(), (r, c)), shape=(self.m, self.n)).tocsc()

    def __repr__(self):
        return (f"{self.m}×{self.n} GenericSparseMatrixCOO{{{self.dtype}, {self.index_dtype}}} "
                f"with {self.nnz()} stored entries on {self.device}")


SPARSE_TYPES = (GenericSparseMatrixCSR, GenericSparseMatrixCSC, GenericSparseMatrixCOO)


class Transpose:
    """LinearAlgebra.Transpose of a sparse matrix (a lazy wrapper)."""

    conj = False

    def __init__(self, parent):
        self.parent = parent

    @property
    def shape(self):
        return (self.parent.n, self.parent.m)

    def size(self, dim=None):
        return self.shape if dim is None else self.shape[dim - 1]

    @property
    def T(self):
        return self.parent

    def __matmul__(self, other):
        from . import ops
        return ops.times(self, other)

    __mul__ = __matmul__

    def __add__(self, other):
        from . import ops
        return ops.add(self, other)

    def __sub__(self, other):
        from . import ops
        return ops.add(self, ops.negate(other))

    def __radd__(self, other):
        from . import ops
        return ops.add(other, self)

    def __rsub__(self, other):
        from . import ops
        return ops.add(other, ops.negate(self))

    def __neg__(self):
        from . import ops
        return type(self)(ops.negate(self.parent))

    def collect(self):
        return self.parent.collect().T


class Adjoint(Transpose):
    """LinearAlgebra.Adjoint: the transpose with conjugated values (equal to
    Transpose for real element types)."""

    conj = True

    def collect(self):
        return np.conj(self.parent.collect().T)


def scalar_dtype(x):
    """The element type a scalar brings to Julia's promote_type: None for Bool and
    integers (they never widen Float32 or Float64), float64 for a Python float
    (a Julia Float64 literal), the scalar's own type for numpy and torch scalars."""
    if isinstance(x, (bool, int, np.bool_, np.integer)):
        return None
    if isinstance(x, np.floating):
        return torch.float32 if x.dtype == np.float32 else torch.float64
    if isinstance(x, np.complexfloating):
        return torch.complex64 if x.dtype == np.complex64 else torch.complex128
    if isinstance(x, torch.Tensor):
        if x.numel() != 1:
            raise ArgumentError("a scalar is needed")
        return x.dtype if (x.dtype.is_floating_point or x.dtype.is_complex) else None
    if isinstance(x, float):
        return torch.float64
    if isinstance(x, complex):
        return torch.complex128
    raise ArgumentError(f"unsupported scalar {type(x).__name__}")


_KIND = {torch.int32: (0, 32), torch.int64: (0, 64), torch.float32: (1, 32),
         torch.float64: (1, 64), torch.complex64: (2, 32), torch.complex128: (2, 64)}
_OF_KIND = {(0, 32): torch.int32, (0, 64): torch.int64, (1, 32): torch.float32,
            (1, 64): torch.float64, (2, 32): torch.complex64, (2, 64): torch.complex128}


def promote_types(a, b):
    """Julia's promote_type over Int32, Int64, Float32, Float64, ComplexF32 and
    ComplexF64: integers widen to the wider integer and give way to any float or
    complex type; floats widen; a complex type takes the wider real part."""
    ka, wa = _KIND[a]
    kb, wb = _KIND[b]
    if ka == 0 and kb == 0:
        return _OF_KIND[(0, max(wa, wb))]
    if ka == 0:
        return b
    if kb == 0:
        return a
    return _OF_KIND[(max(ka, kb), max(wa, wb))]


def promote_dtype(*dts):
    """promote_type over the element types (None = a type that widens nothing)."""
    dts = [d for d in dts if d is not None]
    if not dts:
        return None
    out = dts[0]
    for d in dts[1:]:
        out = promote_types(out, d)
    return out


class UniformScaling:
    """LinearAlgebra.UniformScaling λI of any size.  `I` is λ = true.

    A ± λI, λI ± A (src/core.jl:89-94) and A*λI, λI*A (src/core.jl:46-47)
    dispatch here.  λ's element type joins A's by `scalar_dtype` / `promote_dtype`
    (Julia's promote_type(Tv, typeof(λ)))."""

    # numpy scalars defer to __rmul__ (so np.float32(α) * I keeps α's type)
    __array_ufunc__ = None

    def __init__(self, lam=True, dtype=...):
        self.lam = lam
        # dtype: the element type λ brings (None: one that widens nothing)
        self.dtype = scalar_dtype(lam) if dtype is ... else dtype

    @property
    def value(self):
        """λ as a Python number: int for Bool and integers, complex for complex
        scalars, float otherwise."""
        v = self.lam.item() if isinstance(self.lam, torch.Tensor) else self.lam
        if isinstance(v, (bool, np.bool_, int, np.integer)):
            return int(v)
        if isinstance(v, (complex, np.complexfloating)):
            return complex(v)
        return float(v)

    def __mul__(self, other):
        if isinstance(other, (AbstractGenericSparseMatrix, Transpose)):
            from . import ops
            return ops.scale(other, self.value)
        o = other.item() if isinstance(other, torch.Tensor) else other
        return UniformScaling(self.value * o, promote_dtype(self.dtype, scalar_dtype(other)))

    def __rmul__(self, other):
        o = other.item() if isinstance(other, torch.Tensor) else other
        return UniformScaling(o * self.value, promote_dtype(self.dtype, scalar_dtype(other)))

    def __neg__(self):
        return UniformScaling(-self.value, self.dtype)

    def __add__(self, other):
        from . import ops
        return ops.add(other, self)

    __radd__ = __add__

    def __sub__(self, other):
        from . import ops
        return ops.add(ops.negate(other), self)

    def __rsub__(self, other):
        from . import ops
        return ops.add(other, -self)

    def __repr__(self):
        return f"UniformScaling({self.lam!r})"


I = UniformScaling(True)


class Diagonal:
    """LinearAlgebra.Diagonal with its values in a device vector `diag` (n x n), or,
    from `Diagonal.fill`, one value for the whole diagonal (FillArrays'
    Diagonal(Fill(v, n)), ext/GenericSparseArraysFillArraysExt/kron_coo.jl:5-48)."""

    def __init__(self, diag: torch.Tensor):
        if not isinstance(diag, torch.Tensor) or diag.dim() != 1:
            raise ArgumentError("Diagonal needs a 1-D tensor of values")
        self.diag = diag
        self.fill_value = None
        self._n = int(diag.shape[0])
        self._dtype = diag.dtype
        self._device = diag.device

    @classmethod
    def fill(cls, value, n: int, dtype=torch.float64, device="cuda"):
        """Diagonal(Fill(value, n)): n copies of one value, stored as that value."""
        D = cls.__new__(cls)
        D.diag = None
        D.fill_value = float(value)
        D._n = int(n)
        D._dtype = dtype
        D._device = torch.device(device)
        return D

    @property
    def shape(self):
        return (self._n, self._n)

    def size(self, dim=None):
        return self.shape if dim is None else self.shape[dim - 1]

    @property
    def dtype(self):
        return self._dtype

    @property
    def device(self):
        return self._device

    def collect(self) -> np.ndarray:
        if self.diag is None:
            return np.diag(np.full(self._n, self.fill_value))
        return np.diag(self.diag.detach().cpu().numpy())


def transpose(A):
    return A.parent if isinstance(A, Transpose) and not isinstance(A, Adjoint) else Transpose(A)


def adjoint(A):
    return A.parent if isinstance(A, Adjoint) else Adjoint(A)


def adapt(device, A):
    """Adapt.adapt_structure(to, A) (src/matrix_csr/matrix_csr.jl:65-71): re-home the buffers."""
    dev = torch.device(device)
    if isinstance(A, GenericSparseMatrixCSR):
        return GenericSparseMatrixCSR(A.m, A.n, A.rowptr.to(dev), A.colval.to(dev),
                                      A.nzval.to(dev))
    if isinstance(A, GenericSparseMatrixCSC):
        return GenericSparseMatrixCSC(A.m, A.n, A.colptr.to(dev), A.rowval.to(dev),
                                      A.nzval.to(dev))
    if isinstance(A, GenericSparseMatrixCOO):
        return GenericSparseMatrixCOO(A.m, A.n, A.rowind.to(dev), A.colind.to(dev),
                                      A.nzval.to(dev))
    raise ArgumentError(f"cannot adapt {type(A)}")


def nnz(A) -> int:
    return A.nnz()


def nonzeros(A):
    return A.nonzeros()


__all__ = [
    "GenericSparseMatrixCSR", "GenericSparseMatrixCSC", "GenericSparseMatrixCOO",
    "Transpose", "Adjoint", "transpose", "adjoint", "adapt", "nnz", "nonzeros", "dense",
    "UniformScaling", "I", "Diagonal",
]

# keep a module-level registry of live plans so interpreter shutdown frees them first
_live = weakref.WeakSet()
