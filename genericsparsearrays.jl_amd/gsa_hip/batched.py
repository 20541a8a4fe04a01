"""Many uneven CSR products in one persistent launch (gsa_batch_* in the C ABI).

    batch = BatchedSpMM([(A1, B1, C1), (A2, B2, C2), ...], rows_per_tile=512)
    batch.run(alpha, beta)            # one launch for every instance
    g = batch.graph(alpha, beta)      # HIP-graph capture of that launch
    g.replay()

Each A is a GenericSparseMatrixCSR (Float32 or Float64) or a BF16CSR; B and C
are device vectors or matrices, column-major (Julia's layout) or row-major
(Julia's transpose(Bt)); row-major B keeps the k values one stored entry needs
in one cache line.  All
instances of one batch share the value type and the index type.  This
replaces a loop of mul!(C_i, A_i, B_i, α, β) calls
(src/matrix_csr/matrix_csr.jl:145-211) with one launch.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib as L
from ._lib import ArgumentError, DimensionMismatch
from .matrix import GenericSparseMatrixCSR, itype_of, ptr, stream_of
from .precision import BF16CSR


def _vtype(A) -> int:
    if isinstance(A, BF16CSR):
        return L.GSA_BF16
    if A.nzval.dtype == torch.float64:
        return L.GSA_F64
    if A.nzval.dtype == torch.float32:
        return L.GSA_F32
    raise ArgumentError(f"unsupported value type {A.nzval.dtype}")


def _layout(t: torch.Tensor, rows: int, k: int):
    """(ld, rowmajor) of a dense vector or matrix operand."""
    if t.dim() == 1:
        if t.stride(0) != 1:
            raise ArgumentError("vector operands must be contiguous")
        return max(rows, 1), False
    if t.stride(0) == 1 and (k == 1 or t.stride(1) >= max(rows, 1)):
        return max(t.stride(1), rows, 1), False
    if t.stride(1) == 1 and (rows <= 1 or t.stride(0) >= k):
        return max(t.stride(0), k, 1), True
    raise ArgumentError("dense operands of a batch must be column- or row-major")


def describe(items):
    """(vtype, itype, ctypes array of gsa_batch_desc) for (A, B, C) triples."""
    if not items:
        raise ArgumentError("empty batch")
    vt = _vtype(items[0][0])
    it = itype_of(items[0][0].rowptr)
    dense_dt = torch.float64 if vt == L.GSA_F64 else torch.float32
    arr = (L.BatchDesc * len(items))()
    for i, (A, B, C) in enumerate(items):
        if _vtype(A) != vt or itype_of(A.rowptr) != it:
            raise ArgumentError("all instances of a batch share value and index types")
        if B.dtype != dense_dt or C.dtype != dense_dt:
            raise ArgumentError(f"instance {i}: dense operands must be {dense_dt}")
        m, n = A.shape
        k = 1 if B.dim() == 1 else B.shape[1]
        if B.shape[0] != n:
            raise DimensionMismatch(f"instance {i}: second dimension of A, {n}, does not match "
                                    f"the first dimension of B, {B.shape[0]}")
        if C.shape[0] != m or (1 if C.dim() == 1 else C.shape[1]) != k:
            raise DimensionMismatch(f"instance {i}: C is not {m} x {k}")
        if max(m, n, k) >= 2 ** 31:
            raise ArgumentError(f"instance {i}: dimensions must fit Int32 in a batch")
        d = arr[i]
        d.rowptr, d.colval, d.nzval = A.rowptr.data_ptr(), A.colval.data_ptr(), A.nzval.data_ptr()
        d.row_scale = A.row_scale.data_ptr() if isinstance(A, BF16CSR) else None
        d.B, d.C = B.data_ptr(), C.data_ptr()
        d.m, d.n, d.k = m, n, k
        ldb, brow = _layout(B, n, k)
        ldc, crow = _layout(C, m, k)
        d.ldb, d.ldc = ldb, ldc
        d.flags = (L.GSA_BATCH_B_ROWMAJOR if brow else 0) | (L.GSA_BATCH_C_ROWMAJOR if crow else 0)
    return vt, it, arr


class BatchGraph:
    """An instantiated HIP graph of one batched launch."""

    def __init__(self, handle, batch):
        self.handle = handle
        self._batch = batch  # keeps the plan and the operands alive

    def replay(self, stream=None):
        s = stream if stream is not None else stream_of(self._batch.device)
        L.call("gsa_graph_launch", self.handle, s)

    def __del__(self):
        if self.handle is not None:
            try:
                L.lib().gsa_graph_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


class BatchedSpMM:
    def __init__(self, items, rows_per_tile: int = 512):
        self.items = list(items)
        self.vtype, self.itype, self.desc = describe(self.items)
        self.device = self.items[0][2].device
        self.rows_per_tile = int(rows_per_tile)
        h = ctypes.c_void_p()
        L.call("gsa_batch_plan_create", stream_of(self.device), self.vtype, self.itype,
               len(self.items), self.desc, self.rows_per_tile, ctypes.byref(h))
        self.plan = h

    def run(self, alpha=1.0, beta=0.0, stream=None):
        s = stream if stream is not None else stream_of(self.device)
        L.call("gsa_batch_spmm", s, self.plan, float(alpha), float(beta))

    def graph(self, alpha=1.0, beta=0.0) -> BatchGraph:
        """Capture one launch; the capture uses a private stream so the caller's
        stream is never in capture mode."""
        cap = torch.cuda.Stream(device=self.device)
        cap.wait_stream(torch.cuda.current_stream(self.device))
        h = ctypes.c_void_p()
        L.call("gsa_batch_graph_create", ctypes.c_void_p(cap.cuda_stream), self.plan,
               float(alpha), float(beta), ctypes.byref(h))
        torch.cuda.current_stream(self.device).wait_stream(cap)
        return BatchGraph(h, self)

    def __del__(self):
        if getattr(self, "plan", None) is not None:
            try:
                L.lib().gsa_plan_destroy(self.plan)
            except Exception:
                pass
            self.plan = None
