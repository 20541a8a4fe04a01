"""Reduced-precision value storage for CSR matrices (serving variants).

BF16CSR keeps the structure of a GenericSparseMatrixCSR and stores each value
as BFloat16 with one power-of-two Float32 scale per row (scale_i >= max |a_ij|,
so dividing by it is exact and only the BF16 rounding loses precision).
Products accumulate in Float32; rows over 256 stored entries accumulate with
TwoSum.  The reference has no such format — its value type is any Julia
number type — so this is an addition, with its error stated against the
Float64 oracle by the tests and the benchmark.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L
from ._lib import ArgumentError, DimensionMismatch
from .matrix import (GenericSparseMatrixCSR, itype_of, vtype_of, ptr, stream_of, require_device,
                     _PlanCache)
from .ops import _dense_operand


class BF16CSR:
    """CSR with BF16 values (bit patterns in an int16 tensor) and row scales."""

    def __init__(self, A: GenericSparseMatrixCSR):
        if not isinstance(A, GenericSparseMatrixCSR):
            raise ArgumentError("BF16CSR needs a GenericSparseMatrixCSR")
        dev = require_device(A.rowptr, A.colval, A.nzval)
        self.m, self.n = A.m, A.n
        self.rowptr, self.colval = A.rowptr, A.colval
        self.nzval = torch.empty(A.nnz(), dtype=torch.int16, device=dev)
        self.row_scale = torch.empty(A.m, dtype=torch.float32, device=dev)
        L.call("gsa_csr_to_bf16", stream_of(dev), vtype_of(A.nzval), itype_of(A.rowptr), A.m,
               ptr(A.rowptr), ptr(A.nzval), ptr(self.nzval), ptr(self.row_scale))
        self._plan = _PlanCache()

    def plan(self):
        return self._plan.get(self.rowptr, self.m, self.nnz(), self.colval, self.n)

    @property
    def shape(self):
        return (self.m, self.n)

    def nnz(self):
        return int(self.nzval.numel())

    @property
    def device(self):
        return self.nzval.device

    def dequantized_values(self) -> np.ndarray:
        """Float64 values scale_i * bf16_ij (host), for error accounting."""
        bits = self.nzval.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
        v = bits.view(np.float32).astype(np.float64)
        rows = np.repeat(np.arange(self.m), np.diff(self.rowptr.cpu().numpy()))
        return v * self.row_scale.cpu().numpy().astype(np.float64)[rows]


def mul_bf16(C: torch.Tensor, A: BF16CSR, B: torch.Tensor, alpha=1.0, beta=0.0):
    """C = beta*C + alpha * A_bf16 * B with Float32 B and C.  B may be a vector,
    column-major or row-major (Julia's transpose(Bt)); C a vector or column-major."""
    if B.dtype != torch.float32 or C.dtype != torch.float32:
        raise ArgumentError("BF16 products take Float32 dense operands")
    k = 1 if B.dim() == 1 else B.shape[1]
    if B.shape[0] != A.n:
        raise DimensionMismatch(f"second dimension of A, {A.n}, does not match the first "
                                f"dimension of B, {B.shape[0]}")
    if C.shape[0] != A.m or (1 if C.dim() == 1 else C.shape[1]) != k:
        raise DimensionMismatch("C does not match A*B")
    dev = require_device(B, C, A.nzval)
    Bt, k, ldb, opB = _dense_operand(B)
    if C.dim() == 2 and not (C.stride(0) == 1 and (k == 1 or C.stride(1) >= max(A.m, 1))):
        raise ArgumentError("C must be column-major")
    ldc = max(A.m, 1) if C.dim() == 1 else max(C.stride(1), A.m, 1)
    L.call("gsa_spmm_csr_bf16", stream_of(dev), itype_of(A.rowptr), A.m, A.n, k, ptr(A.rowptr),
           ptr(A.colval), ptr(A.nzval), ptr(A.row_scale), A.nnz(), ptr(Bt), ldb, opB, ptr(C), ldc,
           L.GSA_OP_N, float(alpha), float(beta), A.plan() if A.nnz() > 0 else None)
    return C
