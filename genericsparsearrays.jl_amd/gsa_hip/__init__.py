"""gsa_hip — MI355X (gfx950) sparse kernel layer with GenericSparseArrays.jl's API.

Types:      GenericSparseMatrixCSR, GenericSparseMatrixCSC, GenericSparseMatrixCOO,
            GenericSparseVector (dot, sum, norm, normalize, scalar maps)
Values:     Int32, Int64, Float32, Float64, ComplexF32, ComplexF64
Wrappers:   transpose / adjoint (lazy Transpose / Adjoint)
Operators:  mul_ (mul!), A @ x / A * x, A * B (SpGEMM), A + B, A - B, kron,
            dropzeros, dropzeros_, conversions through the constructors,
            tr, sum, dot(x, A, y), issymmetric, ishermitian, A ± λI (UniformScaling, I),
            kron with a Diagonal.
Host interop: convert.csr_from_scipy / csc_from_scipy / coo_from_scipy (ingest),
            SparseMatrixCSC(A) and convert.to_scipy(A) (egress).
Every operator runs in libgsa_hip.so; there is no CPU fallback.
"""
from ._lib import ArgumentError, DimensionMismatch, GSAError, load as load_library
from .matrix import (GenericSparseMatrixCSR, GenericSparseMatrixCSC, GenericSparseMatrixCOO,
                     Transpose, Adjoint, transpose, adjoint, adapt, nnz, nonzeros, dense,
                     UniformScaling, I, Diagonal)
from .ops import (mul_, times, add, kron, dropzeros, dropzeros_, spgemm, scale, negate, tr,
                  issymmetric, ishermitian, add_uniform_scaling)
from .ops import dot as _dot3, sum_ as _sum_matrix
from .vector import GenericSparseVector, norm, normalize, normalize_
from . import vector as _vector
from . import convert
from .convert import sparse_matrix_csc as SparseMatrixCSC


def dot(*args):
    """LinearAlgebra.dot: dot(x, A, y) for a sparse matrix A (ops.dot), or
    dot(x, y) with a GenericSparseVector on either side (vector.dot)."""
    if len(args) == 3:
        return _dot3(*args)
    if len(args) == 2:
        return _vector.dot(*args)
    raise ArgumentError("dot takes (x, y) or (x, A, y)")


def sum_(A):
    """Base.sum of a sparse matrix or a GenericSparseVector: sum(nonzeros(A))."""
    if isinstance(A, GenericSparseVector):
        return _vector.spvec_sum(A)
    return _sum_matrix(A)


sum = sum_  # Base.sum (shadows the builtin in this namespace only)

__all__ = [
    "ArgumentError", "DimensionMismatch", "GSAError", "load_library",
    "GenericSparseMatrixCSR", "GenericSparseMatrixCSC", "GenericSparseMatrixCOO",
    "Transpose", "Adjoint", "transpose", "adjoint", "adapt", "nnz", "nonzeros", "dense",
    "mul_", "times", "add", "kron", "dropzeros", "dropzeros_", "spgemm", "scale", "negate",
    "convert", "UniformScaling", "I", "Diagonal", "tr", "sum_", "dot", "issymmetric",
    "ishermitian", "add_uniform_scaling", "GenericSparseVector", "norm", "normalize",
    "normalize_", "SparseMatrixCSC",
]
