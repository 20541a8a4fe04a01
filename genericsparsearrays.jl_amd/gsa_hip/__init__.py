"""gsa_hip — MI355X (gfx950) sparse kernel layer with GenericSparseArrays.jl's API.

Types:      GenericSparseMatrixCSR, GenericSparseMatrixCSC, GenericSparseMatrixCOO
Wrappers:   transpose / adjoint (lazy Transpose / Adjoint)
Operators:  mul_ (mul!), A @ x / A * x, A * B (SpGEMM), A + B, A - B, kron,
            dropzeros, dropzeros_, conversions through the constructors,
            tr, sum, dot(x, A, y), issymmetric, ishermitian, A ± λI (UniformScaling, I),
            kron with a Diagonal.
Every operator runs in libgsa_hip.so; there is no CPU fallback.
"""
from ._lib import ArgumentError, DimensionMismatch, GSAError, load as load_library
from .matrix import (GenericSparseMatrixCSR, GenericSparseMatrixCSC, GenericSparseMatrixCOO,
                     Transpose, Adjoint, transpose, adjoint, adapt, nnz, nonzeros, dense,
                     UniformScaling, I, Diagonal)
from .ops import (mul_, times, add, kron, dropzeros, dropzeros_, spgemm, scale, negate, tr,
                  sum_, dot, issymmetric, ishermitian, add_uniform_scaling)
from .ops import sum_ as sum  # Base.sum(A) (shadows the builtin in this namespace only)
from . import convert

__all__ = [
    "ArgumentError", "DimensionMismatch", "GSAError", "load_library",
    "GenericSparseMatrixCSR", "GenericSparseMatrixCSC", "GenericSparseMatrixCOO",
    "Transpose", "Adjoint", "transpose", "adjoint", "adapt", "nnz", "nonzeros", "dense",
    "mul_", "times", "add", "kron", "dropzeros", "dropzeros_", "spgemm", "scale", "negate",
    "convert", "UniformScaling", "I", "Diagonal", "tr", "sum_", "dot", "issymmetric",
    "ishermitian", "add_uniform_scaling",
]
