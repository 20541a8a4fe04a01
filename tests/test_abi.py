"""CPU-side checks of the boundary: the C-ABI library loads and exports every
symbol include/gsa_hip.h declares, and the host mirror raises the reference's
errors before any device work (no GPU needed)."""
import os
import re

import numpy as np
import pytest
import torch

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gsa_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsa_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_entry_points():
    names = declared_functions()
    assert "gsa_spmm_csr" in names and "gsa_coo_to_csr" in names and len(names) >= 30


def test_library_exports_every_declared_symbol():
    import gsa_hip._lib as L
    lib = L.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares exactly the header's entry points
    assert sorted(L.exported_symbols()) == declared_functions()


def test_version_and_error_string_need_no_gpu():
    import gsa_hip._lib as L
    lib = L.load()
    assert lib.gsa_version() >= 100
    assert isinstance(lib.gsa_last_error(), bytes)


def _csr_cpu():
    import gsa_hip as G
    rowptr = torch.tensor([1, 3, 4], dtype=torch.int64)
    colval = torch.tensor([1, 2, 1], dtype=torch.int64)
    nzval = torch.tensor([1.0, 3.0, 2.0], dtype=torch.float64)
    return G.GenericSparseMatrixCSR(2, 2, rowptr, colval, nzval)


def test_constructor_checks_mirror_reference():
    import gsa_hip as G
    A = _csr_cpu()
    assert A.shape == (2, 2) and A.nnz() == 3 and len(A) == 4
    # rowptr length must be m+1 (test/shared/matrix_csr.jl:61-67)
    with pytest.raises(G.ArgumentError):
        G.GenericSparseMatrixCSR(2, 2, torch.tensor([1, 3]), torch.tensor([1]),
                                 torch.tensor([1.0], dtype=torch.float64))
    with pytest.raises(G.ArgumentError):
        G.GenericSparseMatrixCSR(-1, 2, torch.tensor([1]), torch.tensor([], dtype=torch.int64),
                                 torch.tensor([], dtype=torch.float64))
    with pytest.raises(G.ArgumentError):
        G.GenericSparseMatrixCSC(2, 2, torch.tensor([1, 2, 3]), torch.tensor([1, 2]),
                                 torch.tensor([1.0], dtype=torch.float64))


def test_constructor_copies_buffers():
    import gsa_hip as G
    rowptr = torch.tensor([1, 2], dtype=torch.int64)
    A = G.GenericSparseMatrixCSR(1, 1, rowptr, torch.tensor([1]), torch.tensor([2.0]))
    rowptr[0] = 7
    assert int(A.rowptr[0]) == 1  # matrix_csr.jl:55-61 copies every buffer


def test_mul_dimension_mismatch_messages():
    import gsa_hip as G
    A = _csr_cpu()
    with pytest.raises(G.DimensionMismatch, match="second dimension of A, 2, does not match"):
        G.mul_(torch.zeros(2, dtype=torch.float64), A, torch.zeros(3, dtype=torch.float64))
    with pytest.raises(G.DimensionMismatch, match="first dimension of A, 2"):
        G.mul_(torch.zeros(3, dtype=torch.float64), A, torch.zeros(2, dtype=torch.float64))


def test_cpu_storage_is_rejected_not_computed():
    import gsa_hip as G
    A = _csr_cpu()
    with pytest.raises(G.ArgumentError, match="no CPU path|same backend"):
        G.mul_(torch.zeros(2, dtype=torch.float64), A, torch.ones(2, dtype=torch.float64))


def test_element_type_promotion_check():
    import gsa_hip as G
    A = _csr_cpu()
    with pytest.raises(G.ArgumentError, match="promotable"):
        G.mul_(torch.zeros(2, dtype=torch.float32), A, torch.ones(2, dtype=torch.float32))


def test_add_and_spgemm_dimension_checks():
    import gsa_hip as G
    A = _csr_cpu()
    B = G.GenericSparseMatrixCSR(3, 2, torch.tensor([1, 1, 1, 1]),
                                 torch.tensor([], dtype=torch.int64),
                                 torch.tensor([], dtype=torch.float64))
    with pytest.raises(G.DimensionMismatch, match="dimensions must match"):
        A + B
    with pytest.raises(G.DimensionMismatch, match="second dimension of A"):
        G.spgemm(A, B)


def test_synthetic_generators_shapes():
    from gsa_hip import synth
    A = synth.sprand(1000, 1000, 1e-2, seed=1)
    assert A.rowptr[0] == 1 and A.rowptr[-1] == A.nnz + 1 and A.colval.min() >= 1
    rows = np.diff(A.rowptr)
    for r in range(0, 1000, 97):  # distinct sorted columns per row
        c = A.colval[A.rowptr[r] - 1:A.rowptr[r + 1] - 1]
        assert np.all(np.diff(c) > 0)
    assert abs(synth.zipf_mean(1.846) - 10.0) < 0.01
    assert abs(synth.zipf_mean(2.0) - 5.53) < 0.01
    L = synth.laplacian2d(10)
    assert L.nnz == 5 * 100 - 4 * 10
    P = synth.powerlaw_csr(20000, 20000, seed=3)
    assert 1 <= np.diff(P.rowptr).min() and np.diff(P.rowptr).max() <= 5000


def test_uniform_scaling_and_diagonal_host_rules():
    """UniformScaling's element type joins A's as promote_type does; the checks of
    tr, dot, issymmetric and A + λI run on the host before any device work."""
    import numpy as np
    import gsa_hip as G
    assert G.I.dtype is None and G.UniformScaling(2.5).dtype == torch.float64
    assert (np.float32(0.5) * G.I).dtype == torch.float32
    assert (2 * G.I).dtype is None and (2 * G.I).value == 2.0
    assert (-(3.0 * G.I)).value == -3.0
    R = G.GenericSparseMatrixCSR(2, 3, torch.tensor([1, 2, 2]), torch.tensor([1]),
                                 torch.tensor([1.0], dtype=torch.float64))
    with pytest.raises(G.DimensionMismatch, match="square to compute the trace"):
        G.tr(R)
    with pytest.raises(G.DimensionMismatch, match="square to add UniformScaling"):
        R + G.I
    assert G.issymmetric(R) is False and G.ishermitian(G.transpose(R)) is False
    with pytest.raises(G.DimensionMismatch, match="first dimension of A, 2, does not match the "
                                                 "length of x, 3"):
        G.dot(torch.zeros(3, dtype=torch.float64), R, torch.zeros(3, dtype=torch.float64))
    with pytest.raises(G.DimensionMismatch, match="second dimension of A, 3"):
        G.dot(torch.zeros(2, dtype=torch.float64), R, torch.zeros(2, dtype=torch.float64))
    D = G.Diagonal(torch.ones(4, dtype=torch.float64))
    F = G.Diagonal.fill(2.0, 5, dtype=torch.float32, device="cpu")
    assert D.shape == (4, 4) and F.shape == (5, 5) and F.dtype == torch.float32
    assert np.array_equal(F.collect(), 2.0 * np.eye(5))
    with pytest.raises(G.ArgumentError):
        G.kron(D, F)
    with pytest.raises(G.ArgumentError, match="no CPU path|HIP kernel layer needs device"):
        G.tr(G.GenericSparseMatrixCSR(2, 2, torch.tensor([1, 2, 3]), torch.tensor([1, 2]),
                                      torch.tensor([1.0, 2.0], dtype=torch.float64)))
