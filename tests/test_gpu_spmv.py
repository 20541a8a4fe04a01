"""Parity of the HIP mul! (SpMV / SpMM, CSR / CSC / COO) with the oracle.

Tolerances (Float64): |y - y_ref| <= 1e-12 * Σ_j |a_ij||x_j| per entry; rows the
kernel sums on one lane (<= 32 stored entries in the staged window) and the
reference's order must match bit for bit.  Float32: 2e-6 normwise (row sums
here stay under 1000 entries, so any summation order stays inside
n*eps32*Σ|a||x| ~ 1e-4 and in practice ~1e-6).
"""
import numpy as np
import pytest
import torch

import gsa_hip as G
import oracle as O
from gsa_hip import synth
from util import dev_csr, dev_csc_of, host, norm_bound, assert_normwise

pytestmark = pytest.mark.gpu


def test_doc_example(dev):
    A = G.GenericSparseMatrixCSR(3, 3, torch.tensor([1, 3, 4, 5], device=dev),
                                 torch.tensor([1, 2, 1, 3], device=dev),
                                 torch.tensor([2.0, 4.0, 3.0, 5.0], dtype=torch.float64, device=dev))
    x = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64, device=dev)
    assert host(A @ x).tolist() == [10.0, 3.0, 15.0]
    Acsc = G.GenericSparseMatrixCSC(A)
    assert host(Acsc @ x).tolist() == [10.0, 3.0, 15.0]
    Acoo = G.GenericSparseMatrixCOO(A)
    assert host(Acoo @ x).tolist() == [10.0, 3.0, 15.0]


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_config1_sprand_bitexact(dev, itype):
    H = synth.sprand(10_000, 10_000, 1e-3, seed=1)
    x = np.random.default_rng(1).random(10_000)
    A = dev_csr(H, itype=itype)
    y = host(A @ torch.as_tensor(x, device=dev))
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x)
    assert np.array_equal(y, ref)  # every row is short: the reference's order exactly
    np.testing.assert_allclose(y, H.to_scipy() @ x, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.5, 0.0), (1.0, 1.0), (0.5, 2.5), (2.0, -1.0)])
def test_alpha_beta_semantics(dev, alpha, beta):
    H = synth.sprand(3000, 2000, 4e-3, seed=5, values="normal")
    rng = np.random.default_rng(6)
    x = rng.standard_normal(2000)
    c0 = rng.standard_normal(3000)
    if beta == 0.0:
        c0[::7] = np.nan  # β == 0 clears NaN (fill!, matrix_csr.jl:192)
    A = dev_csr(H)
    C = torch.as_tensor(c0.copy(), device=dev)
    G.mul_(C, A, torch.as_tensor(x, device=dev), alpha, beta)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x, C=c0, alpha=alpha, beta=beta)
    assert np.array_equal(host(C), ref)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_powerlaw_parity(dev, dtype, itype):
    npd = np.float64 if dtype == torch.float64 else np.float32
    H = synth.powerlaw_csr(200_000, 200_000, alpha=1.846, seed=7, dtype=npd)
    x = np.random.default_rng(8).standard_normal(200_000).astype(npd)
    A = dev_csr(H, dtype=dtype, itype=itype)
    y = host(A @ torch.as_tensor(x, device=dev))
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x)
    bound = norm_bound(H.rowptr, H.colval, H.nzval, x)
    assert_normwise(y, ref, bound, 1e-12 if dtype == torch.float64 else 2e-6)
    lens = np.diff(H.rowptr)
    short = lens <= 32
    assert np.array_equal(y[short], ref[short])


def test_powerlaw_full_size_f64(dev):
    """1M x 1M truncated Zipf rows (alpha 1.846, mean 10), the benchmark matrix."""
    H = synth.powerlaw_csr(1_000_000, 1_000_000, alpha=1.846, seed=2)
    x = np.random.default_rng(2).standard_normal(1_000_000)
    A = dev_csr(H)
    y = host(A @ torch.as_tensor(x, device=dev))
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x, threads=O.max_threads())
    bound = norm_bound(H.rowptr, H.colval, H.nzval, x)
    worst = assert_normwise(y, ref, bound, 1e-12)
    print(f"max normalised error {worst:.3e}")


@pytest.mark.parametrize("lengths", [
    [0, 0, 0],                      # all rows empty
    [5000] + [1] * 3000,            # a long row that spills past the window
    [2048, 1, 2303, 2304, 2305, 0, 4097, 3],  # lengths around window and LDS limits
    [0] * 5000 + [7] + [0] * 5000,  # a block owning thousands of empty rows
    [60] * 1000,                    # rows on the one-wavefront path
    [100_000],                      # one very long row
])
def test_row_length_edges(dev, lengths):
    rng = np.random.default_rng(11)
    H = synth.csr_from_lengths(np.array(lengths), 120_000, rng, "normal")
    x = rng.standard_normal(H.n)
    A = dev_csr(H)
    y = host(A @ torch.as_tensor(x, device=dev))
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x)
    assert_normwise(y, ref, norm_bound(H.rowptr, H.colval, H.nzval, x), 1e-12)


def test_empty_matrices(dev):
    for m, n in ((0, 5), (5, 0), (0, 0), (4, 3)):
        A = G.GenericSparseMatrixCSR(m, n, torch.ones(m + 1, dtype=torch.int64, device=dev),
                                     torch.zeros(0, dtype=torch.int64, device=dev),
                                     torch.zeros(0, dtype=torch.float64, device=dev))
        y = A @ torch.ones(n, dtype=torch.float64, device=dev)
        assert y.shape == (m,) and (m == 0 or float(y.abs().max()) == 0.0)


@pytest.mark.parametrize("k", [2, 3, 8, 16, 32, 40])
@pytest.mark.parametrize("layout", ["colmajor", "rowmajor"])
def test_spmm_parity(dev, k, layout):
    H = synth.powerlaw_csr(30_000, 20_000, alpha=1.846, seed=12)
    rng = np.random.default_rng(13)
    B = rng.standard_normal((20_000, k))
    A = dev_csr(H)
    Bd = G.dense(B, device=dev) if layout == "colmajor" else torch.as_tensor(B, device=dev)
    C = host(A @ Bd)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, B)
    bound = O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(H.nzval), np.abs(B))
    assert_normwise(C, ref, bound, 1e-12)
    short = np.diff(H.rowptr) <= 64
    assert np.array_equal(C[short], ref[short])


def test_spmm_beta_into_strided_output(dev):
    H = synth.sprand(500, 400, 0.02, seed=14, values="normal")
    rng = np.random.default_rng(15)
    B = rng.standard_normal((400, 4))
    C0 = rng.standard_normal((500, 4))
    A = dev_csr(H)
    Cd = torch.as_tensor(C0.copy(), device=dev)  # row-major: written through a temp
    G.mul_(Cd, A, G.dense(B, device=dev), 2.0, 0.5)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, B, C=C0, alpha=2.0, beta=0.5)
    assert np.array_equal(host(Cd), ref)


@pytest.mark.parametrize("k", [1, 4])
def test_transpose_and_csc_coo_parity(dev, k):
    H = synth.powerlaw_csr(40_000, 30_000, alpha=1.846, seed=16)
    rng = np.random.default_rng(17)
    xN = rng.standard_normal((30_000, k)).squeeze()
    xT = rng.standard_normal((40_000, k)).squeeze()
    A = dev_csr(H)
    Ac = dev_csc_of(H)
    Ao = G.GenericSparseMatrixCOO(A)
    dN = torch.as_tensor(xN, device=dev) if k == 1 else G.dense(xN, device=dev)
    dT = torch.as_tensor(xT, device=dev) if k == 1 else G.dense(xT, device=dev)
    refN = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, xN)
    refT = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, H.nzval, xT)
    bN = O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(H.nzval), np.abs(xN))
    bT = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, np.abs(H.nzval), np.abs(xT))
    # CSR-T (scatter), CSC-N (scatter), CSC-T (gather), COO N/T (scatter)
    assert_normwise(host(G.transpose(A) @ dT), refT, bT, 1e-12)
    assert_normwise(host(G.adjoint(A) @ dT), refT, bT, 1e-12)
    assert_normwise(host(Ac @ dN), refN, bN, 1e-12)
    assert_normwise(host(G.transpose(Ac) @ dT), refT, bT, 1e-12)
    assert_normwise(host(Ao @ dN), refN, bN, 1e-12)
    assert_normwise(host(G.transpose(Ao) @ dT), refT, bT, 1e-12)


def test_plan_reuse_and_invalidation(dev):
    H = synth.powerlaw_csr(50_000, 50_000, seed=18)
    x = torch.as_tensor(np.random.default_rng(19).standard_normal(50_000), device=dev)
    A = dev_csr(H)
    y1 = host(A @ x)
    h1 = A._plan.handle.value
    y2 = host(A @ x)
    assert A._plan.handle.value == h1 and np.array_equal(y1, y2)
    A.rowptr[5:] += 0  # in-place write bumps the version: plan rebuilt
    y3 = host(A @ x)
    assert np.array_equal(y1, y3)


def test_int_overflowing_dims_rejected():
    with pytest.raises(G.ArgumentError):
        G.GenericSparseMatrixCSR(2 ** 31, 1, torch.ones(1, dtype=torch.int32),
                                 torch.zeros(0, dtype=torch.int32), torch.zeros(0))
