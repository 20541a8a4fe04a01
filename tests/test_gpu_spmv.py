"""Parity of the HIP mul! (SpMV / SpMM, CSR / CSC / COO) with the oracle.

Tolerances (Float64): |y - y_ref| <= 1e-12 * Σ_j |a_ij||x_j| per entry; rows the
kernel sums on one lane (<= 32 stored entries in the staged window) and the
reference's order must match bit for bit.  Float32: 2e-6 normwise (row sums
here stay under 1000 entries, so any summation order stays inside
n*eps32*Σ|a||x| ~ 1e-4 and in practice ~1e-6).
"""
import ctypes

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
    short = np.diff(H.rowptr) <= 64  # one lane per (row, column), left to right
    assert np.array_equal(C[short], ref[short])


@pytest.mark.parametrize("k", [2, 3, 4, 5, 8, 16])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_spmm_rows_across_blocks(dev, k, dtype):
    """SpMM with rows around the one-lane and block limits, rows split over the
    workgroup and spilled past the staging, empty rows; B row-major with a
    leading dimension of k + 1, and column-major (transposed into scratch)."""
    npd = np.float64 if dtype == torch.float64 else np.float32
    rng = np.random.default_rng(61)
    lens = np.r_[[575, 1, 576, 33, 1152, 0, 32, 31, 1153, 577, 5000], rng.integers(0, 40, 2000)]
    H = synth.csr_from_lengths(lens, 9000, rng, "normal", dtype=npd)
    B = rng.standard_normal((H.n, k)).astype(npd)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval.astype(np.float64), B.astype(np.float64))
    bound = O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(H.nzval).astype(np.float64),
                         np.abs(B).astype(np.float64))
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    A = dev_csr(H, dtype=dtype)
    short = np.diff(H.rowptr) <= 32
    Bpad = np.zeros((H.n, k + 1), dtype=npd)
    Bpad[:, :k] = B
    for Bd in (G.dense(B, device=dev), torch.as_tensor(Bpad, device=dev)[:, :k]):
        C = host(A @ Bd)
        assert_normwise(C, ref, bound, tol)
        if dtype == torch.float64:
            assert np.array_equal(C[short], ref[short])


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.5, 0.0), (1.0, 1.0), (0.5, 2.5)])
def test_spmm_alpha_beta(dev, alpha, beta):
    rng = np.random.default_rng(62)
    H = synth.csr_from_lengths(np.r_[rng.integers(0, 40, 3000), [4000, 700, 513]], 5000, rng,
                               "normal")
    k = 4
    B = rng.standard_normal((H.n, k))
    C0 = rng.standard_normal((H.m, k))
    if beta == 0.0:
        C0[::5, :] = np.nan  # cleared by β == 0, long and spilled rows included
    A = dev_csr(H)
    Cd = G.dense(C0.copy(), device=dev)
    G.mul_(Cd, A, G.dense(B, device=dev), alpha, beta)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, B, C=C0, alpha=alpha, beta=beta)
    bound = abs(alpha) * O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(H.nzval), np.abs(B)) \
        + np.nan_to_num(np.abs(beta * C0))
    got = host(Cd)
    assert not np.isnan(got).any()
    assert_normwise(got, ref, bound, 1e-12)


@pytest.mark.parametrize("k", [3, 32, 70])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.5, 0.5), (2.0, 1.0)])
def test_spmm_column_major_through_row_scratch(dev, k, alpha, beta):
    """A column-major C of 2^21 outputs or more is written through row-major
    scratch and a tile transpose that applies α and β: the same bits as the
    row-major output of the same kernel, the oracle's bits on rows of at most 64
    entries, NaN in C cleared by β == 0, and a leading dimension larger than m kept."""
    rng = np.random.default_rng(63)
    bulk = (1 << 21) // k + 1000  # rows enough for m·k >= 2^21
    lens = np.r_[rng.integers(0, 12, bulk), [0, 65, 4000, 700, 2300, 1], rng.integers(0, 9, 300)]
    H = synth.csr_from_lengths(lens, 7000, rng, "normal")
    B = rng.standard_normal((H.n, k))
    ld = H.m + 3
    C0 = rng.standard_normal((H.m, k))
    if beta == 0.0:
        C0[::7, :] = np.nan
    A = dev_csr(H)
    big = torch.full((k, ld), 7.0, dtype=torch.float64, device=dev)
    big[:, :H.m] = torch.as_tensor(C0.T.copy(), device=dev)
    Cd = big.t()[:H.m]  # column-major, ldc = m + 3
    G.mul_(Cd, A, G.dense(B, device=dev), alpha, beta)
    got = host(Cd)
    assert torch.all(big[:, H.m:] == 7.0)  # the padding rows are not written
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, B, C=C0, alpha=alpha, beta=beta)
    bound = abs(alpha) * O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(H.nzval), np.abs(B)) \
        + np.nan_to_num(np.abs(beta * C0))
    assert not np.isnan(got).any()
    assert_normwise(got, ref, bound, 1e-12)
    short = np.diff(H.rowptr) <= 64
    assert np.array_equal(got[short], ref[short])
    # the row-major output of the same product (no scratch, same kernel): same bits
    from gsa_hip import _lib as L
    from gsa_hip.matrix import ptr, stream_of
    Cr = torch.as_tensor(C0.copy(), device=dev)
    s = stream_of(torch.device(dev))
    Bt = torch.as_tensor(B, device=dev)
    L.call("gsa_spmm_csr_ex", s, L.GSA_F64, L.GSA_I64, H.m, H.n, k, ptr(A.rowptr), ptr(A.colval),
           ptr(A.nzval), A.nnz(), ptr(Bt), k, L.GSA_OP_T, ptr(Cr), k, L.GSA_OP_T, alpha, beta,
           A.plan())
    assert np.array_equal(host(Cr), got)


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


# ---- rows across SpMV windows -----------------------------------------------
# The SpMV cuts the entries into fixed 576-entry windows; a row longer than 32
# entries that runs past its window is summed from partial sums (its owner's
# tail and each later window's head); a shorter one is finished by its owner
# from 32 entries staged past the window, so it stays bit-identical.

def _spmv(dev, A, x, planned):
    """y = A x through gsa_spmm_csr, with a plan (gsa_plan_create) or without."""
    import ctypes
    from gsa_hip import _lib as L
    from gsa_hip.matrix import stream_of
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    h = ctypes.c_void_p()
    st = stream_of(dev)
    it = L.GSA_I64 if A.rowptr.dtype == torch.int64 else L.GSA_I32
    if planned:
        L.call("gsa_plan_create", st, it, A.m, A.nnz(), p(A.rowptr), ctypes.byref(h))
    try:
        y = torch.empty(A.m, dtype=torch.float64, device=dev)
        L.call("gsa_spmm_csr", st, L.GSA_F64, it, L.GSA_OP_N, A.m, A.n, 1, p(A.rowptr),
               p(A.colval), p(A.nzval), A.nnz(), p(x), max(A.n, 1), L.GSA_OP_N, p(y), max(A.m, 1),
               None, None, h if planned else None)
        return host(y)
    finally:
        if planned:
            L.lib().gsa_plan_destroy(h)


@pytest.mark.parametrize("lengths", [
    [5000] + [1] * 3000,
    [2048, 1, 2303, 2304, 2305, 0, 4097, 3],
    [0] * 5000 + [7] + [0] * 5000,
    [100_000],
    [575, 1, 576, 33, 1152, 0, 32, 31, 1153, 577],  # lengths around the window and W_LONG
    [600] * 300,                                    # every row crosses a window end
    [32] * 4000,                                    # short rows crossing: finished by owners
    [33] * 4000,                                    # just long enough for head/tail sums
])
@pytest.mark.parametrize("planned", [True, False])
def test_rows_across_windows(dev, lengths, planned):
    rng = np.random.default_rng(25)
    H = synth.csr_from_lengths(np.array(lengths), 120_000, rng, "normal")
    x = rng.standard_normal(H.n)
    A = dev_csr(H)
    y = _spmv(dev, A, torch.as_tensor(x, device=dev), planned)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x)
    assert_normwise(y, ref, norm_bound(H.rowptr, H.colval, H.nzval, x), 1e-12)
    short = np.diff(H.rowptr) <= 32
    assert np.array_equal(y[short], ref[short])


def test_unsorted_columns(dev):
    """Column order inside a row is not required (the reference does not sort)."""
    rng = np.random.default_rng(27)
    H = synth.csr_from_lengths(np.array([3000, 5, 2500, 0, 900, 40] * 3), 80_000, rng, "normal")
    for r in range(H.m):  # shuffle every row's entries (columns and values together)
        a, b = H.rowptr[r] - 1, H.rowptr[r + 1] - 1
        perm = rng.permutation(b - a)
        H.colval[a:b] = H.colval[a:b][perm]
        H.nzval[a:b] = H.nzval[a:b][perm]
    x = rng.standard_normal(H.n)
    A = dev_csr(H)
    y = _spmv(dev, A, torch.as_tensor(x, device=dev), True)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x)
    assert_normwise(y, ref, norm_bound(H.rowptr, H.colval, H.nzval, x), 1e-12)


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.5, 0.0), (1.0, 1.0), (0.5, 2.5)])
def test_rows_across_windows_alpha_beta(dev, alpha, beta):
    rng = np.random.default_rng(24)
    H = synth.csr_from_lengths(np.r_[rng.integers(0, 40, 3000), [4000, 700, 513]], 20_000, rng,
                               "normal")
    x = rng.standard_normal(H.n)
    c0 = rng.standard_normal(H.m)
    if beta == 0.0:
        c0[::5] = np.nan  # cleared by β == 0 on every path, folded rows included
    A = dev_csr(H)
    C = torch.as_tensor(c0.copy(), device=dev)
    G.mul_(C, A, torch.as_tensor(x, device=dev), alpha, beta)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x, C=c0, alpha=alpha, beta=beta)
    # β C contributes |β c0| (0 where β == 0 clears a NaN)
    bound = abs(alpha) * norm_bound(H.rowptr, H.colval, H.nzval, x) + np.nan_to_num(np.abs(beta * c0))
    got = host(C)
    assert not np.isnan(got).any()
    assert_normwise(got, ref, bound, 1e-12)
    short = np.diff(H.rowptr) <= 32
    assert np.array_equal(got[short], ref[short])


def test_bf16_f32_rows_across_windows(dev):
    """Float32 and BF16-stored values through the same windows (TwoSum on the
    wavefront-reduced rows, per-row scale for BF16)."""
    from gsa_hip.precision import BF16CSR, mul_bf16
    rng = np.random.default_rng(28)
    H = synth.csr_from_lengths(np.r_[rng.integers(0, 30, 5000), [3000, 800]], 30_000, rng,
                               "normal", dtype=np.float32)
    x = rng.standard_normal(H.n).astype(np.float32)
    A = dev_csr(H, dtype=torch.float32)
    y = host(A @ torch.as_tensor(x, device=dev))
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval.astype(np.float64), x.astype(np.float64))
    bound = norm_bound(H.rowptr, H.colval, H.nzval, x)
    assert_normwise(y, ref, bound, 2e-6)
    B = BF16CSR(A)
    yb = torch.empty(H.m, dtype=torch.float32, device=dev)
    mul_bf16(yb, B, torch.as_tensor(x, device=dev))
    deq = B.dequantized_values()
    refb = O.spmm_csr_N(H.m, H.rowptr, H.colval, deq, x.astype(np.float64))
    assert_normwise(host(yb), refb, norm_bound(H.rowptr, H.colval, deq, x), 2e-6)


# ---- scatter-shaped products (CSR-T, CSC-N, COO) ------------------------------
# The reference adds with @atomic, so its order is undefined: tolerance only.

@pytest.mark.parametrize("lengths", [
    [5000] + [1] * 3000,
    [1] * 3000,                        # windows owning more than 256 rows (global search)
    [0] * 5000 + [7] + [0] * 5000,
    [575, 1, 576, 33, 1152, 0, 32, 31, 1153, 577],
    [100_000],
])
@pytest.mark.parametrize("k", [1, 3])
def test_transposed_scatter_windows(dev, lengths, k):
    rng = np.random.default_rng(31)
    H = synth.csr_from_lengths(np.array(lengths), 7_000, rng, "normal")
    vec = lambda a: a[:, 0] if k == 1 else a  # keeps a length-1 vector a vector
    xT = vec(rng.standard_normal((H.m, k)))
    A = dev_csr(H)
    Ac = dev_csc_of(H)
    d = torch.as_tensor(xT, device=dev) if k == 1 else G.dense(xT, device=dev)
    ref = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, H.nzval, xT)
    bound = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, np.abs(H.nzval), np.abs(xT))
    assert_normwise(host(G.transpose(A) @ d), ref, bound, 1e-12)
    xN = vec(rng.standard_normal((H.n, k)))
    dN = torch.as_tensor(xN, device=dev) if k == 1 else G.dense(xN, device=dev)
    refN = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, xN)
    bN = O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(H.nzval), np.abs(xN))
    assert_normwise(host(Ac @ dN), refN, bN, 1e-12)  # CSC-N scatters on the column arrays


@pytest.mark.parametrize("order", ["rowmajor", "shuffled"])
@pytest.mark.parametrize("k", [1, 4])
def test_coo_runs_duplicates_and_order(dev, order, k):
    """COO with long runs of one row, duplicate (row, col) pairs and any order:
    duplicates are summed (matrix_coo_kernels.jl:1-43)."""
    rng = np.random.default_rng(32)
    m, n = 3000, 2000
    lens = np.r_[[3000, 700], rng.integers(0, 12, m - 2)]
    rows = np.repeat(np.arange(1, m + 1), lens)
    cols = rng.integers(1, n + 1, rows.size)
    dup = rng.choice(rows.size, rows.size // 10, replace=False)
    rows = np.r_[rows, rows[dup]]
    cols = np.r_[cols, cols[dup]]
    if order == "rowmajor":
        o = np.argsort(rows, kind="stable")
    else:
        o = rng.permutation(rows.size)
    rows, cols = rows[o].astype(np.int64), cols[o].astype(np.int64)
    vals = rng.standard_normal(rows.size)
    A = G.GenericSparseMatrixCOO(m, n, torch.as_tensor(rows, device=dev),
                                 torch.as_tensor(cols, device=dev), torch.as_tensor(vals, device=dev))
    x = rng.standard_normal((n, k)).squeeze()
    d = torch.as_tensor(x, device=dev) if k == 1 else G.dense(x, device=dev)
    ref = O.spmm_coo(0, m, rows, cols, vals, x)
    bound = O.spmm_coo(0, m, rows, cols, np.abs(vals), np.abs(x))
    assert_normwise(host(A @ d), ref, bound, 1e-12)
    xt = rng.standard_normal((m, k)).squeeze()
    dt = torch.as_tensor(xt, device=dev) if k == 1 else G.dense(xt, device=dev)
    refT = O.spmm_coo(1, n, rows, cols, vals, xt)
    bT = O.spmm_coo(1, n, rows, cols, np.abs(vals), np.abs(xt))
    assert_normwise(host(G.transpose(A) @ dt), refT, bT, 1e-12)


def test_row_view_follows_value_updates(dev):
    """The COO/CSC row view holds positions, not values: products after an in-place
    value update use the new values; a changed index array rebuilds the view."""
    rng = np.random.default_rng(33)
    H = synth.powerlaw_csr(5000, 4000, alpha=1.846, seed=34, hi=700)
    A = dev_csr(H)
    Ac = G.GenericSparseMatrixCSC(A)
    Ao = G.GenericSparseMatrixCOO(Ac)  # column-major entry order: a sorted view with positions
    x = rng.standard_normal(H.n)
    xd = torch.as_tensor(x, device=dev)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x)
    bound = norm_bound(H.rowptr, H.colval, H.nzval, x)
    for M in (Ac, Ao):
        assert_normwise(host(M @ xd), ref, bound, 1e-12)
        M.nzval.mul_(2.0)
        assert_normwise(host(M @ xd), 2.0 * ref, 2.0 * bound, 1e-12)
        M.nzval.mul_(0.5)
    # permute the COO entries in place: the version bump rebuilds the view
    o = torch.as_tensor(rng.permutation(Ao.nnz()), device=dev)
    Ao.rowind.copy_(Ao.rowind[o]); Ao.colind.copy_(Ao.colind[o]); Ao.nzval.copy_(Ao.nzval[o])
    assert_normwise(host(Ao @ xd), ref, bound, 1e-12)


# ---- row bins: CSC-N and COO-N SpMV without atomics ------------------------------
# With a plan, y = A x for CSC (and for COO not in row order) streams the storage
# once into (row in bin, product) pairs grouped by 2048-row bins, then one
# workgroup per bin sums them in LDS (Float64) and writes its rows once.

def _flags(plan):
    import ctypes
    from gsa_hip import _lib as L
    f = ctypes.c_int(-1)
    L.call("gsa_plan_describe", plan, ctypes.byref(f))
    return f.value


def _csc_and_coo(H, rng, dev, dtype=torch.float64, itype=torch.int64):
    """The matrix as CSC, as COO in column-major order and as COO shuffled."""
    A = dev_csr(H, dtype=dtype, itype=itype)
    Ac = G.GenericSparseMatrixCSC(A)
    Acol = G.GenericSparseMatrixCOO(Ac)
    o = torch.as_tensor(rng.permutation(H.nnz), device=dev)
    Ashuf = G.GenericSparseMatrixCOO(H.m, H.n, Acol.rowind[o], Acol.colind[o], Acol.nzval[o])
    return {"csc": Ac, "coo_colmajor": Acol, "coo_shuffled": Ashuf}


@pytest.mark.parametrize("shape", [
    "powerlaw",      # 3 bins, the last one partial; rows up to 700 entries
    "empty_bins",    # whole bins without entries, rows only in the middle bin
    "one_row",       # a single row holding every entry
    "tile_edges",    # nnz one past two 4096-entry tiles
    "wide",          # more than 512 bins (the 2048-bin pass-1 layout)
])
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_row_bins_parity(dev, shape, itype):
    from gsa_hip import _lib as L
    rng = np.random.default_rng(51)
    if shape == "powerlaw":
        H = synth.powerlaw_csr(5000, 4000, alpha=1.846, seed=52, hi=700)
    elif shape == "empty_bins":
        lens = np.zeros(3 * 2048, dtype=np.int64)
        lens[2048:4096] = rng.integers(0, 30, 2048)
        H = synth.csr_from_lengths(lens, 5000, rng, "normal")
    elif shape == "one_row":
        H = synth.csr_from_lengths(np.array([20_000]), 30_000, rng, "normal")
    elif shape == "tile_edges":
        lens = np.full(2 * 4096 + 1, 1, dtype=np.int64)
        H = synth.csr_from_lengths(lens, 3000, rng, "normal")
    else:
        H = synth.csr_from_lengths(rng.integers(0, 4, 1_200_000), 50_000, rng, "normal")
    x = rng.standard_normal(H.n)
    xd = torch.as_tensor(x, device=dev)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x)
    bound = norm_bound(H.rowptr, H.colval, H.nzval, x)
    for name, M in _csc_and_coo(H, rng, dev, itype=itype).items():
        flags = _flags(M.plan())
        # a COO already in row order keeps the row view (pointer only)
        in_row_order = name != "csc" and bool((M.rowind[1:] >= M.rowind[:-1]).all())
        want = L.GSA_PLAN_ROW_VIEW if in_row_order else L.GSA_PLAN_ROW_BINS
        assert flags & want, (name, flags)
        y1 = host(M @ xd)
        y2 = host(M @ xd)
        assert np.array_equal(y1, y2), name  # fixed summation order
        assert_normwise(y1, ref, bound, 1e-12)


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.5, 0.0), (1.0, 1.0), (0.5, 2.5)])
def test_row_bins_alpha_beta(dev, alpha, beta):
    rng = np.random.default_rng(53)
    H = synth.powerlaw_csr(7000, 3000, alpha=1.846, seed=54, hi=900)
    x = rng.standard_normal(H.n)
    c0 = rng.standard_normal(H.m)
    if beta == 0.0:
        c0[::5] = np.nan  # cleared by β == 0, empty rows included
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x, C=c0, alpha=alpha, beta=beta)
    bound = abs(alpha) * norm_bound(H.rowptr, H.colval, H.nzval, x) + np.nan_to_num(np.abs(beta * c0))
    for name, M in _csc_and_coo(H, rng, dev).items():
        C = torch.as_tensor(c0.copy(), device=dev)
        G.mul_(C, M, torch.as_tensor(x, device=dev), alpha, beta)
        got = host(C)
        assert not np.isnan(got).any(), name
        assert_normwise(got, ref, bound, 1e-12)


def test_row_bins_float32_and_strided_x(dev):
    """Float32 values (accumulated in Float64 in the bins), and x read with a
    stride through the C ABI (opB = T, ldb = 3)."""
    import ctypes
    from gsa_hip import _lib as L
    from gsa_hip.matrix import stream_of
    rng = np.random.default_rng(55)
    H = synth.powerlaw_csr(6000, 5000, alpha=1.846, seed=56, hi=2000, dtype=np.float32)
    x = rng.standard_normal(H.n).astype(np.float32)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval.astype(np.float64), x.astype(np.float64))
    bound = norm_bound(H.rowptr, H.colval, H.nzval, x)
    Ms = _csc_and_coo(H, rng, dev, dtype=torch.float32)
    for name, M in Ms.items():
        assert_normwise(host(M @ torch.as_tensor(x, device=dev)), ref, bound, 2e-6)
    Ac = Ms["csc"]
    X3 = torch.zeros((H.n, 3), dtype=torch.float32, device=dev)
    X3[:, 1] = torch.as_tensor(x, device=dev)
    y = torch.empty(H.m, dtype=torch.float32, device=dev)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    L.call("gsa_spmm_csc", stream_of(dev), L.GSA_F32, L.GSA_I64, L.GSA_OP_N, H.m, H.n, 1,
           p(Ac.colptr), p(Ac.rowval), p(Ac.nzval), Ac.nnz(), ctypes.c_void_p(X3.data_ptr() + 4),
           3, L.GSA_OP_T, p(y), H.m, None, None, Ac.plan())
    assert_normwise(host(y), ref, bound, 2e-6)


def test_row_bins_too_many_rows_use_the_row_view(dev):
    """More rows than 2048 bins of 2048 hold: the plan keeps the row view."""
    from gsa_hip import _lib as L
    rng = np.random.default_rng(57)
    m = 2048 * 2048 + 5
    lens = np.zeros(m, dtype=np.int64)
    lens[rng.choice(m, 20_000, replace=False)] = rng.integers(1, 6, 20_000)
    H = synth.csr_from_lengths(lens, 1000, rng, "normal")
    x = rng.standard_normal(H.n)
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, x)
    bound = norm_bound(H.rowptr, H.colval, H.nzval, x)
    Ac = dev_csc_of(H)
    flags = _flags(Ac.plan())
    assert flags & L.GSA_PLAN_ROW_VIEW and not flags & L.GSA_PLAN_ROW_BINS
    assert_normwise(host(Ac @ torch.as_tensor(x, device=dev)), ref, bound, 1e-12)


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-0.5, 2.0)])
def test_transposed_bins(dev, itype, alpha, beta):
    """transpose(A) * x for CSR and COO through the plan's transposed bins (two
    streaming passes, no atomics); the same call without a plan still scatters."""
    import ctypes
    from gsa_hip import _lib as L
    from gsa_hip.matrix import stream_of
    rng = np.random.default_rng(61)
    H = synth.powerlaw_csr(9000, 7000, alpha=1.846, seed=62, hi=2500)
    x = rng.standard_normal(H.m)
    c0 = rng.standard_normal(H.n)
    if beta == 0.0:
        c0[::7] = np.nan
    ref = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, H.nzval, x, C=c0, alpha=alpha, beta=beta)
    bound = abs(alpha) * O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, np.abs(H.nzval), np.abs(x)) \
        + np.nan_to_num(np.abs(beta * c0))
    A = dev_csr(H, itype=itype)
    Ao = G.GenericSparseMatrixCOO(A)
    o = torch.as_tensor(rng.permutation(H.nnz), device=dev)
    Ashuf = G.GenericSparseMatrixCOO(H.m, H.n, Ao.rowind[o], Ao.colind[o], Ao.nzval[o])
    xd = torch.as_tensor(x, device=dev)
    for M in (A, Ao, Ashuf):
        C = torch.as_tensor(c0.copy(), device=dev)
        G.mul_(C, G.transpose(M), xd, alpha, beta)
        assert _flags(M._plan.handle) & L.GSA_PLAN_TRANSPOSE_BINS
        got = host(C)
        assert not np.isnan(got).any()
        assert_normwise(got, ref, bound, 1e-12)
        C2 = torch.as_tensor(c0.copy(), device=dev)
        G.mul_(C2, G.adjoint(M), xd, alpha, beta)
        assert np.array_equal(host(C2), got)  # fixed order: the same bits again
    # the scatter without a plan, and strided x (opB = T, ldb = 2) through the bins
    C = torch.as_tensor(c0.copy(), device=dev)
    L.call("gsa_spmm_csr", stream_of(dev), L.GSA_F64, G.matrix.itype_of(A.rowptr), L.GSA_OP_T,
           H.m, H.n, 1, G.matrix.ptr(A.rowptr), G.matrix.ptr(A.colval), G.matrix.ptr(A.nzval),
           H.nnz, G.matrix.ptr(xd), 1, L.GSA_OP_N, G.matrix.ptr(C), H.n,
           ctypes.byref(ctypes.c_double(alpha)), ctypes.byref(ctypes.c_double(beta)), None)
    assert_normwise(host(C), ref, bound, 1e-12)
    xs = torch.zeros((H.m, 2), dtype=torch.float64, device=dev)
    xs[:, 0] = xd
    C = torch.as_tensor(c0.copy(), device=dev)
    L.call("gsa_spmm_csr", stream_of(dev), L.GSA_F64, G.matrix.itype_of(A.rowptr), L.GSA_OP_T,
           H.m, H.n, 1, G.matrix.ptr(A.rowptr), G.matrix.ptr(A.colval), G.matrix.ptr(A.nzval),
           H.nnz, G.matrix.ptr(xs), 2, L.GSA_OP_T, G.matrix.ptr(C), H.n,
           ctypes.byref(ctypes.c_double(alpha)), ctypes.byref(ctypes.c_double(beta)),
           A.plan_transposed())
    assert_normwise(host(C), ref, bound, 1e-12)


def test_transposed_bins_float32(dev):
    rng = np.random.default_rng(63)
    H = synth.powerlaw_csr(5000, 6000, alpha=1.846, seed=64, hi=2000, dtype=np.float32)
    x = rng.standard_normal(H.m).astype(np.float32)
    A = dev_csr(H, dtype=torch.float32, itype=torch.int32)
    ref = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, H.nzval.astype(np.float64), x.astype(np.float64))
    bound = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, np.abs(H.nzval.astype(np.float64)),
                         np.abs(x.astype(np.float64)))
    got = host(G.transpose(A) @ torch.as_tensor(x, device=dev))
    assert_normwise(got, ref, bound, 2e-6)


@pytest.mark.parametrize("k", [1, 3, 16, 40])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_spmm_rowmajor_c_matches_columnmajor(dev, k, dtype):
    """gsa_spmm_csr_ex with a row-major C (opC = T, ldc >= k) gives the same bits as
    the column-major product, β modes included; the BF16 variant likewise."""
    from gsa_hip import _lib as L
    from gsa_hip.matrix import stream_of, ptr, itype_of, vtype_of
    from gsa_hip.precision import BF16CSR
    rng = np.random.default_rng(81)
    H = synth.powerlaw_csr(20_000, 15_000, alpha=1.846, seed=82, hi=3000)
    A = dev_csr(H, dtype=dtype)
    A.plan()
    B = G.dense(rng.standard_normal((H.n, k)), device=dev, dtype=dtype) if k > 1 else \
        torch.as_tensor(rng.standard_normal(H.n), device=dev, dtype=dtype)
    c0 = rng.standard_normal((H.m, k))
    for alpha, beta in ((1.0, 0.0), (0.5, -2.0), (1.0, 1.0)):
        Cc = G.dense(c0, device=dev, dtype=dtype) if k > 1 else \
            torch.as_tensor(c0[:, 0].copy(), device=dev, dtype=dtype)
        G.mul_(Cc, A, B, alpha, beta)
        ldc = k + 3 if k > 1 else 1
        Cr = torch.zeros((H.m, ldc), dtype=dtype, device=dev)
        Cr[:, :k] = torch.as_tensor(c0, dtype=dtype, device=dev)
        Bt, kk, ldb, opB = G.ops._dense_operand(B)
        L.call("gsa_spmm_csr_ex", stream_of(dev), vtype_of(A.nzval), itype_of(A.rowptr), H.m,
               H.n, k, ptr(A.rowptr), ptr(A.colval), ptr(A.nzval), H.nnz, ptr(Bt), ldb, opB,
               ptr(Cr), ldc, L.GSA_OP_T, alpha, beta, A.plan())
        got = host(Cr)[:, :k]
        exp = host(Cc).reshape(H.m, k)
        assert np.array_equal(got, exp)
        if k > 1:
            assert np.array_equal(host(Cr)[:, k:], np.zeros((H.m, 3)))  # padding untouched
    if dtype == torch.float32:
        Ab = BF16CSR(A)
        Cc = torch.zeros((k, H.m), dtype=torch.float32, device=dev).t() if k > 1 else \
            torch.zeros(H.m, dtype=torch.float32, device=dev)
        G.precision.mul_bf16(Cc, Ab, B, 1.0, 0.0)
        Cr = torch.zeros((H.m, max(k, 1)), dtype=torch.float32, device=dev)
        Bt, kk, ldb, opB = G.ops._dense_operand(B)
        L.call("gsa_spmm_csr_bf16", stream_of(dev), itype_of(A.rowptr), H.m, H.n, k,
               ptr(Ab.rowptr), ptr(Ab.colval), ptr(Ab.nzval), ptr(Ab.row_scale), H.nnz, ptr(Bt),
               ldb, opB, ptr(Cr), max(k, 1), L.GSA_OP_T, 1.0, 0.0, Ab.plan())
        assert np.array_equal(host(Cr), host(Cc).reshape(H.m, k))


@pytest.mark.parametrize("k", [3, 16, 40])
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_gather_views_for_scatter_products(dev, k, itype):
    """CSC * B, transpose(CSR) * B and COO (both directions, duplicates, shuffled
    storage) with several right-hand sides run through the plan's gather views;
    each matches the oracle's scatter within tolerance, α/β included, and the
    plan-less call (the atomic scatter) agrees."""
    from gsa_hip import _lib as L
    rng = np.random.default_rng(91)
    H = synth.powerlaw_csr(9000, 7000, alpha=1.846, seed=92, hi=2500)
    A = dev_csr(H, itype=itype)
    Ac = G.GenericSparseMatrixCSC(A)
    Ao = G.GenericSparseMatrixCOO(A)
    o = torch.as_tensor(rng.permutation(H.nnz), device=dev)
    dup = torch.as_tensor(rng.choice(H.nnz, 500), device=dev)
    Ashuf = G.GenericSparseMatrixCOO(H.m, H.n, torch.cat([Ao.rowind[o], Ao.rowind[dup]]),
                                     torch.cat([Ao.colind[o], Ao.colind[dup]]),
                                     torch.cat([Ao.nzval[o], Ao.nzval[dup] * 0.5]))
    r_sh, c_sh, v_sh = (host(t) for t in (Ashuf.rowind, Ashuf.colind, Ashuf.nzval))
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        # A * B: CSC and COO
        Bn = rng.standard_normal((H.n, k))
        c0 = rng.standard_normal((H.m, k))
        if beta == 0.0:
            c0[::7] = np.nan
        refN = O.spmm_csr_N(H.m, H.rowptr, H.colval, H.nzval, Bn, C=c0, alpha=alpha, beta=beta)
        bN = abs(alpha) * norm_bound(H.rowptr, H.colval, H.nzval, np.abs(Bn)) + np.nan_to_num(np.abs(beta * c0))
        refS = O.spmm_coo(False, H.m, r_sh, c_sh, v_sh, Bn, C=c0, alpha=alpha, beta=beta)
        for M, ref in ((Ac, refN), (Ao, refN), (Ashuf, refS)):
            C = G.dense(c0, device=dev)
            G.mul_(C, M, G.dense(Bn, device=dev), alpha, beta)
            assert _flags(M._plan.handle) & L.GSA_PLAN_GATHER_VIEW
            got = host(C)
            assert not np.isnan(got).any()
            assert_normwise(got, ref, bN if M is not Ashuf else np.maximum(bN, 1e-300) * 2, 1e-12)
        # transpose(A) * B: CSR and COO, B row-major (Julia's transpose of a k x m matrix)
        Bt = rng.standard_normal((H.m, k))
        ct = rng.standard_normal((H.n, k))
        refT = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, H.nzval, Bt, C=ct, alpha=alpha, beta=beta)
        bT = abs(alpha) * O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, np.abs(H.nzval), np.abs(Bt)) \
            + np.abs(beta * ct)
        Brow = torch.as_tensor(Bt, device=dev)  # row-major: read through opB = T
        for M in (A, Ao):
            C = G.dense(ct, device=dev)
            G.mul_(C, G.transpose(M), Brow, alpha, beta)
            assert _flags(M._plan.handle) & L.GSA_PLAN_GATHER_VIEW
            assert_normwise(host(C), refT, bT, 1e-12)
    # the plan-less product (atomic scatter) gives the same to tolerance
    C1 = G.dense(np.zeros((H.m, k)), device=dev)
    G.mul_(C1, Ac, G.dense(Bn, device=dev))
    C2 = G.dense(np.zeros((H.m, k)), device=dev)
    Bd = G.dense(Bn, device=dev)
    L.call("gsa_spmm_csc", G.matrix.stream_of(dev), L.GSA_F64, G.matrix.itype_of(Ac.colptr),
           L.GSA_OP_N, H.m, H.n, k, G.matrix.ptr(Ac.colptr), G.matrix.ptr(Ac.rowval),
           G.matrix.ptr(Ac.nzval), H.nnz, G.matrix.ptr(Bd), max(Bd.stride(1), H.n), L.GSA_OP_N,
           G.matrix.ptr(C2), H.m, None, None, None)
    assert_normwise(host(C1), host(C2), norm_bound(H.rowptr, H.colval, H.nzval, np.abs(Bn)), 1e-12)


def test_one_plan_serves_two_streams_at_once(dev):
    """A plan is read-only: the SpMVs of one matrix on two streams at the same
    time (one cached plan; CSR windows with long rows crossing window ends, and
    the CSC row bins) give the single-stream bits, as do two batched launches of
    one batch plan.  The reference's mul! may be called from any task."""
    from gsa_hip import batched
    rng = np.random.default_rng(301)
    H = synth.powerlaw_csr(300_000, 300_000, alpha=1.846, seed=302)
    A = dev_csr(H)
    Ac = G.GenericSparseMatrixCSC(A)
    xs = [torch.as_tensor(rng.standard_normal(H.n), device=dev) for _ in range(2)]
    ref = {(id(M), i): host(M @ xs[i]) for M in (A, Ac) for i in range(2)}
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for M in (A, Ac):
        outs = [[torch.empty(H.m, dtype=torch.float64, device=dev) for _ in range(6)]
                for _ in range(2)]
        for st in streams:
            st.wait_stream(torch.cuda.current_stream())
        for rep in range(6):  # interleaved launches keep both streams busy together
            for i, st in enumerate(streams):
                with torch.cuda.stream(st):
                    G.mul_(outs[i][rep], M, xs[i])
        torch.cuda.synchronize()
        for i in range(2):
            for rep in range(6):
                assert np.array_equal(host(outs[i][rep]), ref[(id(M), i)]), (type(M), i, rep)
    # one batch plan, two streams
    items = []
    for j in range(40):
        Hj = synth.powerlaw_csr(2000 + 100 * j, 2000 + 100 * j, alpha=1.846, seed=400 + j,
                                dtype=np.float32)
        Aj = dev_csr(Hj, dtype=torch.float32, itype=torch.int32)
        Bj = torch.as_tensor(rng.standard_normal((Hj.n, 4)).astype(np.float32), device=dev)
        items.append((Aj, Bj, torch.empty((Hj.m, 4), dtype=torch.float32, device=dev).t()
                      .contiguous().t()))
    bt = batched.BatchedSpMM(items)
    bt.run()
    torch.cuda.synchronize()
    want = [host(c).copy() for _, _, c in items]
    for rep in range(4):
        for st in streams:
            with torch.cuda.stream(st):
                bt.run()  # same values into the same outputs from both streams
    torch.cuda.synchronize()
    for (_, _, c), w in zip(items, want):
        assert np.array_equal(host(c), w)


def test_stream_destroy_releases_cached_scratch(dev):
    """gsa_stream_destroy waits for the stream, then frees the scratch cached for
    it: nothing stays cached under the dead handle."""
    import ctypes
    from gsa_hip import _lib as L
    lib = L.load()
    s = ctypes.c_void_p()
    L.call("gsa_stream_create", ctypes.byref(s))
    H = synth.powerlaw_csr(100_000, 100_000, alpha=1.846, seed=311)
    A = dev_csr(H)
    x = torch.ones(H.n, dtype=torch.float64, device=dev)
    y = torch.zeros(H.m, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    p = G.matrix.ptr
    L.call("gsa_spmm_csr", s, L.GSA_F64, L.GSA_I64, L.GSA_OP_N, H.m, H.n, 1, p(A.rowptr),
           p(A.colval), p(A.nzval), H.nnz, p(x), H.n, L.GSA_OP_N, p(y), H.m, None, None, None)
    L.call("gsa_stream_sync", s)
    assert lib.gsa_scratch_cached(s) > 0  # the window table and partial sums
    np.testing.assert_allclose(host(y), H.to_scipy() @ np.ones(H.n), rtol=1e-12, atol=1e-12)
    L.call("gsa_stream_destroy", s)
    assert lib.gsa_scratch_cached(s) == 0


def test_graph_capture_owns_its_scratch(dev):
    """A mul! captured into a HIP graph takes its scratch (window table, partial
    sums) as device memory owned by the graph: the executable keeps it after the
    capture's graph object is gone, replays equal the eager bits, and destroying
    the executable releases it."""
    from gsa_hip import _lib as L
    lib = L.load()
    H = synth.powerlaw_csr(200_000, 200_000, alpha=1.846, seed=321)
    A = dev_csr(H)
    x = torch.as_tensor(np.random.default_rng(322).standard_normal(H.n), device=dev)
    want = host(A @ x)
    y = torch.empty(H.m, dtype=torch.float64, device=dev)
    A.plan()
    torch.cuda.synchronize()
    base = lib.gsa_scratch_graph_blocks()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(3):
            G.mul_(y, A, x)
    torch.cuda.synchronize()
    held = lib.gsa_scratch_graph_blocks() - base
    # the executable still owns the blocks (checked before any replay)
    assert held >= 3, held
    for rep in range(4):
        y.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(host(y), want), rep
        assert np.array_equal(host(A @ x), want)  # eager work between replays
    del g
    import gc
    gc.collect()
    torch.cuda.synchronize()
    import time
    for _ in range(50):  # the release may run on a runtime thread
        if lib.gsa_scratch_graph_blocks() == base:
            break
        time.sleep(0.02)
    assert lib.gsa_scratch_graph_blocks() == base
