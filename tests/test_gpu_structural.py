"""Parity of the HIP conversions, +, dropzeros, kron and SpGEMM with the oracle.

Index arrays and patterns must match bit for bit; values too (every value is
either copied, one IEEE add/multiply, or — SpGEMM — summed in the reference's
order).  Inputs: the golden fixtures, seeded random matrices and, at full
size, the benchmark-scale matrices with size-independent checks.
"""
import os

import numpy as np
import pytest
import torch

import gsa_hip as G
import oracle as O
from gsa_hip import synth
from conftest import GOLDEN
from util import dev_csr, dev_csc_of, host

pytestmark = pytest.mark.gpu


def g(name):
    return np.load(os.path.join(GOLDEN, name))


def csr_tuple(A):
    return host(A.rowptr), host(A.colval), host(A.nzval)


def csc_tuple(A):
    return host(A.colptr), host(A.rowval), host(A.nzval)


def coo_tuple(A):
    return host(A.rowind), host(A.colind), host(A.nzval)


def same(got, exp):
    for a, b in zip(got, exp):
        assert a.shape == b.shape and np.array_equal(a, b), (a[:10], b[:10])


def from_arrays(cls, m, n, a, b, c, dev, itype=torch.int64):
    return cls(m, n, torch.as_tensor(a).to(dev, itype), torch.as_tensor(b).to(dev, itype),
               torch.as_tensor(c).to(dev), copy=False)


# ------------------------------------------------------------ conversions ----
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_conversions_golden(dev, itype):
    d = g("conv_random.npz")
    m, n = int(d["m"]), int(d["n"])
    csr = (d["csr_0"], d["csr_1"], d["csr_2"])
    csc = (d["csc_0"], d["csc_1"], d["csc_2"])
    rowmaj = (d["coo_rowmajor_0"], d["coo_rowmajor_1"], d["coo_rowmajor_2"])
    colmaj = (d["coo_colmajor_0"], d["coo_colmajor_1"], d["coo_colmajor_2"])
    A = from_arrays(G.GenericSparseMatrixCSR, m, n, *csr, dev, itype)
    Ac = from_arrays(G.GenericSparseMatrixCSC, m, n, *csc, dev, itype)
    same(coo_tuple(G.GenericSparseMatrixCOO(A)), rowmaj)
    same(coo_tuple(G.GenericSparseMatrixCOO(Ac)), colmaj)
    same(csc_tuple(G.GenericSparseMatrixCSC(A)), csc)
    same(csr_tuple(G.GenericSparseMatrixCSR(Ac)), csr)
    for coo in (rowmaj, colmaj):
        Ao = from_arrays(G.GenericSparseMatrixCOO, m, n, *coo, dev, itype)
        same(csr_tuple(G.GenericSparseMatrixCSR(Ao)), csr)
        same(csc_tuple(G.GenericSparseMatrixCSC(Ao)), csc)


def test_conversion_edges_and_storage_pin(dev):
    d = g("conv_edges.npz")
    for name in ("single", "fullrow", "fullcol"):
        csr = tuple(d[f"{name}_csr_{i}"] for i in range(3))
        csc = tuple(d[f"{name}_csc_{i}"] for i in range(3))
        Ac = from_arrays(G.GenericSparseMatrixCSC, 3, 4, *csc, dev)
        same(csr_tuple(G.GenericSparseMatrixCSR(Ac)), csr)
        assert np.array_equal(G.GenericSparseMatrixCSR(Ac).collect(), d[f"{name}_dense"])
    p = g("csr_storage_2x2.npz")  # test/shared/matrix_csr.jl:49-59
    Ao = from_arrays(G.GenericSparseMatrixCOO, 2, 2, p["coo_rows"], p["coo_cols"], p["coo_vals"], dev)
    same(csr_tuple(G.GenericSparseMatrixCSR(Ao)), (p["rowptr"], p["colval"], p["nzval"]))


def test_coo_to_csr_unsorted_duplicates_and_empty(dev):
    rng = np.random.default_rng(20)
    r = rng.integers(1, 301, size=5000)
    c = rng.integers(1, 201, size=5000)
    v = rng.standard_normal(5000)
    Ao = from_arrays(G.GenericSparseMatrixCOO, 300, 200, r, c, v, dev)
    same(csr_tuple(G.GenericSparseMatrixCSR(Ao)), O.coo_to_csr(300, 200, r, c, v))
    same(csc_tuple(G.GenericSparseMatrixCSC(Ao)), O.coo_to_csc(300, 200, r, c, v))
    E = from_arrays(G.GenericSparseMatrixCOO, 4, 3, np.zeros(0, np.int64), np.zeros(0, np.int64),
                    np.zeros(0), dev)
    Ecsr = G.GenericSparseMatrixCSR(E)
    assert host(Ecsr.rowptr).tolist() == [1] * 5 and Ecsr.nnz() == 0
    # storage already in (row, col) order, duplicates and empty rows included: the
    # sort is skipped and the arrays copied; the same result as the sort
    o = np.lexsort((c, r))
    rs, cs, vs = r[o], c[o], v[o]
    As = from_arrays(G.GenericSparseMatrixCOO, 300, 200, rs, cs, vs, dev)
    same(csr_tuple(G.GenericSparseMatrixCSR(As)), O.coo_to_csr(300, 200, rs, cs, vs))
    o = np.lexsort((r, c))  # column-major: sorted for COO -> CSC
    rc, cc, vc = r[o], c[o], v[o]
    Ac = from_arrays(G.GenericSparseMatrixCOO, 300, 200, rc, cc, vc, dev)
    same(csc_tuple(G.GenericSparseMatrixCSC(Ac)), O.coo_to_csc(300, 200, rc, cc, vc))
    same(csr_tuple(G.GenericSparseMatrixCSR(Ac)), O.coo_to_csr(300, 200, rc, cc, vc))


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_transposition_rows_across_tiles(dev, itype):
    """CSR <-> CSC finds each entry's row from the row pointer, 2048 entries per
    tile: empty rows (also in runs, and at tile starts), rows spanning several
    tiles, rows of one entry.  Index arrays bit-exact with the oracle's route
    through COO; the same rows through COO + of row-major operands."""
    rng = np.random.default_rng(71)
    lengths = rng.integers(0, 6, 20000)
    lengths[rng.random(lengths.size) < 0.4] = 0
    lengths[[100, 7000, 7001, 15000]] = [3000, 5000, 2048, 4096]
    lengths[7002:7400] = 0
    H = synth.csr_from_lengths(lengths, 6000, rng, "normal")
    m, n = H.m, H.n
    A = from_arrays(G.GenericSparseMatrixCSR, m, n, H.rowptr, H.colval, H.nzval, dev, itype)
    exp = O.csr_to_csc(m, n, H.rowptr, H.colval, H.nzval)
    Ac = G.GenericSparseMatrixCSC(A)
    same(csc_tuple(Ac), exp)
    same(csr_tuple(G.GenericSparseMatrixCSR(Ac)), (H.rowptr, H.colval, H.nzval))
    Ao = G.GenericSparseMatrixCOO(A)
    coo = coo_tuple(Ao)
    half = (coo[0], coo[1], coo[2] * 0.5)
    Bo = from_arrays(G.GenericSparseMatrixCOO, m, n, *half, dev, itype)
    same(coo_tuple(Ao + Bo), O.coo_add(m, coo, half))


def test_transposed_conversions(dev):
    H = synth.sprand(300, 200, 0.03, seed=21, values="normal")
    A = dev_csr(H)
    Ac = dev_csc_of(H)
    # Transpose(CSC) -> CSR is a reinterpretation (conversions.jl:106-123)
    At = G.GenericSparseMatrixCSR(G.transpose(Ac))
    assert At.shape == (200, 300)
    np.testing.assert_array_equal(At.collect(), H.to_scipy().toarray().T)
    # Transpose(CSR) -> CSC, Transpose(COO) -> CSR
    np.testing.assert_array_equal(G.GenericSparseMatrixCSC(G.transpose(A)).collect(),
                                  H.to_scipy().toarray().T)
    Ao = G.GenericSparseMatrixCOO(A)
    np.testing.assert_array_equal(G.GenericSparseMatrixCSR(G.transpose(Ao)).collect(),
                                  H.to_scipy().toarray().T)


def test_conversions_full_size_powerlaw(dev):
    """1M x 1M power-law matrix: CSR <-> CSC <-> COO round trips, bit-exact."""
    H = synth.powerlaw_csr(1_000_000, 1_000_000, alpha=1.846, seed=2)
    A = dev_csr(H)
    Ac = G.GenericSparseMatrixCSC(A)
    Ao = G.GenericSparseMatrixCOO(A)
    # round trips return the original arrays
    same(csr_tuple(G.GenericSparseMatrixCSR(Ac)), (H.rowptr, H.colval, H.nzval))
    same(csr_tuple(G.GenericSparseMatrixCSR(Ao)), (H.rowptr, H.colval, H.nzval))
    Aoc = G.GenericSparseMatrixCOO(Ac)  # column-major COO
    same(csr_tuple(G.GenericSparseMatrixCSR(Aoc)), (H.rowptr, H.colval, H.nzval))
    # CSC is the oracle's (checked on a 200k-row sample of the pipeline's pieces)
    cp, rv, v = csc_tuple(Ac)
    assert cp[0] == 1 and cp[-1] == H.nnz + 1 and np.all(np.diff(cp) >= 0)
    S = H.to_scipy().tocsc()
    S.sort_indices()
    assert np.array_equal(cp - 1, S.indptr) and np.array_equal(rv - 1, S.indices)
    assert np.array_equal(v, S.data)
    r, c, vv = coo_tuple(Ao)
    assert np.array_equal(np.diff(H.rowptr), np.bincount(r - 1, minlength=H.m))
    assert np.array_equal(c, H.colval) and np.array_equal(vv, H.nzval)


# ------------------------------------------------------------ + / dropzeros ---
def test_add_golden_csr_csc_coo(dev):
    d = g("add_random.npz")
    m, n = int(d["m"]), int(d["n"])
    A = from_arrays(G.GenericSparseMatrixCSR, m, n, d["A_rowptr"], d["A_colval"], d["A_nzval"], dev)
    B = from_arrays(G.GenericSparseMatrixCSR, m, n, d["B_rowptr"], d["B_colval"], d["B_nzval"], dev)
    C = A + B
    assert isinstance(C, G.GenericSparseMatrixCSR)
    same(csr_tuple(C), (d["C_rowptr"], d["C_colval"], d["C_nzval"]))
    Cc = G.GenericSparseMatrixCSC(A) + G.GenericSparseMatrixCSC(B)
    np.testing.assert_array_equal(Cc.collect(), C.collect())
    Ao = from_arrays(G.GenericSparseMatrixCOO, m, n, *(d[f"Acoo_{i}"] for i in range(3)), dev)
    Bo = from_arrays(G.GenericSparseMatrixCOO, m, n, *(d[f"Bcoo_{i}"] for i in range(3)), dev)
    same(coo_tuple(Ao + Bo), tuple(d[f"Ccoo_{i}"] for i in range(3)))


def test_add_overlap_and_self_cancel(dev):
    d = g("overlap_add.npz")
    m, n = int(d["m"]), int(d["n"])
    A = from_arrays(G.GenericSparseMatrixCSR, m, n, d["A_rowptr"], d["A_colval"], d["A_nzval"], dev)
    B = from_arrays(G.GenericSparseMatrixCSR, m, n, d["B_rowptr"], d["B_colval"], d["B_nzval"], dev)
    same(csr_tuple(A + B), (d["C_rowptr"], d["C_colval"], d["C_nzval"]))
    for M in (A, G.GenericSparseMatrixCSC(A), G.GenericSparseMatrixCOO(A)):
        Z = M - M  # nnz 0 (test/shared/matrix_csr.jl:379-381)
        assert Z.nnz() == 0 and not Z.collect().any()


def test_add_transposed_operands(dev):
    H1 = synth.sprand(60, 60, 0.1, seed=22, values="normal")
    H2 = synth.sprand(60, 60, 0.15, seed=23, values="normal")
    A, B = dev_csr(H1), dev_csr(H2)
    D1, D2 = H1.to_scipy().toarray(), H2.to_scipy().toarray()
    for opA, dA in ((lambda X: X, D1), (G.transpose, D1.T)):
        for opB, dB in ((lambda X: X, D2), (G.transpose, D2.T), (G.adjoint, D2.T)):
            C = opA(A) + opB(B)
            assert isinstance(C, G.GenericSparseMatrixCSR)
            np.testing.assert_array_equal(C.collect(), dA + dB)


def test_add_laplacian_perturbed_full_size(dev):
    """L + L' (L' = L with 10% of entries dropped) on the 1000x1000 grid."""
    L = synth.laplacian2d(1000)
    Lp = synth.drop_entries(L, 0.1, seed=3)
    A, B = dev_csr(L), dev_csr(Lp)
    C = A + B
    same(csr_tuple(C), O.csr_add(L.m, (L.rowptr, L.colval, L.nzval), (Lp.rowptr, Lp.colval, Lp.nzval)))
    D = A - B  # exercises dropzeros: only the dropped entries remain
    same(csr_tuple(D), O.csr_add(L.m, (L.rowptr, L.colval, L.nzval), (Lp.rowptr, Lp.colval, -Lp.nzval)))
    assert D.nnz() == L.nnz - Lp.nnz


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_add_long_rows_workgroup_merge(dev, itype):
    """Rows of more than 128 entries (A's and B's together) in blocks too large for
    LDS are merged by the whole workgroup (binary searches and scans) when their
    columns strictly increase, and by one thread otherwise; both equal the oracle's
    two-pointer merge (CSR) and stable run sums (COO from CSR) bit for bit."""
    rng = np.random.default_rng(61)
    H = synth.powerlaw_csr(4000, 6000, seed=6, hi=3000)
    P = synth.drop_entries(H, 0.3, seed=7)
    vals = P.nzval.copy()
    flip = rng.random(vals.size) < 0.5
    vals[flip] = -vals[flip]          # these cancel A's entries exactly
    vals[~flip] = rng.standard_normal(int((~flip).sum()))
    m, n = H.m, H.n
    A_t = (H.rowptr, H.colval, H.nzval)
    B_t = (P.rowptr, P.colval, vals)
    assert np.diff(H.rowptr).max() + np.diff(P.rowptr).max() > 128
    A = from_arrays(G.GenericSparseMatrixCSR, m, n, *A_t, dev, itype)
    B = from_arrays(G.GenericSparseMatrixCSR, m, n, *B_t, dev, itype)
    same(csr_tuple(A + B), O.csr_add(m, A_t, B_t))
    same(csr_tuple(A + A), O.csr_add(m, A_t, A_t))
    # the same sums as row-major COO operands (the row-merge path of COO +)
    coo = lambda T: (np.repeat(np.arange(1, T[0].size), np.diff(T[0])), T[1], T[2])
    Ao = from_arrays(G.GenericSparseMatrixCOO, m, n, *coo(A_t), dev, itype)
    Bo = from_arrays(G.GenericSparseMatrixCOO, m, n, *coo(B_t), dev, itype)
    same(coo_tuple(Ao + Bo), O.coo_add(m, coo(A_t), coo(B_t)))
    # a long row with repeated columns: the serial merge keeps the two-pointer semantics
    lens = np.r_[rng.integers(0, 8, 300), [1500], rng.integers(0, 8, 300)]
    rp = np.r_[0, np.cumsum(lens)] + 1
    cv = np.concatenate([np.sort(rng.integers(1, 700, k)) for k in lens])  # repeats allowed
    nz = rng.standard_normal(cv.size)
    R_t = (rp, cv, nz)
    R = from_arrays(G.GenericSparseMatrixCSR, lens.size, 700, *R_t, dev, itype)
    same(csr_tuple(R + R), O.csr_add(lens.size, R_t, R_t))
    Ro = from_arrays(G.GenericSparseMatrixCOO, lens.size, 700, *coo(R_t), dev, itype)
    same(coo_tuple(Ro + Ro), O.coo_add(lens.size, coo(R_t), coo(R_t)))


def test_dropzeros_csr_csc_coo(dev):
    rng = np.random.default_rng(24)
    H = synth.csr_from_lengths(np.r_[rng.integers(0, 20, 3000), [3000, 0, 5]], 5000, rng, "normal")
    H.nzval[rng.random(H.nnz) < 0.3] = 0.0
    A = dev_csr(H)
    exp = O.csr_dropzeros(H.m, H.rowptr, H.colval, H.nzval)
    same(csr_tuple(G.dropzeros(A)), exp)
    Ao = G.GenericSparseMatrixCOO(A)
    same(coo_tuple(G.dropzeros(Ao)), O.coo_dropzeros(*coo_tuple(Ao)))
    Ac = G.GenericSparseMatrixCSC(A)
    np.testing.assert_array_equal(G.dropzeros(Ac).collect(), H.to_scipy().toarray())
    # dropzeros! keeps the stored zeros of an all-zero matrix (matrix_csr.jl:698-703)
    Z = dev_csr(H)
    Z.nzval.zero_()
    G.dropzeros_(Z)
    assert Z.nnz() == H.nnz
    assert G.dropzeros(Z).nnz() == 0


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_dropzeros_block_edges(dev, itype, dtype):
    """dropzeros runs over blocks of 2048 stored entries: here the entries are
    exactly four blocks (so the plan has an empty fifth), the first block is all
    zeros, a row starts on a block boundary, zero runs cross block edges inside
    a 4096-entry row, and empty rows sit between and after, in CSR and CSC."""
    rng = np.random.default_rng(25)
    lens = np.array([2048, 1, 2047, 0, 0, 4096, 0, 0])
    assert lens.sum() == 4 * 2048
    ptr = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int64)
    n = 4200
    col = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int64) + 1
    val = rng.standard_normal(col.size).astype(dtype)
    if np.issubdtype(dtype, np.complexfloating):
        val = val + 1j * rng.standard_normal(col.size)
    val[:2048] = 0                                    # a block of zeros (row 1)
    val[2048 + 1 + 100:2048 + 1 + 2047] = 0           # row 3 keeps its first 100
    z = 4096 + np.arange(4096)
    val[z[(z % 2048 > 2000) | (z % 2048 < 50)]] = 0   # runs across block edges in row 6
    val[4096 + 7::13] = 0
    for M in (from_arrays(G.GenericSparseMatrixCSR, lens.size, n, ptr, col, val, dev, itype),):
        exp = O.csr_dropzeros(lens.size, ptr, col, val)
        D = G.dropzeros(M)
        same(csr_tuple(D), exp)
        assert D.rowptr.dtype == itype
        Mc = G.GenericSparseMatrixCSC(M)
        Dc = G.dropzeros(Mc)
        dense = np.zeros((lens.size, n), dtype=val.dtype)
        dense[np.repeat(np.arange(lens.size), lens), col - 1] = val
        np.testing.assert_array_equal(Dc.collect(), dense)
        assert Dc.nnz() == exp[1].size


# ------------------------------------------------------------ kron ---------
def test_kron_golden(dev):
    d = g("kron_random.npz")
    mA, nA, mB, nB = (int(d[k]) for k in ("mA", "nA", "mB", "nB"))
    A = from_arrays(G.GenericSparseMatrixCSR, mA, nA, d["A_rowptr"], d["A_colval"], d["A_nzval"], dev)
    B = from_arrays(G.GenericSparseMatrixCSR, mB, nB, d["B_rowptr"], d["B_colval"], d["B_nzval"], dev)
    C = G.kron(A, B)
    assert isinstance(C, G.GenericSparseMatrixCSR) and C.shape == (mA * mB, nA * nB)
    same(csr_tuple(C), (d["C_rowptr"], d["C_colval"], d["C_nzval"]))
    K = G.kron(G.GenericSparseMatrixCOO(A), G.GenericSparseMatrixCOO(B))
    same(coo_tuple(K), (d["Kcoo_0"], d["Kcoo_1"], d["Kcoo_2"]))
    Kc = G.kron(G.GenericSparseMatrixCSC(A), G.GenericSparseMatrixCSC(B))
    np.testing.assert_array_equal(Kc.collect(), C.collect())


def test_kron_transposed_and_noncanonical(dev):
    H1 = synth.sprand(20, 15, 0.2, seed=25, values="normal")
    H2 = synth.sprand(12, 10, 0.2, seed=26, values="normal")
    A, B = dev_csr(H1), dev_csr(H2)
    import scipy.sparse as sp
    for opA, sA in ((lambda X: X, H1.to_scipy()), (G.transpose, H1.to_scipy().T)):
        for opB, sB in ((lambda X: X, H2.to_scipy()), (G.adjoint, H2.to_scipy().T)):
            C = G.kron(opA(A), opB(B))
            np.testing.assert_array_equal(C.collect(), sp.kron(sA, sB).toarray())
    # a CSR operand with descending columns in a row takes the sort route; the
    # result is the reference's canonical order
    rp = np.array([1, 3, 4]); cv = np.array([3, 1, 2]); v = np.array([1.0, 2.0, 3.0])
    N = from_arrays(G.GenericSparseMatrixCSR, 2, 3, rp, cv, v, dev)
    C = G.kron(N, B)
    same(csr_tuple(C), O.kron_csr(2, 3, H2.m, H2.n, (rp, cv, v), (H2.rowptr, H2.colval, H2.nzval)))


def test_kron_full_size(dev):
    """Two 1000x1000 matrices at density 0.01: C has ~1e8 entries (Int64)."""
    H1 = synth.sprand(1000, 1000, 0.01, seed=31)
    H2 = synth.sprand(1000, 1000, 0.01, seed=32)
    A, B = dev_csr(H1), dev_csr(H2)
    C = G.kron(A, B)
    assert C.nnz() == H1.nnz * H2.nnz
    rp = host(C.rowptr)
    # row (iA, iB) length is lenA(iA) * lenB(iB); rows strictly ascending
    lens = np.outer(np.diff(H1.rowptr), np.diff(H2.rowptr)).ravel()
    assert rp[0] == 1 and np.array_equal(np.diff(rp), lens)
    cv = C.colval
    same_row = torch.ones(cv.numel() - 1, dtype=torch.bool, device=dev)
    starts = torch.as_tensor(rp[1:-1] - 1, device=dev)
    starts = starts[(starts > 0) & (starts < cv.numel())]
    same_row[starts - 1] = False
    assert bool(((cv[1:] > cv[:-1]) | ~same_row).all())
    # spot-check rows against the closed form
    for i in (0, 12345, 500000, 999999):
        iA, iB = divmod(i, 1000)
        a = slice(H1.rowptr[iA] - 1, H1.rowptr[iA + 1] - 1)
        b = slice(H2.rowptr[iB] - 1, H2.rowptr[iB + 1] - 1)
        ec = ((H1.colval[a] - 1)[:, None] * 1000 + H2.colval[b][None, :]).ravel()
        ev = (H1.nzval[a][:, None] * H2.nzval[b][None, :]).ravel()
        got_c = host(C.colval[rp[i] - 1:rp[i + 1] - 1])
        got_v = host(C.nzval[rp[i] - 1:rp[i + 1] - 1])
        assert np.array_equal(got_c, ec) and np.array_equal(got_v, ev)


# ------------------------------------------------------------ SpGEMM --------
def test_spgemm_golden_and_formats(dev):
    d = g("spgemm_random.npz")
    m, k, n = int(d["m"]), int(d["k"]), int(d["n"])
    A = from_arrays(G.GenericSparseMatrixCSR, m, k, d["A_rowptr"], d["A_colval"], d["A_nzval"], dev)
    B = from_arrays(G.GenericSparseMatrixCSR, k, n, d["B_rowptr"], d["B_colval"], d["B_nzval"], dev)
    C = A * B
    assert isinstance(C, G.GenericSparseMatrixCSR)
    same(csr_tuple(C), (d["C_rowptr"], d["C_colval"], d["C_nzval"]))
    Cc = G.GenericSparseMatrixCSC(A) * G.GenericSparseMatrixCSC(B)
    np.testing.assert_array_equal(Cc.collect(), C.collect())
    Co = G.GenericSparseMatrixCOO(A) * G.GenericSparseMatrixCOO(B)
    np.testing.assert_array_equal(Co.collect(), C.collect())
    for opA in (G.transpose, G.adjoint):
        At = G.GenericSparseMatrixCSR(G.transpose(G.GenericSparseMatrixCSC(A)))  # Aᵀ stored
        np.testing.assert_array_equal((opA(At) * B).collect(), C.collect())


def test_spgemm_cancellation_zero_kept(dev):
    A = from_arrays(G.GenericSparseMatrixCSR, 1, 2, np.array([1, 3]), np.array([1, 2]),
                    np.array([1.0, 1.0]), dev)
    B = from_arrays(G.GenericSparseMatrixCSR, 2, 1, np.array([1, 2, 3]), np.array([1, 1]),
                    np.array([1.0, -1.0]), dev)
    C = A * B
    assert C.nnz() == 1 and host(C.nzval)[0] == 0.0


@pytest.mark.parametrize("which", ["LL_small", "RR_small", "powerlaw"])
def test_spgemm_random_vs_oracle(dev, which):
    if which == "LL_small":
        H1 = H2 = synth.laplacian2d(60)
    elif which == "RR_small":
        H1 = H2 = synth.fixed_row_csr(5000, 5000, 8, seed=27)
    else:
        H1 = synth.powerlaw_csr(3000, 3000, alpha=2.0, seed=28, hi=300)
        H2 = synth.powerlaw_csr(3000, 3000, alpha=2.0, seed=29, hi=300)
    C = dev_csr(H1) * dev_csr(H2)
    same(csr_tuple(C), O.spgemm_csr(H1.m, H2.n, (H1.rowptr, H1.colval, H1.nzval),
                                    (H2.rowptr, H2.colval, H2.nzval)))


def test_spgemm_laplacian_full_size(dev):
    """L*L on the 1000x1000 grid: 12,980,004 entries, bit-exact with the oracle."""
    L = synth.laplacian2d(1000)
    A = dev_csr(L)
    C = A * A
    assert C.nnz() == 12_980_004
    same(csr_tuple(C), O.spgemm_csr(L.m, L.n, (L.rowptr, L.colval, L.nzval),
                                    (L.rowptr, L.colval, L.nzval)))


@pytest.mark.parametrize("case", ["wave_and_block", "fallback_over_2048", "long_A_rows_empty_B",
                                  "empty_rows"])
def test_spgemm_row_paths(dev, case):
    """Each output row is sorted in a wavefront (<= 64 products), a workgroup
    (<= 2048), or the global path (any row over 2048): all bit-exact with the
    oracle's Gustavson order."""
    rng = np.random.default_rng(41)
    if case == "wave_and_block":
        H1 = synth.csr_from_lengths(np.r_[rng.integers(0, 6, 400), [30, 60, 64, 65, 90]], 300, rng, "normal")
        H2 = synth.csr_from_lengths(rng.integers(0, 25, 300), 500, rng, "normal")
    elif case == "fallback_over_2048":
        H1 = synth.csr_from_lengths(np.r_[rng.integers(0, 5, 200), [150]], 300, rng, "normal")
        H2 = synth.csr_from_lengths(rng.integers(10, 40, 300), 700, rng, "normal")
    elif case == "long_A_rows_empty_B":
        H1 = synth.csr_from_lengths(np.r_[[250, 200, 3], rng.integers(0, 4, 50)], 300, rng, "normal")
        lens = np.zeros(300, dtype=np.int64)
        lens[::37] = rng.integers(1, 9, lens[::37].size)
        H2 = synth.csr_from_lengths(lens, 400, rng, "normal")
    else:
        H1 = synth.csr_from_lengths(np.r_[[0] * 50, [3, 0, 5], [0] * 50], 40, rng, "normal")
        H2 = synth.csr_from_lengths(np.r_[[0] * 20, rng.integers(0, 6, 20)], 30, rng, "normal")
    C = dev_csr(H1) * dev_csr(H2)
    same(csr_tuple(C), O.spgemm_csr(H1.m, H2.n, (H1.rowptr, H1.colval, H1.nzval),
                                    (H2.rowptr, H2.colval, H2.nzval)))


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_scipy_ingest_transposes_on_the_device(dev, itype):
    """GenericSparseMatrixCSR(::SparseMatrixCSC) (conversions.jl:25-30): the CSC
    arrays as given (rows unsorted within a column, a duplicate) transposed on the
    device; the result is the oracle's CSC -> CSR of the same arrays."""
    import scipy.sparse as sp
    rng = np.random.default_rng(71)
    S = sp.random(300, 200, density=0.05, random_state=rng, format="csc")
    S = sp.csc_matrix((S.data, S.indices, S.indptr), shape=S.shape)
    for c in range(0, 200, 3):  # reverse the rows of every third column
        a, b = S.indptr[c], S.indptr[c + 1]
        S.indices[a:b] = S.indices[a:b][::-1].copy()
        S.data[a:b] = S.data[a:b][::-1].copy()
    # a duplicate (row, column) pair: column 1's first entry stored twice
    a = S.indptr[1]
    ind = np.insert(S.indices, a, S.indices[a])
    dat = np.insert(S.data, a, 0.25)
    ptr_ = S.indptr.copy()
    ptr_[2:] += 1
    S = sp.csc_matrix((dat, ind, ptr_), shape=(300, 200))
    S.has_sorted_indices = False
    cp, rv, nz = S.indptr.astype(np.int64) + 1, S.indices.astype(np.int64) + 1, S.data.copy()
    assert nz.size == ind.size  # scipy kept the duplicate
    A = G.convert.csr_from_scipy(S, device=dev, index_dtype=itype)
    same(csr_tuple(A), O.csc_to_csr(300, 200, cp, rv, nz))
    Ac = G.convert.csc_from_scipy(S, device=dev, index_dtype=itype)
    same(csc_tuple(Ac), (cp, rv, nz))  # the arrays as given (conversions.jl:7-8)
    R = sp.random(250, 180, density=0.05, random_state=rng, format="csr")
    Acr = G.convert.csc_from_scipy(R, device=dev, index_dtype=itype)
    rp, cv, nzr = R.indptr.astype(np.int64) + 1, R.indices.astype(np.int64) + 1, R.data.copy()
    same(csc_tuple(Acr), O.csr_to_csc(250, 180, rp, cv, nzr))


def test_index_shift_widens_narrows_and_checks(dev):
    """gsa_index_shift: zero-based host arrays made one-based on the device and
    back, Int32 <-> Int64, in place when the types match; an Int32 output that
    cannot hold a value is an ArgumentError before anything is returned."""
    from gsa_hip import _lib as L
    from gsa_hip.matrix import itype_of, ptr, stream_of
    rng = np.random.default_rng(5)
    a = rng.integers(0, 2**31 - 2, 100_003).astype(np.int64)

    def shift(t, out_dtype, delta, out=None):
        out = torch.empty(t.numel(), dtype=out_dtype, device=dev) if out is None else out
        L.call("gsa_index_shift", stream_of(dev), itype_of(t), itype_of(out), t.numel(), ptr(t),
               ptr(out), delta)
        return out

    t64 = torch.as_tensor(a, device=dev)
    t32 = shift(t64, torch.int32, 1)                 # narrowed, in range
    assert np.array_equal(host(t32), (a + 1).astype(np.int32))
    w = shift(t32, torch.int64, -1)                  # widened back
    assert np.array_equal(host(w), a)
    shift(w, torch.int64, 7, out=w)                  # in place
    assert np.array_equal(host(w), a + 7)
    big = torch.as_tensor(np.array([1, 2**31 - 1], dtype=np.int64), device=dev)
    with pytest.raises(G.ArgumentError):
        shift(big, torch.int32, 1)                   # 2^31 does not fit Int32
    shift(torch.empty(0, dtype=torch.int64, device=dev), torch.int32, 1)  # empty: no-op


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
@pytest.mark.parametrize("vdtype", [np.float64, np.float32, np.complex128, np.int32])
def test_host_egress_and_ingest_round_trip(dev, itype, vdtype):
    """SparseMatrixCSC(A) (conversions.jl:10-79) for every format and wrapper
    equals scipy's canonical CSC of the same matrix bit for bit; to_scipy gives
    back the stored arrays; ingest from scipy CSR/CSC/COO reproduces them."""
    import scipy.sparse as sp
    rng = np.random.default_rng(17)
    m, n = 700, 450
    S = sp.random(m, n, density=0.02, random_state=rng, format="csr")
    if np.issubdtype(vdtype, np.complexfloating):
        S.data = (S.data + 1j * rng.standard_normal(S.nnz)).astype(vdtype)
    elif np.issubdtype(vdtype, np.integer):
        S.data = rng.integers(-9, 10, S.nnz).astype(vdtype)
    else:
        S.data = S.data.astype(vdtype)
    S.sort_indices()
    ref = S.tocsc()
    ref.sort_indices()

    def same_csc(C, R):
        assert C.shape == R.shape and C.format == "csc"
        assert np.array_equal(C.indptr, R.indptr) and np.array_equal(C.indices, R.indices)
        assert np.array_equal(C.data, R.data) and C.data.dtype == R.data.dtype

    A = G.convert.csr_from_scipy(S, device=dev, index_dtype=itype)
    assert A.rowptr.dtype == itype and A.colval.dtype == itype
    back = G.convert.to_scipy(A)
    assert np.array_equal(back.indptr, S.indptr) and np.array_equal(back.indices, S.indices)
    assert np.array_equal(back.data, S.data)
    same_csc(G.SparseMatrixCSC(A), ref)                               # CSR: device transpose
    Ac = G.convert.csc_from_scipy(ref, device=dev, index_dtype=itype)
    same_csc(G.SparseMatrixCSC(Ac), ref)                              # CSC: its arrays
    Ao = G.convert.coo_from_scipy(S, device=dev, index_dtype=itype)
    same_csc(G.SparseMatrixCSC(Ao), ref)                              # COO: sparse(...)
    Tt = S.T.tocsc()
    Tt.sort_indices()
    same_csc(G.SparseMatrixCSC(G.transpose(A)), Tt)
    Ta = S.conj().T.tocsc()
    Ta.sort_indices()
    same_csc(G.SparseMatrixCSC(G.adjoint(A)), Ta)
    same_csc(G.SparseMatrixCSC(G.transpose(Ac)), Tt)
    same_csc(G.SparseMatrixCSC(G.adjoint(Ao)), Ta)
    # COO with duplicates: sparse() adds them (in storage order)
    r = np.array([0, 2, 0, 1, 0]); c = np.array([1, 0, 1, 1, 1])
    v = np.array([1, 2, 3, 4, 5]).astype(vdtype)
    D = G.GenericSparseMatrixCOO(3, 2, torch.as_tensor(r + 1).to(dev, itype),
                                 torch.as_tensor(c + 1).to(dev, itype), torch.as_tensor(v).to(dev))
    exp = sp.coo_matrix((v, (r, c)), shape=(3, 2)).tocsc()
    exp.sort_indices()
    same_csc(G.SparseMatrixCSC(D), exp)
    assert np.array_equal(G.convert.to_scipy(D).row, r)               # storage order kept


def _rows_to_csr(rows):
    """CSR arrays (one-based) from a list of (columns, values) per row."""
    lens = np.array([len(c) for c, _ in rows], dtype=np.int64)
    rp = np.r_[0, np.cumsum(lens)] + 1
    cv = np.concatenate([np.asarray(c, np.int64) for c, _ in rows] + [np.zeros(0, np.int64)])
    nz = np.concatenate([np.asarray(v, np.float64) for _, v in rows] + [np.zeros(0)])
    return rp.astype(np.int64), cv, nz


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_add_tiles_chunks_and_fallbacks(dev, itype):
    """The row merge works in row blocks (mean rows of up to 12 entries) or in
    even tiles of 1792 entries-plus-twice-the-rows (longer mean rows); either
    way it merges runs of whole rows of at most 2048 staged entries, longer rows
    in chunks of 2048 output-order positions, and small blocks of short rows one
    row per thread.  Rows just under and over those sizes, rows that need
    several runs, long rows whose equal-column pairs straddle chunk boundaries
    (some cancelling), runs of empty rows longer than a tile, and rows with
    repeated or unordered columns (merged one per thread, in a staged run and as
    a long row), through both kernels, all equal the oracle's two-pointer merge
    (CSR) and run sums (row-major COO) bit for bit."""
    rng = np.random.default_rng(77)
    n = 30_000
    A_rows, B_rows = [], []

    def pair(la, lb, overlap):
        ca = np.sort(rng.choice(np.arange(2, n + 1), la, replace=False))
        keep = rng.random(la) < overlap
        extra = np.setdiff1d(rng.choice(np.arange(2, n + 1), lb, replace=False), ca)
        cb = np.unique(np.r_[ca[keep][:lb], extra[:max(0, lb - int(keep[:lb].sum()))]])
        va = rng.standard_normal(la)
        vb = rng.standard_normal(cb.size)
        shared = np.isin(cb, ca)
        idx = np.searchsorted(ca, cb[shared])
        cancel = rng.random(idx.size) < 0.3
        vb_sh = vb[shared]
        vb_sh[cancel] = -va[idx[cancel]]
        vb[shared] = vb_sh
        A_rows.append((ca, va))
        B_rows.append((cb, vb))

    def empty(k):
        for _ in range(k):
            A_rows.append(([], []))
            B_rows.append(([], []))

    for _ in range(300):  # short rows
        pair(int(rng.integers(0, 12)), int(rng.integers(0, 12)), 0.5)
    for la, lb in ((600, 423), (512, 512), (1023, 0), (0, 1024), (1024, 1023), (1100, 1000),
                   (895, 896), (896, 896), (1791, 0), (0, 1792), (1000, 1049), (1024, 1025)):
        pair(la, lb, 0.6)
    for _ in range(12):  # rows of 300-900 entries: several sub-tiles per tile
        pair(int(rng.integers(150, 450)), int(rng.integers(150, 450)), 0.5)
    empty(2500)
    for _ in range(50):
        pair(int(rng.integers(0, 6)), int(rng.integers(0, 6)), 0.5)
    # long rows whose every A column is in B, B starting one column lower: the
    # output order alternates B, A, B, A, ... so the pairs straddle chunk edges
    for la in (2047, 3000, 5000):
        ca = np.sort(rng.choice(np.arange(3, n + 1), la, replace=False))
        cb = np.r_[2, ca]
        va = rng.standard_normal(la)
        vb = rng.standard_normal(cb.size)
        vb[1::7] = -va[::7]  # cancels at some of the straddling pairs
        A_rows.append((ca, va))
        B_rows.append((cb, vb))
        pair(int(rng.integers(0, 9)), int(rng.integers(0, 9)), 0.5)
    pair(9000, 7000, 0.4)  # many chunks in one row
    empty(3)
    # one short row with an unordered pair and one long row with a repeated column:
    # their runs (and that long row) are merged one row per thread
    C_rows = [(c.copy(), v.copy()) for c, v in A_rows]
    i_short = next(i for i, (c, _) in enumerate(C_rows) if 4 <= len(c) <= 11)
    c = C_rows[i_short][0]
    c[1], c[2] = c[2], c[1]
    i_long = next(i for i, (c, _) in enumerate(C_rows) if len(c) >= 3000)
    c = C_rows[i_long][0]
    c[100] = c[99]
    coo = lambda T: (np.repeat(np.arange(1, T[0].size), np.diff(T[0])), T[1], T[2])
    # as built (a mean of about 20 entries per row: the even tiles) and with
    # 10000 empty rows after them (a mean of about 5: the row blocks)
    for extra in (0, 10_000):
        pad = [([], [])] * extra
        A_t, B_t, C_t = (_rows_to_csr(R + pad) for R in (A_rows, B_rows, C_rows))
        m = len(A_rows) + extra
        A = from_arrays(G.GenericSparseMatrixCSR, m, n, *A_t, dev, itype)
        B = from_arrays(G.GenericSparseMatrixCSR, m, n, *B_t, dev, itype)
        same(csr_tuple(A + B), O.csr_add(m, A_t, B_t))
        same(csr_tuple(B + A), O.csr_add(m, B_t, A_t))
        same(csr_tuple(A - A), O.csr_add(m, A_t, (A_t[0], A_t[1], -A_t[2])))
        Ao = from_arrays(G.GenericSparseMatrixCOO, m, n, *coo(A_t), dev, itype)
        Bo = from_arrays(G.GenericSparseMatrixCOO, m, n, *coo(B_t), dev, itype)
        same(coo_tuple(Ao + Bo), O.coo_add(m, coo(A_t), coo(B_t)))
        Cm = from_arrays(G.GenericSparseMatrixCSR, m, n, *C_t, dev, itype)
        same(csr_tuple(Cm + B), O.csr_add(m, C_t, B_t))
        same(csr_tuple(B + Cm), O.csr_add(m, B_t, C_t))


def test_add_powerlaw_full_size(dev):
    """A + transpose(A) on the 1M-row power-law matrix (rows up to 5000 entries,
    Float64/Int64), bit for bit against the oracle's two-pointer merge."""
    H = synth.powerlaw_csr(1_000_000, 1_000_000, alpha=1.846, seed=2)
    A = dev_csr(H)
    Ac = G.GenericSparseMatrixCSC(A)
    At = G.GenericSparseMatrixCSR(H.n, H.m, Ac.colptr, Ac.rowval, Ac.nzval, copy=False)
    T_t = csc_tuple(Ac)
    A_t = (H.rowptr, H.colval, H.nzval)
    C = A + At
    same(csr_tuple(C), O.csr_add(H.m, A_t, T_t))
    assert C.nnz() > H.nnz


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_coo_add_merge_and_sort_paths(dev, itype):
    """COO + COO (matrix_coo.jl:300-384): operands already in column-major order take
    the merge; others the radix sort.  Both equal the oracle's stable sort of the
    concatenation, with duplicates inside an operand and across the two operands
    (the first operand's entry first in each run), cancellations dropped."""
    rng = np.random.default_rng(91)
    m, n = 700, 500

    def colmajor(nnz, dup):
        r = rng.integers(1, m + 1, nnz)
        c = rng.integers(1, n + 1, nnz)
        r = np.r_[r, r[:dup]]
        c = np.r_[c, c[:dup]]
        o = np.lexsort((r, c))  # column-major, stable
        return r[o].astype(np.int64), c[o].astype(np.int64), rng.integers(-3, 4, o.size).astype(np.float64)
    A = colmajor(20_000, 300)
    B = colmajor(15_000, 200)
    B = (np.r_[B[0], A[0][:500]], np.r_[B[1], A[1][:500]], np.r_[B[2], -A[2][:500]])
    o = np.lexsort((B[0], B[1]))
    B = tuple(x[o] for x in B)  # shares 500 entries of A with negated values: exact zeros
    exp = O.coo_add(m, A, B)
    dA = from_arrays(G.GenericSparseMatrixCOO, m, n, *A, dev, itype)
    dB = from_arrays(G.GenericSparseMatrixCOO, m, n, *B, dev, itype)
    same(coo_tuple(dA + dB), exp)  # merge path
    pA = rng.permutation(A[0].size)
    Ash = tuple(x[pA] for x in A)
    dAs = from_arrays(G.GenericSparseMatrixCOO, m, n, *Ash, dev, itype)
    same(coo_tuple(dAs + dB), O.coo_add(m, Ash, B))  # sort path
    # row-major operands (a COO made from a CSR, duplicates kept in storage order):
    # sorted by column alone, then merged
    oA = np.lexsort((A[1], A[0]))
    Ar = tuple(x[oA] for x in A)
    oB = np.lexsort((B[1], B[0]))
    Br = tuple(x[oB] for x in B)
    dAr = from_arrays(G.GenericSparseMatrixCOO, m, n, *Ar, dev, itype)
    dBr = from_arrays(G.GenericSparseMatrixCOO, m, n, *Br, dev, itype)
    same(coo_tuple(dAr + dBr), O.coo_add(m, Ar, Br))
    same(coo_tuple(dAr + dB), O.coo_add(m, Ar, B))
    same(coo_tuple(dA + dBr), O.coo_add(m, A, Br))
    # both row-major with Float64 values: every run summed in the oracle's order, bit for bit
    Af = (Ar[0], Ar[1], rng.standard_normal(Ar[0].size))
    Bf = (Br[0], Br[1], rng.standard_normal(Br[0].size))
    dAf = from_arrays(G.GenericSparseMatrixCOO, m, n, *Af, dev, itype)
    dBf = from_arrays(G.GenericSparseMatrixCOO, m, n, *Bf, dev, itype)
    same(coo_tuple(dAf + dBf), O.coo_add(m, Af, Bf))
    Z = (np.zeros(0), np.zeros(0), np.zeros(0))
    E = from_arrays(G.GenericSparseMatrixCOO, m, n, np.zeros(0, np.int64), np.zeros(0, np.int64),
                    np.zeros(0), dev, itype)
    same(coo_tuple(E + dB), O.coo_add(m, Z, B))
    same(coo_tuple(dA + E), O.coo_add(m, A, Z))
    same(coo_tuple(E + dBr), O.coo_add(m, Z, Br))
    same(coo_tuple(dAr + E), O.coo_add(m, Ar, Z))
    # a row-major operand whose entries all cancel
    Neg = (Ar[0], Ar[1], -Ar[2])
    dNeg = from_arrays(G.GenericSparseMatrixCOO, m, n, *Neg, dev, itype)
    same(coo_tuple(dAr + dNeg), O.coo_add(m, Ar, Neg))
