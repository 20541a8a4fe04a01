"""Complex and integer value types (ComplexF64, ComplexF32, Int64, Int32) through
every family the reference runs them on (test/runtests.jl:34-68;
test/shared/matrix_{csr,csc,coo}.jl, conversions.jl, vector.jl), and
GenericSparseVector.

Checks, against the oracle (pinned in tests/test_oracle.py):
- integer-valued complex inputs and integer inputs make every product and sum
  exact, so mul! in every format and wrapper (N, transpose, adjoint, with an
  adjoint dense operand) must equal the oracle bit for bit whatever the order;
  integer products wrap modulo 2^64 / 2^32 (Julia's Int arithmetic);
- on non-integer ComplexF64 data: |y - y_ref| <= 1e-12 * sum_j |a_ij||x_j| per
  entry, and rows of at most 32 entries bit-identical (the reference's loop,
  with Julia's complex product (ac - bd, ad + bc), no FMA);
- +, kron, SpGEMM, dropzeros and the conversions are bit-exact for every type
  (the same operation order as the reference's kernels);
- reductions (tr, sum, dot) to 1e-12 (complex Float64) / 1e-5 (ComplexF32)
  normwise, exactly for integers; issymmetric / ishermitian exactly.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import gsa_hip as G
import oracle as O
from gsa_hip import synth
from util import host, assert_normwise

pytestmark = pytest.mark.gpu

TYPES = [np.complex128, np.complex64, np.int64, np.int32]
TORCH = {np.dtype(np.complex128): torch.complex128, np.dtype(np.complex64): torch.complex64,
         np.dtype(np.int64): torch.int64, np.dtype(np.int32): torch.int32,
         np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32}


def int_valued(rng, size, dtype, big=False):
    if np.issubdtype(dtype, np.complexfloating):
        v = rng.integers(-4, 5, size) + 1j * rng.integers(-4, 5, size)
        v[v == 0] = 1 + 2j
        return v.astype(dtype)
    if big:
        info = np.iinfo(dtype)
        v = rng.integers(info.min // 4, info.max // 4, size)
    else:
        v = rng.integers(-4, 5, size)
    v[v == 0] = 3
    return v.astype(dtype)


def host_csr(m, n, lengths_or_density, seed, dtype, big=False, normal=False):
    """Host CSR arrays (one-based) with sorted distinct columns."""
    rng = np.random.default_rng(seed)
    if np.isscalar(lengths_or_density):
        S = sp.random(m, n, density=lengths_or_density, random_state=rng, format="csr")
        S.sort_indices()
        rp, cv = S.indptr.astype(np.int64) + 1, S.indices.astype(np.int64) + 1
    else:
        H = synth.csr_from_lengths(np.asarray(lengths_or_density), n, rng, "normal")
        rp, cv = H.rowptr.astype(np.int64), H.colval.astype(np.int64)
    k = cv.size
    if normal:
        v = (rng.standard_normal(k) + 1j * rng.standard_normal(k)).astype(dtype)
    else:
        v = int_valued(rng, k, dtype, big)
    return rp, cv, v


def dev_mat(cls, m, n, rp, cv, v, dev, itype=torch.int64):
    return cls(m, n, torch.as_tensor(rp).to(dev, itype), torch.as_tensor(cv).to(dev, itype),
               torch.as_tensor(v).to(dev), copy=False)


def dense_of(m, n, rp, cv, v):
    S = sp.csr_matrix((v, cv - 1, rp - 1), shape=(m, n))
    return S.toarray()


@pytest.mark.parametrize("dtype", TYPES)
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_mul_every_format_and_wrapper_exact(dev, dtype, itype):
    """Integer-valued data: every product exact, so CSR / CSC / COO, N / T / C,
    SpMV and SpMM (k = 3), and an adjoint dense operand equal the oracle exactly."""
    big = np.issubdtype(dtype, np.integer)
    m, n = 700, 500
    lens = np.r_[np.random.default_rng(1).integers(0, 12, m - 3), [300, 650, 0]]
    rp, cv, v = host_csr(m, n, lens, 11, dtype, big=big)
    rng = np.random.default_rng(12)
    A = dev_mat(G.GenericSparseMatrixCSR, m, n, rp, cv, v, dev, itype)
    Ac = G.GenericSparseMatrixCSC(A)
    Ao = G.GenericSparseMatrixCOO(A)
    Aoc = G.GenericSparseMatrixCOO(Ac)  # column-major entries
    cx = np.issubdtype(dtype, np.complexfloating)
    for k in (1, 3):
        x = int_valued(rng, (n, k) if k > 1 else n, dtype, big)
        xt = int_valued(rng, (m, k) if k > 1 else m, dtype, big)
        xd = G.dense(x, device=dev) if k > 1 else torch.as_tensor(x, device=dev)
        xtd = G.dense(xt, device=dev) if k > 1 else torch.as_tensor(xt, device=dev)
        wantN = O.spmm_csr_N(m, rp, cv, v, x)
        wantT = O.spmm_csr_T(m, n, rp, cv, v, xt)
        wantC = O.spmm_csr_T(m, n, rp, cv, v, xt, conj_a=True)
        for M in (A, Ac, Ao, Aoc):
            assert np.array_equal(host(M @ xd), wantN), (type(M).__name__, k, "N")
            assert np.array_equal(host(G.transpose(M) @ xtd), wantT), (type(M).__name__, k, "T")
            assert np.array_equal(host(G.adjoint(M) @ xtd), wantC), (type(M).__name__, k, "C")
    # α, β (complex α with a nonzero imaginary part; β scales C)
    x = int_valued(rng, n, dtype, big)
    c0 = int_valued(rng, m, dtype)
    alpha, beta = (dtype(2 - 1j), dtype(1 + 3j)) if cx else (dtype(3), dtype(-2))
    C = torch.as_tensor(c0.copy(), device=dev)
    G.mul_(C, A, torch.as_tensor(x, device=dev), alpha.item(), beta.item())
    want = O.spmm_csr_N(m, rp, cv, v, x, C=c0, alpha=alpha, beta=beta)
    assert np.array_equal(host(C), want)
    if cx:  # mul!(C, A, adjoint(B)): B stored k x n, read conjugated (CONJB)
        Bs = int_valued(rng, (3, n), dtype)
        Bd = G.dense(Bs, device=dev)  # column-major 3 x n
        got = host(A @ Bd.mH)
        assert np.array_equal(got, O.spmm_csr_N(m, rp, cv, v, np.conj(Bs.T)))


def test_complex_f64_rounding_and_short_rows(dev):
    """Non-integer ComplexF64 on a 100k-row power-law matrix: 1e-12 normwise
    against the oracle, rows of at most 32 entries bit-identical; adjoint SpMV
    through the gather view to the same tolerance."""
    H = synth.powerlaw_csr(100_000, 100_000, alpha=1.846, seed=21)
    rng = np.random.default_rng(22)
    v = (rng.standard_normal(H.nnz) + 1j * rng.standard_normal(H.nnz)).astype(np.complex128)
    x = (rng.standard_normal(H.n) + 1j * rng.standard_normal(H.n)).astype(np.complex128)
    A = dev_mat(G.GenericSparseMatrixCSR, H.m, H.n, H.rowptr, H.colval, v, dev)
    y = host(A @ torch.as_tensor(x, device=dev))
    ref = O.spmm_csr_N(H.m, H.rowptr, H.colval, v, x)
    bound = O.spmm_csr_N(H.m, H.rowptr, H.colval, np.abs(v), np.abs(x))
    err = np.abs(y - ref)
    assert np.all(err <= 1e-12 * bound + 1e-300), float((err / np.maximum(bound, 1e-300)).max())
    short = np.diff(H.rowptr) <= 32
    assert np.array_equal(y[short], ref[short])
    yc = host(G.adjoint(A) @ torch.as_tensor(x, device=dev))
    refc = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, v, x, conj_a=True)
    bc = O.spmm_csr_T(H.m, H.n, H.rowptr, H.colval, np.abs(v), np.abs(x))
    assert np.all(np.abs(yc - refc) <= 1e-12 * bc + 1e-300)
    # ComplexF32: 1e-5 normwise against the ComplexF64 oracle on the same values
    v32, x32 = v.astype(np.complex64), x.astype(np.complex64)
    A32 = dev_mat(G.GenericSparseMatrixCSR, H.m, H.n, H.rowptr, H.colval, v32, dev)
    y32 = host(A32 @ torch.as_tensor(x32, device=dev)).astype(np.complex128)
    ref64 = O.spmm_csr_N(H.m, H.rowptr, H.colval, v32.astype(np.complex128),
                         x32.astype(np.complex128))
    assert np.all(np.abs(y32 - ref64) <= 1e-5 * bound + 1e-300)


@pytest.mark.parametrize("dtype", TYPES)
def test_add_merge_paths_every_type(dev, dtype):
    """CSR + and row-major COO + with rows of about 100 entries (merge-path runs),
    one row of 4500 entries (chunks, with equal pairs across chunk edges) and
    cancellations, as built (a mean of about 200 entries per row: the even tiles)
    and after 5000 empty rows (a mean under 12: the row blocks), bit for bit
    against the oracle for every value type."""
    rng = np.random.default_rng(57)
    n = 6000
    lensA = np.r_[rng.integers(60, 140, 40), [2500], rng.integers(0, 9, 10)]
    lensB = np.r_[rng.integers(60, 140, 40), [2000], rng.integers(0, 9, 10)]
    rowsA = [np.sort(rng.choice(np.arange(1, n + 1), k, replace=False)) for k in lensA]
    # B shares about half of A's columns in every row (the long row: every other one)
    rowsB = []
    for ca, k in zip(rowsA, lensB):
        own = rng.choice(np.arange(1, n + 1), k, replace=False)
        rowsB.append(np.unique(np.r_[ca[::2][: k // 2], own[: k - k // 2]]))
    vA = [int_valued(rng, c.size, dtype) for c in rowsA]
    vB = [int_valued(rng, c.size, dtype) for c in rowsB]
    for ca, va, cb, vb in zip(rowsA, vA, rowsB, vB):  # a third of the shared places cancel
        shared = np.flatnonzero(np.isin(cb, ca))[::3]
        vb[shared] = -va[np.searchsorted(ca, cb[shared])]
    for extra in (0, 5000):
        lA = [c.size for c in rowsA] + [0] * extra
        lB = [c.size for c in rowsB] + [0] * extra
        m = len(lA)
        rpA = np.r_[0, np.cumsum(lA)].astype(np.int64) + 1
        rpB = np.r_[0, np.cumsum(lB)].astype(np.int64) + 1
        cvA, cvB = np.concatenate(rowsA).astype(np.int64), np.concatenate(rowsB).astype(np.int64)
        nzA, nzB = np.concatenate(vA), np.concatenate(vB)
        A = dev_mat(G.GenericSparseMatrixCSR, m, n, rpA, cvA, nzA, dev)
        B = dev_mat(G.GenericSparseMatrixCSR, m, n, rpB, cvB, nzB, dev)
        want = O.csr_add(m, (rpA, cvA, nzA), (rpB, cvB, nzB))
        C = A + B
        for got, w in zip((host(C.rowptr), host(C.colval), host(C.nzval)), want):
            assert np.array_equal(got, w)
        rA = np.repeat(np.arange(1, m + 1), np.diff(rpA)).astype(np.int64)
        rB = np.repeat(np.arange(1, m + 1), np.diff(rpB)).astype(np.int64)
        Ao = dev_mat(G.GenericSparseMatrixCOO, m, n, rA, cvA, nzA, dev)
        Bo = dev_mat(G.GenericSparseMatrixCOO, m, n, rB, cvB, nzB, dev)
        wantc = O.coo_add(m, (rA, cvA, nzA), (rB, cvB, nzB))
        Co = Ao + Bo
        for got, w in zip((host(Co.rowind), host(Co.colind), host(Co.nzval)), wantc):
            assert np.array_equal(got, w)


@pytest.mark.parametrize("dtype", TYPES)
def test_structural_ops_bit_exact(dev, dtype):
    """+ (CSR, CSC, COO, with cancellation), dropzeros, kron (COO, CSR), SpGEMM
    (CSR, with a cancellation zero kept) and the conversions, bit for bit."""
    cx = np.issubdtype(dtype, np.complexfloating)
    m, n = 300, 240
    rpA, cvA, vA = host_csr(m, n, 0.03, 31, dtype, normal=cx)
    rpB, cvB, vB = host_csr(m, n, 0.03, 32, dtype, normal=cx)
    # make some sums cancel exactly: B takes -A's value at a few shared places
    SA = sp.csr_matrix((vA, cvA - 1, rpA - 1), shape=(m, n))
    SB = sp.csr_matrix((vB, cvB - 1, rpB - 1), shape=(m, n)).tolil()
    r, c = SA.nonzero()
    for i in range(0, r.size, 7):
        SB[r[i], c[i]] = -SA[r[i], c[i]]
    SB = SB.tocsr()
    SB.sort_indices()
    rpB, cvB, vB = SB.indptr.astype(np.int64) + 1, SB.indices.astype(np.int64) + 1, \
        SB.data.astype(dtype)
    A = dev_mat(G.GenericSparseMatrixCSR, m, n, rpA, cvA, vA, dev)
    B = dev_mat(G.GenericSparseMatrixCSR, m, n, rpB, cvB, vB, dev)
    want = O.csr_add(m, (rpA, cvA, vA), (rpB, cvB, vB))
    C = A + B
    for got, w in zip((host(C.rowptr), host(C.colval), host(C.nzval)), want):
        assert np.array_equal(got, w)
    Cc = G.GenericSparseMatrixCSC(A) + G.GenericSparseMatrixCSC(B)
    np.testing.assert_array_equal(Cc.collect(), C.collect())
    Ao, Bo = G.GenericSparseMatrixCOO(A), G.GenericSparseMatrixCOO(B)
    rA, cA, vAo = (host(t) for t in (Ao.rowind, Ao.colind, Ao.nzval))
    rB, cB, vBo = (host(t) for t in (Bo.rowind, Bo.colind, Bo.nzval))
    wantc = O.coo_add(m, (rA, cA, vAo), (rB, cB, vBo))
    Co = Ao + Bo
    for got, w in zip((host(Co.rowind), host(Co.colind), host(Co.nzval)), wantc):
        assert np.array_equal(got, w)
    assert (A - A).nnz() == 0
    # dropzeros keeps the order, drops exact zeros (both parts for complex)
    vz = vA.copy()
    vz[::5] = 0
    Z = dev_mat(G.GenericSparseMatrixCSR, m, n, rpA, cvA, vz, dev)
    wz = O.csr_dropzeros(m, rpA, cvA, vz)
    D = G.dropzeros(Z)
    for got, w in zip((host(D.rowptr), host(D.colval), host(D.nzval)), wz):
        assert np.array_equal(got, w)
    # kron of two small operands
    s1 = host_csr(20, 15, 0.2, 33, dtype, normal=cx)
    s2 = host_csr(12, 9, 0.25, 34, dtype, normal=cx)
    K1 = dev_mat(G.GenericSparseMatrixCSR, 20, 15, *s1, dev)
    K2 = dev_mat(G.GenericSparseMatrixCSR, 12, 9, *s2, dev)
    wk = O.kron_csr(20, 15, 12, 9, s1, s2)
    K = G.kron(K1, K2)
    for got, w in zip((host(K.rowptr), host(K.colval), host(K.nzval)), wk):
        assert np.array_equal(got, w)
    Kc = G.kron(G.GenericSparseMatrixCOO(K1), G.GenericSparseMatrixCOO(K2))
    wkc = O.kron_coo(12, 9, O.csr_to_coo(20, *s1), O.csr_to_coo(12, *s2))
    for got, w in zip((host(Kc.rowind), host(Kc.colind), host(Kc.nzval)), wkc):
        assert np.array_equal(got, w)
    # SpGEMM (Gustavson order; A * A' has cancellations for integer data)
    rpT, cvT, vT = host_csr(n, 200, 0.04, 35, dtype, normal=cx)
    Bt = dev_mat(G.GenericSparseMatrixCSR, n, 200, rpT, cvT, vT, dev)
    P = A * Bt
    wp = O.spgemm_csr(m, 200, (rpA, cvA, vA), (rpT, cvT, vT))
    for got, w in zip((host(P.rowptr), host(P.colval), host(P.nzval)), wp):
        assert np.array_equal(got, w)
    Pc = G.GenericSparseMatrixCSC(A) * G.GenericSparseMatrixCSC(Bt)
    # column-wise accumulation (the CSC kernel's order): equal to rounding
    np.testing.assert_allclose(Pc.collect(), P.collect(), rtol=1e-5, atol=1e-5)
    # conversions: values moved unchanged, indices bit-exact
    cp, rv, vv = O.csr_to_csc(m, n, rpA, cvA, vA)
    Ac = G.GenericSparseMatrixCSC(A)
    assert np.array_equal(host(Ac.colptr), cp) and np.array_equal(host(Ac.rowval), rv)
    assert np.array_equal(host(Ac.nzval), vv)
    back = G.GenericSparseMatrixCSR(Ac)
    assert np.array_equal(host(back.nzval), vA)
    coo = G.GenericSparseMatrixCOO(A)
    assert np.array_equal(host(coo.nzval), vA)
    assert np.array_equal(host(G.GenericSparseMatrixCSR(coo).nzval), vA)


def test_complex_transposed_conversions_conjugate(dev):
    """GenericSparseMatrixCSC(adjoint(A)) holds the conjugated values
    (conversions.jl:238-258); transpose keeps them."""
    rp, cv, v = host_csr(50, 40, 0.1, 41, np.complex128, normal=True)
    A = dev_mat(G.GenericSparseMatrixCSR, 50, 40, rp, cv, v, dev)
    D = dense_of(50, 40, rp, cv, v)
    np.testing.assert_array_equal(G.GenericSparseMatrixCSC(G.adjoint(A)).collect(), D.conj().T)
    np.testing.assert_array_equal(G.GenericSparseMatrixCSC(G.transpose(A)).collect(), D.T)
    np.testing.assert_array_equal((G.adjoint(A) + G.adjoint(A)).collect(), 2 * D.conj().T)


@pytest.mark.parametrize("dtype", TYPES)
def test_reductions_every_type(dev, dtype):
    """tr, sum, dot(x, A, y) (A, transpose, adjoint; CSR, CSC, COO), A + λI."""
    cx = np.issubdtype(dtype, np.complexfloating)
    n = 400
    rp, cv, v = host_csr(n, n, 0.02, 51, dtype, normal=cx)
    # a stored diagonal in about half the rows
    S = sp.csr_matrix((v, cv - 1, rp - 1), shape=(n, n)).tolil()
    rng = np.random.default_rng(52)
    for i in range(0, n, 2):
        S[i, i] = int_valued(rng, 1, dtype)[0]
    S = S.tocsr()
    S.sort_indices()
    rp, cv, v = S.indptr.astype(np.int64) + 1, S.indices.astype(np.int64) + 1, S.data.astype(dtype)
    D = S.toarray()
    A = dev_mat(G.GenericSparseMatrixCSR, n, n, rp, cv, v, dev)
    forms = (A, G.GenericSparseMatrixCSC(A), G.GenericSparseMatrixCOO(A))
    tol = 1e-12 if dtype == np.complex128 else 1e-5
    for M in forms:
        t = G.tr(M)
        s = G.sum(M)
        if cx:
            sc = np.abs(np.diag(D)).sum() + 1e-300
            assert abs(t - np.trace(D)) <= tol * sc
            assert abs(s - D.sum()) <= tol * (np.abs(D).sum() + 1e-300)
            assert abs(G.tr(G.adjoint(M)) - np.conj(np.trace(D))) <= tol * sc
        else:
            with np.errstate(over="ignore"):
                assert t == int(np.trace(D).astype(dtype))
                assert s == int(D.astype(np.int64).sum())  # Int32 sums widen to Int64
    x = int_valued(rng, n, dtype)
    y = int_valued(rng, n, dtype)
    xd, yd = torch.as_tensor(x, device=dev), torch.as_tensor(y, device=dev)
    for M in forms:
        for W, Dm in ((M, D), (G.transpose(M), D.T), (G.adjoint(M), D.conj().T)):
            got = G.dot(xd, W, yd)
            with np.errstate(over="ignore"):
                want = np.conj(x) @ (Dm @ y)
            if cx:
                scale = np.abs(x) @ (np.abs(Dm) @ np.abs(y)) + 1e-300
                assert abs(got - want) <= tol * scale, (type(M).__name__, got, want)
            else:
                assert got == int(np.asarray(want).astype(dtype)), (type(M).__name__, got, want)
    # A + λI: λ complex for complex matrices; stored zeros are kept (CSR, CSC)
    lam = 2 - 3j if cx else 5
    for M in list(forms):
        R = M + G.UniformScaling(lam)
        rt = host(R.nzval).dtype  # promote_type(Tv, typeof(λ)): a Python complex is ComplexF64
        assert rt == (np.complex128 if cx else dtype)
        np.testing.assert_array_equal(R.collect(), D.astype(rt) + lam * np.eye(n, dtype=rt))
    # kron with a Diagonal of this type
    dv = int_valued(rng, 7, dtype)
    Kd = G.kron(G.Diagonal(torch.as_tensor(dv, device=dev)), A)
    if cx:
        np.testing.assert_allclose(Kd.collect(), np.kron(np.diag(dv), D), rtol=1e-6)
    else:
        np.testing.assert_array_equal(Kd.collect(), np.kron(np.diag(dv), D))


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_issymmetric_ishermitian_complex(dev, dtype):
    """The reference's test (test/shared/matrix_csr.jl:95-120): a non-symmetric
    matrix is neither; A + transpose(A) is symmetric, not Hermitian;
    A + adjoint(A) is Hermitian, not symmetric."""
    rp, cv, v = host_csr(50, 50, 0.1, 61, dtype, normal=True)
    S = sp.csr_matrix((v, cv - 1, rp - 1), shape=(50, 50)).tolil()
    S[0, 1] = 1.0 + 0.0j
    S[1, 0] = 2.0 + 1.0j
    S = S.tocsr()
    for X, sym, herm in ((S, False, False), ((S + S.T).tocsr(), True, False),
                         ((S + S.conj().T).tocsr(), False, True)):
        X.sort_indices()
        M = dev_mat(G.GenericSparseMatrixCSR, 50, 50, X.indptr + 1, X.indices + 1,
                    X.data.astype(dtype), dev)
        for W in (M, G.GenericSparseMatrixCSC(M), G.GenericSparseMatrixCOO(M)):
            assert G.issymmetric(W) == sym
            assert G.ishermitian(W) == herm
        assert G.issymmetric(G.transpose(M)) == sym
        assert G.ishermitian(G.adjoint(M)) == herm


@pytest.mark.parametrize("dtype", TYPES + [np.float64, np.float32])
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_sparse_vector(dev, dtype, itype):
    """GenericSparseVector (test/shared/vector.jl): conversions, sum, dot both
    ways (no atomics), scalar maps, conj/real/imag, norms, normalize."""
    cx = np.issubdtype(dtype, np.complexfloating)
    fl = not np.issubdtype(dtype, np.integer)
    rng = np.random.default_rng(71)
    n = 100_000
    idx = np.sort(rng.choice(n, 1000, replace=False))
    vals = int_valued(rng, idx.size, dtype) if not fl or cx else \
        rng.standard_normal(idx.size).astype(dtype)
    sv = sp.coo_array((vals, (idx,)), shape=(n,))
    V = G.GenericSparseVector.from_scipy(sv, device=dev, index_dtype=itype)
    assert len(V) == n and V.size() == (n,) and V.nnz() == idx.size
    assert np.array_equal(host(V.nonzeroinds()), idx + 1)
    back = V.to_scipy()
    assert np.array_equal(back.toarray(), sv.toarray())
    E = G.GenericSparseVector.from_scipy(sp.coo_array((np.zeros(0, dtype), (np.zeros(0, int),)),
                                                      shape=(10,)), device=dev)
    assert E.nnz() == 0 and len(E) == 10
    dense = sv.toarray()
    y = int_valued(rng, n, dtype) if not fl or cx else rng.standard_normal(n).astype(dtype)
    yd = torch.as_tensor(y, device=dev)
    got = G.dot(V, yd)
    with np.errstate(over="ignore"):
        want = np.vdot(dense, y) if cx else (dense * y).sum()
    if fl and not cx:
        tol = 1e-12 if dtype == np.float64 else 1e-5
        assert abs(got - want) <= tol * (np.abs(dense) @ np.abs(y))
    else:
        assert got == want if cx else got == int(np.asarray(want).astype(dtype))
    rev = G.dot(yd, V)
    assert rev == (np.conj(got) if cx else got)
    s = G.sum(V)
    if fl and not cx:
        assert abs(s - dense.sum()) <= (1e-12 if dtype == np.float64 else 1e-5) * np.abs(dense).sum()
    else:
        assert s == (dense.sum() if cx else int(dense.astype(np.int64).sum()))
    # scalar maps keep the indices
    a = (2 + 1j) if cx else (2 if not fl else 2.5)
    W = V * a
    assert np.array_equal(host(W.nonzeroinds()), idx + 1)
    np.testing.assert_allclose(host(W.nonzeros()), vals * a, rtol=1e-6)
    np.testing.assert_allclose(host((a * V).nonzeros()), vals * a, rtol=1e-6)
    np.testing.assert_allclose(host((-V).nonzeros()), -vals)
    if cx:
        assert np.array_equal(host(V.conj().nonzeros()), np.conj(vals))
        assert np.array_equal(host(V.real().nonzeros()), vals.real)
        assert np.array_equal(host(V.imag().nonzeros()), vals.imag)
    else:
        assert V.conj() is V and V.real() is V and V.imag().nnz() == 0
    if fl:
        for p in (2, 1, np.inf):
            nv = G.norm(V, p)
            np.testing.assert_allclose(nv, np.linalg.norm(vals, p), rtol=1e-6)
        np.testing.assert_allclose(G.norm(G.normalize(V)), 1.0, rtol=1e-6)
        Vc = V.copy()
        G.normalize_(Vc)
        np.testing.assert_allclose(G.norm(Vc), 1.0, rtol=1e-6)
    # errors the reference raises
    with pytest.raises(G.DimensionMismatch):
        G.dot(V, torch.as_tensor(y[:-1], device=dev))
    with pytest.raises(G.ArgumentError):
        G.GenericSparseVector(-1, V.nzind, V.nzval)
    with pytest.raises(G.ArgumentError):
        G.GenericSparseVector(n, V.nzind, V.nzval[:-1])


@pytest.mark.parametrize("dtype", [np.complex128, np.float64, np.int64])
def test_value_maps_every_format(dev, dtype):
    """zero, conj, real, imag, rmul!, lmul!, rdiv! (src/matrix_csr/matrix_csr.jl:80-106,
    src/core.jl:26-40, and their CSC / COO twins): the structure kept (copied),
    the stored values mapped, real inputs returned as they are by conj / real and
    as zero(A) by imag."""
    m, n = 90, 70
    rp, cv, v = host_csr(m, n, 0.05, 5, dtype)
    A = dev_mat(G.GenericSparseMatrixCSR, m, n, rp, cv, v, dev)
    for M in (A, G.GenericSparseMatrixCSC(A), G.GenericSparseMatrixCOO(A)):
        D = M.collect()
        Z = M.zero()
        assert type(Z) is type(M) and Z.shape == M.shape and Z.nnz() == 0
        assert np.array_equal(Z.collect(), np.zeros((m, n), dtype=D.dtype))
        if np.issubdtype(dtype, np.complexfloating):
            assert np.array_equal(M.conj().collect(), np.conj(D))
            assert np.array_equal(M.real().collect(), D.real)
            assert np.array_equal(M.imag().collect(), D.imag)
            assert M.imag().nnz() == M.nnz()  # the structure kept, zeros included
        else:
            assert M.conj() is M and M.real() is M and M.imag().nnz() == 0
        B = M.copy().rmul_(3)
        assert np.array_equal(B.collect(), D * 3)
        assert np.array_equal(M.copy().lmul_(3).collect(), D * 3)
        if not np.issubdtype(dtype, np.integer):
            np.testing.assert_allclose(M.copy().rdiv_(4).collect(), D / 4, rtol=1e-15)
        else:
            with pytest.raises(G.ArgumentError):
                M.copy().rdiv_(4)
        assert np.array_equal(M.collect(), D)  # the source untouched
