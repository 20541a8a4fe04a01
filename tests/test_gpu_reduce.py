"""Parity of tr, sum, dot(x, A, y), issymmetric / ishermitian, A ± λI and kron
with a Diagonal against the oracle.

The sums (tr, sum, dot) are compared within a bound normalised by the sum of
|term| (Float64: 1e-12; Float32 results are rounded once to Float32, so
1e-7 of that sum); each term is the same IEEE product on both sides.  The
symmetry test is a yes/no answer and the structural results (A + λI, kron)
must match index for index and value for value.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import gsa_hip as G
import oracle as O
from gsa_hip import synth
from conftest import GOLDEN
from util import host

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 1e-7}
NPT = {torch.float64: np.float64, torch.float32: np.float32}


def g(name):
    return np.load(os.path.join(GOLDEN, name))


def csr_arrays(S):
    S = sp.csr_matrix(S)
    S.sort_indices()
    return S.indptr.astype(np.int64) + 1, S.indices.astype(np.int64) + 1, S.data.copy()


def csc_arrays(S):
    S = sp.csc_matrix(S)
    S.sort_indices()
    return S.indptr.astype(np.int64) + 1, S.indices.astype(np.int64) + 1, S.data.copy()


def to_dev(cls, m, n, arrs, dev, dtype=torch.float64, itype=torch.int64):
    a, b, c = arrs
    return cls(m, n, torch.as_tensor(a).to(dev, itype), torch.as_tensor(b).to(dev, itype),
               torch.as_tensor(np.asarray(c)).to(dev, dtype), copy=False)


def three_formats(S, dev, dtype, itype, coo_order="col"):
    """CSR, CSC and COO device copies of scipy matrix S (COO column- or row-major)."""
    m, n = S.shape
    A = to_dev(G.GenericSparseMatrixCSR, m, n, csr_arrays(S), dev, dtype, itype)
    Ac = to_dev(G.GenericSparseMatrixCSC, m, n, csc_arrays(S), dev, dtype, itype)
    p, i, v = csc_arrays(S) if coo_order == "col" else csr_arrays(S)
    major = np.repeat(np.arange(1, len(p)), np.diff(p))
    r, c = (i, major) if coo_order == "col" else (major, i)
    Ao = to_dev(G.GenericSparseMatrixCOO, m, n, (r, c, v), dev, dtype, itype)
    return A, Ac, Ao


def powerlaw_square(n, seed, dtype=np.float64):
    H = synth.powerlaw_csr(n, n, seed=seed, hi=3000)
    S = H.to_scipy().astype(dtype)
    # add part of the diagonal so traces have something to sum
    d = np.zeros(n)
    d[::3] = np.random.default_rng(seed).standard_normal(len(d[::3]))
    return sp.csr_matrix(S + sp.diags(d.astype(dtype)))


def close(got, ref, bound, dtype):
    assert abs(got - ref) <= TOL[dtype] * bound + 1e-300, (got, ref, bound)


def host_ptr(A):
    if isinstance(A, G.GenericSparseMatrixCSR):
        return A.m, host(A.rowptr), host(A.colval), host(A.nzval)
    return A.n, host(A.colptr), host(A.rowval), host(A.nzval)


# ------------------------------------------------------------ tr and sum ----
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_trace_and_sum(dev, dtype, itype):
    S = powerlaw_square(20000, 3, NPT[dtype])
    assert np.diff(S.indptr).max() > 64  # long rows take the workgroup path
    for A in three_formats(S, dev, dtype, itype):
        if isinstance(A, G.GenericSparseMatrixCOO):
            ref, bound = O.trace_coo(host(A.rowind), host(A.colind), host(A.nzval))
        else:
            n, p, i, v = host_ptr(A)
            ref, bound = O.trace_ptr(n, p, i, v)
        t0 = G.tr(A)
        close(t0, ref, bound, dtype)
        close(G.tr(G.transpose(A)), ref, bound, dtype)
        vals = host(A.nzval).astype(np.float64)
        close(G.sum(A), vals.sum(), np.abs(vals).sum(), dtype)
        assert G.tr(A) == t0  # fixed order: the same result every call
        if not isinstance(A, G.GenericSparseMatrixCOO):
            A.plan()  # the cached plan's row blocks give the same partition
            assert A.cached_plan() is not None and G.tr(A) == t0


def test_trace_duplicates_and_edges(dev):
    # COO with duplicate diagonal entries: each one counts
    r = torch.tensor([1, 2, 2, 3, 1, 2], device=dev)
    c = torch.tensor([1, 2, 2, 1, 1, 3], device=dev)
    v = torch.tensor([1.0, 2.0, -5.0, 7.0, 4.0, 8.0], dtype=torch.float64, device=dev)
    A = G.GenericSparseMatrixCOO(3, 3, r, c, v)
    assert G.tr(A) == 2.0 and G.sum(A) == 17.0
    E = G.GenericSparseMatrixCSR(4, 4, torch.ones(5, dtype=torch.int64, device=dev),
                                 torch.zeros(0, dtype=torch.int64, device=dev),
                                 torch.zeros(0, dtype=torch.float64, device=dev))
    assert G.tr(E) == 0.0 and G.sum(E) == 0.0
    R = G.GenericSparseMatrixCSR(2, 3, torch.tensor([1, 2, 2], device=dev),
                                 torch.tensor([1], device=dev),
                                 torch.tensor([1.0], dtype=torch.float64, device=dev))
    with pytest.raises(G.DimensionMismatch, match="square to compute the trace"):
        G.tr(R)


# ------------------------------------------------------------ dot(x, A, y) --
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_dot3(dev, dtype, itype):
    H = synth.powerlaw_csr(30000, 20000, seed=4, hi=3000)
    S = H.to_scipy().astype(NPT[dtype])
    m, n = S.shape
    rng = np.random.default_rng(5)
    A, Ac, Ao = three_formats(S, dev, dtype, itype)
    for wrap in (lambda X: X, G.transpose, G.adjoint):
        trans = wrap is not None and wrap(A) is not A
        rows, cols = (n, m) if trans else (m, n)
        x = rng.standard_normal(rows).astype(NPT[dtype])
        y = rng.standard_normal(cols).astype(NPT[dtype])
        xd, yd = torch.as_tensor(x).to(dev), torch.as_tensor(y).to(dev)
        p, i, v = csr_arrays(S)
        ref, bound = O.dot3_ptr(m, p, i, v, x, y, major_left=not trans)
        for M in (A, Ac, Ao):
            close(G.dot(xd, wrap(M), yd), ref, bound, dtype)
    A.plan()
    Ac.plan()
    for M in (A, Ac):  # through the cached plans (Float64: the SpMV windows)
        close(G.dot(xd, G.transpose(M), yd), ref, bound, dtype)
    x2 = rng.standard_normal(m).astype(NPT[dtype])
    y2 = rng.standard_normal(n).astype(NPT[dtype])
    ref2, bound2 = O.dot3_ptr(m, p, i, v, x2, y2, major_left=True)
    for M in (A, Ac):
        close(G.dot(torch.as_tensor(x2).to(dev), M, torch.as_tensor(y2).to(dev)), ref2, bound2,
              dtype)


def test_dot3_promotion_and_checks(dev):
    S = sp.random(300, 200, density=0.05, random_state=np.random.default_rng(6), format="csr")
    A = to_dev(G.GenericSparseMatrixCSR, 300, 200, csr_arrays(S), dev, torch.float32)
    rng = np.random.default_rng(7)
    x, y = rng.standard_normal(300), rng.standard_normal(200)
    # Float32 A with Float64 vectors: promote_type -> Float64 terms
    p, i, v = csr_arrays(S.astype(np.float32).astype(np.float64))
    ref, bound = O.dot3_ptr(300, p, i, v, x, y, True)
    got = G.dot(torch.as_tensor(x).to(dev), A, torch.as_tensor(y).to(dev))
    close(got, ref, bound, torch.float64)
    with pytest.raises(G.DimensionMismatch, match="first dimension of A, 300"):
        G.dot(torch.zeros(299, dtype=torch.float64, device=dev), A,
              torch.zeros(200, dtype=torch.float64, device=dev))
    with pytest.raises(G.DimensionMismatch, match="second dimension of A, 200"):
        G.dot(torch.zeros(300, dtype=torch.float64, device=dev), A,
              torch.zeros(201, dtype=torch.float64, device=dev))
    with pytest.raises(G.ArgumentError, match="same backend"):
        G.dot(torch.zeros(300, dtype=torch.float64), A,
              torch.zeros(200, dtype=torch.float64, device=dev))


# ------------------------------------------------------------ issymmetric ---
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_issymmetric(dev, dtype, itype):
    H = synth.powerlaw_csr(20000, 20000, seed=8, hi=3000)
    B = H.to_scipy().astype(NPT[dtype])
    Ssym = sp.csr_matrix(B + B.T)
    assert np.diff(Ssym.indptr).max() > 64
    # one value changed, one transpose entry removed
    Sval = Ssym.copy()
    r0 = int(np.argmax(np.diff(Sval.indptr)))  # in a long row
    off = np.flatnonzero(Sval.indices[Sval.indptr[r0]:Sval.indptr[r0 + 1]] != r0)
    j = Sval.indptr[r0] + off[5]
    Sval.data[j] = Sval.data[j] + 1
    Smiss = Ssym.tolil()
    rr, cc = Ssym.nonzero()
    k = np.flatnonzero(rr != cc)[17]
    Smiss[rr[k], cc[k]] = 0
    Smiss = sp.csr_matrix(Smiss)
    Smiss.eliminate_zeros()
    for S, expect in ((Ssym, True), (Sval, False), (Smiss, False), (B, False)):
        n, p, i, v = 20000, *csr_arrays(S)
        assert O.symmetric_ptr(n, p, i, v) == expect
        for M in three_formats(S, dev, dtype, itype):
            assert G.issymmetric(M) is expect
            assert G.ishermitian(G.adjoint(M)) is expect
            if not isinstance(M, G.GenericSparseMatrixCOO):
                M.plan()
                assert G.issymmetric(M) is expect
    R = to_dev(G.GenericSparseMatrixCSR, 3, 2, csr_arrays(sp.csr_matrix(np.ones((3, 2)))), dev)
    assert G.issymmetric(R) is False
    E = to_dev(G.GenericSparseMatrixCOO, 4, 4, (np.zeros(0, np.int64),) * 2 + (np.zeros(0),), dev)
    assert G.issymmetric(E) is True


def test_issymmetric_reference_pair(dev):
    # test/shared/matrix_coo.jl:96-120 on real values: A[1,2] = 1, A[2,1] = 2
    rng = np.random.default_rng(9)
    N = sp.random(50, 50, density=0.1, random_state=rng, format="lil")
    N[0, 1], N[1, 0] = 1.0, 2.0
    N = sp.csr_matrix(N)
    for M in three_formats(N, dev, torch.float64, torch.int64):
        assert not G.issymmetric(M) and not G.issymmetric(G.transpose(M))
    Ssym = sp.csr_matrix(N + N.T)
    for M in three_formats(Ssym, dev, torch.float64, torch.int64):
        assert G.issymmetric(M) and G.ishermitian(M)


# ------------------------------------------------------------ A ± λI --------
def _ptr_tuple(A):
    if isinstance(A, G.GenericSparseMatrixCSR):
        return host(A.rowptr), host(A.colval), host(A.nzval)
    return host(A.colptr), host(A.rowval), host(A.nzval)


def _same(got, exp):
    for a, b in zip(got, exp):
        assert a.shape == b.shape and np.array_equal(a, b), (a[:10], b[:10])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_add_uniform_scaling(dev, dtype, itype):
    S = powerlaw_square(20000, 10, NPT[dtype])
    S = S.tolil()
    S[7, 7] = -2.5  # cancels to a stored zero in CSR/CSC, dropped in COO
    S = sp.csr_matrix(S)
    lam = np.float32(2.5) if dtype == torch.float32 else 2.5
    J = G.UniformScaling(lam)
    A, Ac, Ao = three_formats(S, dev, dtype, itype)
    for M in (A, Ac):
        n, p, i, v = host_ptr(M)
        exp = O.add_diag(n, p, i, v, lam)
        R = M + J
        assert type(R) is type(M) and R.nzval.dtype == dtype and R._idx().dtype == itype
        _same(_ptr_tuple(R), exp)
        M.plan()  # through the cached plan's row blocks
        _same(_ptr_tuple(J + M), exp)
    r, c, v = host(Ao.rowind), host(Ao.colind), host(Ao.nzval)
    exp = O.coo_add_diag(20000, (r, c, v), lam)
    R = Ao + J
    _same((host(R.rowind), host(R.colind), host(R.nzval)), exp)
    assert 0 not in host(R.nzval)
    # A - λI and λI - A (src/core.jl:89-94), checked densely on a small case
    S2 = sp.random(40, 40, density=0.2, random_state=np.random.default_rng(11), format="csr")
    S2 = S2.astype(NPT[dtype])
    for M in three_formats(S2, dev, dtype, itype):
        assert np.allclose((M - J).collect(), S2.toarray() - float(lam) * np.eye(40))
        assert np.allclose((J - M).collect(), float(lam) * np.eye(40) - S2.toarray())
        assert np.array_equal((M * G.I).collect(), S2.toarray())
        assert np.array_equal((G.I * M).collect(), S2.toarray())


def test_add_uniform_scaling_golden_and_rules(dev):
    d = g("diag_ops.npz")
    N = to_dev(G.GenericSparseMatrixCSR, 3, 3, (d["N_0"], d["N_1"], d["N_2"]), dev)
    for M in three_formats(sp.csr_matrix(N.collect()), dev, torch.float64, torch.int64):
        assert np.array_equal((M + G.I).collect(), d["E_nodiag"])
        assert np.array_equal((G.I + M).collect(), d["E_nodiag"])
    A4 = to_dev(G.GenericSparseMatrixCSR, 20, 20, (d["A4_rowptr"], d["A4_colval"], d["A4_nzval"]),
                dev)
    _same(_ptr_tuple(A4 + 3.0 * G.I), (d["A4p3_rowptr"], d["A4p3_colval"], d["A4p3_nzval"]))
    # λ = 0 returns a copy; a non-square matrix raises; Float32 + a Python float promotes
    Z = A4 + 0.0 * G.I
    assert Z is not A4 and np.array_equal(Z.collect(), A4.collect())
    R = to_dev(G.GenericSparseMatrixCSR, 3, 2, csr_arrays(sp.csr_matrix(np.ones((3, 2)))), dev)
    with pytest.raises(G.DimensionMismatch, match="square to add UniformScaling"):
        R + G.I
    A32 = to_dev(G.GenericSparseMatrixCSR, 20, 20,
                 (d["A4_rowptr"], d["A4_colval"], d["A4_nzval"].astype(np.float32)), dev,
                 torch.float32)
    assert (A32 + G.I).dtype == torch.float32
    assert (A32 + 0.5 * G.I).dtype == torch.float64
    assert (A32 + np.float32(0.5) * G.I).dtype == torch.float32


# ------------------------------------------------------------ kron(Diagonal) -
@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_kron_diagonal_golden(dev, itype):
    d = g("diag_ops.npz")
    D = G.Diagonal(torch.as_tensor(d["d5"]).to(dev))
    B = to_dev(G.GenericSparseMatrixCOO, 20, 15, (d["B5coo_0"], d["B5coo_1"], d["B5coo_2"]), dev,
               itype=itype)
    A = to_dev(G.GenericSparseMatrixCOO, 30, 25, (d["A5coo_0"], d["A5coo_1"], d["A5coo_2"]), dev,
               itype=itype)
    C = G.kron(D, B)
    assert isinstance(C, G.GenericSparseMatrixCOO) and C.shape == (80, 60)
    _same((host(C.rowind), host(C.colind), host(C.nzval)), (d["KD_0"], d["KD_1"], d["KD_2"]))
    C = G.kron(A, D)
    assert C.shape == (120, 100)
    _same((host(C.rowind), host(C.colind), host(C.nzval)), (d["KA_0"], d["KA_1"], d["KA_2"]))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_kron_diagonal_formats_and_wrappers(dev, dtype):
    rng = np.random.default_rng(12)
    Sb = sp.random(300, 200, density=0.05, random_state=rng, format="csr").astype(NPT[dtype])
    dvals = rng.standard_normal(7).astype(NPT[dtype])
    dvals[3] = 0  # zero diagonal values keep their (zero) entries
    D = G.Diagonal(torch.as_tensor(dvals).to(dev))
    B, Bc, Bo = three_formats(Sb, dev, dtype, torch.int64, coo_order="row")
    for wrap in (lambda X: X, G.transpose):
        for M in (B, Bc, Bo):
            Mo = G.convert.to_coo(wrap(M))
            co = (host(Mo.rowind), host(Mo.colind), host(Mo.nzval))
            mB, nB = Mo.shape
            left = O.kron_diag_coo(mB, nB, dvals, co)
            right = O.kron_coo_diag(co, dvals)
            for got, exp, shape in ((G.kron(D, wrap(M)), left, (7 * mB, 7 * nB)),
                                    (G.kron(wrap(M), D), right, (mB * 7, nB * 7))):
                assert got.shape == shape
                if isinstance(M, G.GenericSparseMatrixCOO):
                    _same((host(got.rowind), host(got.colind), host(got.nzval)), exp)
                elif isinstance(M, G.GenericSparseMatrixCSR):
                    assert isinstance(got, G.GenericSparseMatrixCSR)
                    _same(_ptr_tuple(got), O.coo_to_csr(*shape, *exp))
                else:
                    assert isinstance(got, G.GenericSparseMatrixCSC)
                    _same(_ptr_tuple(got), O.coo_to_csc(*shape, *exp))
    # FillArrays' Diagonal(Fill(2, 4)) of test/shared/matrix_coo.jl:443: a constant diagonal
    D2 = G.Diagonal.fill(2.0, 4, dtype=dtype, device=dev)
    got = G.kron(D2, Bo)
    exp = O.kron_diag_coo(300, 200, np.full(4, 2.0, NPT[dtype]),
                          (host(Bo.rowind), host(Bo.colind), host(Bo.nzval)))
    _same((host(got.rowind), host(got.colind), host(got.nzval)), exp)
    got = G.kron(G.transpose(B), D2)  # CSR operand: CSR result
    Mo = G.convert.to_coo(G.transpose(B))
    exp = O.kron_coo_diag((host(Mo.rowind), host(Mo.colind), host(Mo.nzval)),
                          np.full(4, 2.0, NPT[dtype]))
    _same(_ptr_tuple(got), O.coo_to_csr(200 * 4, 300 * 4, *exp))


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.int64])
def test_sparse_plus_dense(dev, dtype, itype):
    """A + B and B + A for a dense B (src/core.jl:67-89): copy(B) with every stored
    entry added.  Without duplicates each element gets one add, so the result
    equals numpy's B + dense(A) bit for bit, for CSR, CSC, COO, their transpose
    and adjoint, with B column- or row-major; A - B and B - A go through -B / -A."""
    rng = np.random.default_rng(23)
    m, n = 300, 220
    S = sp.random(m, n, density=0.03, random_state=rng, format="csr")
    if np.issubdtype(dtype, np.complexfloating):
        S.data = (S.data + 1j * rng.standard_normal(S.nnz)).astype(dtype)
        Bh = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(dtype)
    elif np.issubdtype(dtype, np.integer):
        S.data = rng.integers(-50, 50, S.nnz).astype(dtype)
        Bh = rng.integers(-1000, 1000, (m, n)).astype(dtype)
    else:
        S.data = rng.standard_normal(S.nnz).astype(dtype)
        Bh = rng.standard_normal((m, n)).astype(dtype)
    A = G.convert.csr_from_scipy(S, device=dev, index_dtype=itype)
    Dn = S.toarray()
    forms = {"csr": A, "csc": G.GenericSparseMatrixCSC(A), "coo": G.GenericSparseMatrixCOO(A)}
    for layout in ("col", "row"):
        Bd = torch.as_tensor(Bh, device=dev)
        if layout == "col":
            Bd = Bd.t().contiguous().t()
        for name, M in forms.items():
            got = M + Bd
            assert got.dtype == Bd.dtype and got.stride() == Bd.stride(), (name, layout)
            assert np.array_equal(host(got), Bh + Dn), (name, layout)
            assert np.array_equal(host(Bd + M), Bh + Dn), (name, layout)
            assert np.array_equal(host(Bd), Bh)  # B untouched
            Bt = torch.as_tensor(np.ascontiguousarray(Bh.T), device=dev)
            assert np.array_equal(host(G.transpose(M) + Bt), Bh.T + Dn.T), name
            assert np.array_equal(host(G.adjoint(M) + Bt), Bh.T + np.conj(Dn.T)), name
            if not np.issubdtype(dtype, np.integer):
                assert np.array_equal(host(M - Bd), Dn + (-Bh)), name
                assert np.array_equal(host(Bd - M), Bh + (-Dn)), name
    # COO duplicates add up (in any order: integer values make it exact)
    if np.issubdtype(dtype, np.integer):
        r = np.array([1, 3, 1, 1]); c = np.array([2, 2, 2, 5]); v = np.array([4, 5, 6, 7])
        Dup = G.GenericSparseMatrixCOO(m, n, torch.as_tensor(r).to(dev, itype),
                                       torch.as_tensor(c).to(dev, itype),
                                       torch.as_tensor(v.astype(dtype)).to(dev))
        exp = Bh.copy()
        np.add.at(exp, (r - 1, c - 1), v.astype(dtype))
        assert np.array_equal(host(Dup + torch.as_tensor(Bh, device=dev)), exp)
    # the reference's checks
    with pytest.raises(G.DimensionMismatch):
        A + torch.zeros((m, n + 1), dtype=torch.as_tensor(Bh).dtype, device=dev)
    if dtype == np.float64:
        with pytest.raises(G.ArgumentError):  # Float64 values into a Float32 dense copy
            A + torch.zeros((m, n), dtype=torch.float32, device=dev)


# ------------------------------------------------ entry blocks across rows --
def _block_crossing_matrix(dtype):
    """A 12000 x 12000 CSR whose rows cross the 2048-entry blocks of the
    entry-parallel traversals in every way: a 5000-entry row with its diagonal
    missing (lambda goes in front), rows whose diagonal or first index past it
    lies blocks away from the row's start, a block owning no row (inside a long
    row), runs of empty rows (also at the end), a long row whose indices all
    lie left of the diagonal (lambda appended), and short rows between."""
    rng = np.random.default_rng(41)
    n = 12000
    rows = {}
    rows[0] = np.arange(1, 5001)                               # diagonal missing, all past
    rows[1] = np.concatenate([[0], np.arange(2, 3001)])       # first past at the 2nd entry
    rows[2] = np.arange(0, 4500)                               # diagonal at offset 2
    rows[3] = np.arange(0, 64)                                 # short, diagonal present
    for r in range(4, 40):                                     # short rows, some with diagonals
        k = int(rng.integers(1, 70))
        rows[r] = np.unique(np.concatenate([rng.choice(n, k, replace=False), [r] if r % 3 else []]))
    for r in range(40, 2040):                                  # a run of empty rows
        rows[r] = np.zeros(0, np.int64)
    rows[2040] = np.concatenate([np.arange(0, 2040), np.arange(2041, 6200)])  # lambda 2040 in
    rows[2041] = np.arange(0, 2041 + 1)                        # diagonal is the last entry
    rows[2042] = np.arange(0, 1700)                            # all left of the diagonal: appended
    rows[2043] = np.arange(7000, 11100)                        # all past, crosses two blocks
    for r in range(2044, 2300):
        k = int(rng.integers(1, 12))
        rows[r] = np.unique(np.concatenate([rng.choice(n, k, replace=False), [r] if r % 2 else []]))
    for r in range(2300, n):                                   # trailing empty rows
        rows[r] = np.zeros(0, np.int64)
    lens = np.array([len(rows[r]) for r in range(n)])
    ptr = np.concatenate([[0], np.cumsum(lens)])
    col = np.concatenate([np.sort(rows[r]).astype(np.int64) for r in range(n)])
    val = rng.standard_normal(len(col)).astype(dtype)
    S = sp.csr_matrix((val, col, ptr), shape=(n, n))
    assert S.has_sorted_indices and S.nnz > 5 * 2048
    return S


@pytest.mark.parametrize("itype", [torch.int64, torch.int32])
def test_entry_blocks_cross_rows(dev, itype):
    """tr, dot, A + lambda*I (count and fill), sparse + dense and issymmetric where
    rows span several entry blocks, against the oracle (and numpy for + dense)."""
    dtype = torch.float64
    S = _block_crossing_matrix(np.float64)
    n = S.shape[0]
    p, i, v = csr_arrays(S)
    A = to_dev(G.GenericSparseMatrixCSR, n, n, (p, i, v), dev, dtype, itype)
    Ac = to_dev(G.GenericSparseMatrixCSC, n, n, csc_arrays(S), dev, dtype, itype)
    for cached in (False, True):
        if cached:
            A.plan()
            Ac.plan()
        t, bound = O.trace_ptr(n, p, i, v)
        close(G.tr(A), t, bound, dtype)
        close(G.tr(Ac), t, bound, dtype)
        exp = O.add_diag(n, p, i, v, 2.5)
        _same(_ptr_tuple(A + G.UniformScaling(2.5)), exp)
        pc, ic, vc = csc_arrays(S)
        _same(_ptr_tuple(Ac + G.UniformScaling(2.5)), O.add_diag(n, pc, ic, vc, 2.5))
        rng = np.random.default_rng(42)
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        ref, b2 = O.dot3_ptr(n, p, i, v, x, y, True)
        xd, yd = torch.as_tensor(x).to(dev), torch.as_tensor(y).to(dev)
        close(G.dot(xd, A, yd), ref, b2, dtype)
        close(G.dot(xd, Ac, yd), ref, b2, dtype)
    # sparse + dense on the rows' leading block of columns (a dense 12000 x 12000
    # would be 1.1 GB): the same rows, columns cut to the first 3000
    Sc = sp.csr_matrix(S[:, :3000])
    Sc.sort_indices()
    Ad = to_dev(G.GenericSparseMatrixCSR, n, 3000, csr_arrays(Sc), dev, dtype, itype)
    Bh = np.random.default_rng(43).standard_normal((n, 3000))
    got = Ad + torch.as_tensor(Bh, device=dev)
    assert np.array_equal(host(got), Bh + Sc.toarray())
    # symmetry: S + S' is symmetric; one changed value deep in a long row is not
    Ssym = sp.csr_matrix(S + S.T)
    Ssym.sort_indices()
    Sbad = Ssym.copy()
    j = Sbad.indptr[2040] + 4000
    Sbad.data[j] += 1.0
    for M, expect in ((Ssym, True), (Sbad, False)):
        pp, ii, vv = csr_arrays(M)
        assert O.symmetric_ptr(n, pp, ii, vv) == expect
        D = to_dev(G.GenericSparseMatrixCSR, n, n, (pp, ii, vv), dev, dtype, itype)
        assert G.issymmetric(D) is expect
        D.plan()
        assert G.issymmetric(D) is expect
