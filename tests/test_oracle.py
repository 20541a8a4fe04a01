"""Pins the CPU oracle (oracle/gsa_oracle.c) before it is trusted as the checker.

Each oracle function is checked against something that does not restate it:
the reference's literal known answers (golden fixtures), scipy.sparse on
seeded inputs, exact dense arithmetic on integer-valued matrices, Python's
stable sort on tiny inputs, and closed forms.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from conftest import GOLDEN


def g(name):
    return np.load(os.path.join(GOLDEN, name))


def to_sp(m, n, rp, cv, nz):
    return sp.csr_matrix((nz, cv - 1, rp - 1), shape=(m, n))


def int_csr(m, n, density, seed, lo=-4, hi=5):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, random_state=rng, format="csr")
    A.data = rng.integers(lo, hi, size=A.data.size).astype(np.float64)
    A.data[A.data == 0] = 2.0
    A.sort_indices()
    return A


def arrs(A):
    A = sp.csr_matrix(A)
    return A.indptr.astype(np.int64) + 1, A.indices.astype(np.int64) + 1, A.data.copy()


# ------------------------------------------------------------- products ----
def test_doc_example_spmv():
    d = g("doc_spmv.npz")
    y = O.spmm_csr_N(3, d["rowptr"], d["colval"], d["nzval"], d["x"])
    assert np.array_equal(y, d["y"])  # [10, 3, 15], docs/src/index.md:47-61


@pytest.mark.parametrize("k", [1, 3])
def test_spmm_N_integer_exact(k):
    A = int_csr(60, 45, 0.2, 1)
    rng = np.random.default_rng(2)
    B = rng.integers(-3, 4, size=(45, k)).astype(np.float64)
    C0 = rng.integers(-3, 4, size=(60, k)).astype(np.float64)
    exp = A.toarray() @ B
    got = O.spmm_csr_N(60, *arrs(A), B)
    assert np.array_equal(got, exp)
    # β scaling and α: C = β C0 + α A B, exact on small integers
    got2 = O.spmm_csr_N(60, *arrs(A), B, C=C0, alpha=2.0, beta=-3.0)
    assert np.array_equal(got2, -3.0 * C0 + 2.0 * exp)
    got3 = O.spmm_csr_N(60, *arrs(A), B, C=C0, alpha=1.0, beta=1.0)
    assert np.array_equal(got3, C0 + exp)


def test_spmm_beta_zero_clears_nan():
    A = int_csr(20, 10, 0.3, 3)
    C0 = np.full((20, 1), np.nan)
    got = O.spmm_csr_N(20, *arrs(A), np.ones(10), C=C0, beta=0.0)
    assert np.all(np.isfinite(got))


def test_spmm_T_integer_exact():
    A = int_csr(50, 70, 0.15, 4)
    rng = np.random.default_rng(5)
    x = rng.integers(-3, 4, size=50).astype(np.float64)
    got = O.spmm_csr_T(50, 70, *arrs(A), x)
    assert np.array_equal(got, A.toarray().T @ x)


def test_spmm_coo_duplicates_exact():
    rng = np.random.default_rng(6)
    r = rng.integers(1, 21, size=200)
    c = rng.integers(1, 16, size=200)
    v = rng.integers(-3, 4, size=200).astype(np.float64)
    D = np.zeros((20, 15))
    np.add.at(D, (r - 1, c - 1), v)
    x = rng.integers(-2, 3, size=15).astype(np.float64)
    assert np.array_equal(O.spmm_coo(0, 20, r, c, v, x), D @ x)
    xt = rng.integers(-2, 3, size=20).astype(np.float64)
    assert np.array_equal(O.spmm_coo(1, 15, r, c, v, xt), D.T @ xt)


def test_spmm_float_close_to_dense():
    A = sp.random(200, 150, density=0.05, random_state=np.random.default_rng(7), format="csr")
    x = np.random.default_rng(8).standard_normal(150)
    got = O.spmm_csr_N(200, *arrs(A), x)
    np.testing.assert_allclose(got, A @ x, rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------- conversions --
def test_conversions_random_vs_scipy():
    d = g("conv_random.npz")
    m, n = int(d["m"]), int(d["n"])
    csr = (d["csr_0"], d["csr_1"], d["csr_2"])
    csc = (d["csc_0"], d["csc_1"], d["csc_2"])
    rowmaj = (d["coo_rowmajor_0"], d["coo_rowmajor_1"], d["coo_rowmajor_2"])
    colmaj = (d["coo_colmajor_0"], d["coo_colmajor_1"], d["coo_colmajor_2"])
    # CSR -> COO is row-major
    for a, b in zip(O.csr_to_coo(m, *csr), rowmaj):
        assert np.array_equal(a, b)
    # CSC -> COO is column-major (pass (n, colptr, rowval) and swap)
    c, r, v = O.csr_to_coo(n, *csc)
    for a, b in zip((r, c, v), colmaj):
        assert np.array_equal(a, b)
    # COO (either order) -> CSR / CSC canonical
    for coo in (rowmaj, colmaj):
        for a, b in zip(O.coo_to_csr(m, n, *coo), csr):
            assert np.array_equal(a, b)
        for a, b in zip(O.coo_to_csc(m, n, *coo), csc):
            assert np.array_equal(a, b)
    for a, b in zip(O.csr_to_csc(m, n, *csr), csc):
        assert np.array_equal(a, b)
    for a, b in zip(O.csc_to_csr(m, n, *csc), csr):
        assert np.array_equal(a, b)


def test_conversion_edges():
    d = g("conv_edges.npz")
    m, n = 3, 4
    for name in ("single", "fullrow", "fullcol"):
        csr = tuple(d[f"{name}_csr_{i}"] for i in range(3))
        csc = tuple(d[f"{name}_csc_{i}"] for i in range(3))
        for a, b in zip(O.csr_to_csc(m, n, *csr), csc):
            assert np.array_equal(a, b)
        for a, b in zip(O.csc_to_csr(m, n, *csc), csr):
            assert np.array_equal(a, b)
        assert np.array_equal(to_sp(m, n, *csr).toarray(), d[f"{name}_dense"])


def test_coo_to_csr_empty_gives_all_ones():
    rp, cv, v = O.coo_to_csr(4, 3, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0))
    assert np.array_equal(rp, np.ones(5, np.int64)) and cv.size == 0


def test_coo_to_csr_duplicates_keep_input_order():
    # brute force on a tiny input: Python's sorted() is stable
    rng = np.random.default_rng(9)
    r = rng.integers(1, 4, size=40)
    c = rng.integers(1, 3, size=40)
    v = np.arange(40, dtype=np.float64)
    order = sorted(range(40), key=lambda i: (r[i], c[i]))
    rp, cv, vv = O.coo_to_csr(3, 2, r, c, v)
    assert np.array_equal(vv, v[order]) and np.array_equal(cv, c[order])
    expected_rp = np.r_[np.searchsorted(r[order], np.arange(1, 4), side="left") + 1, 41]
    assert np.array_equal(rp, expected_rp)


def test_sortperm_and_search_primitives():
    rng = np.random.default_rng(10)
    keys = rng.integers(0, 7, size=500)
    assert np.array_equal(O.sortperm(keys), np.argsort(keys, kind="stable") + 1)
    v = np.sort(rng.integers(0, 50, size=100))
    for x in (-1, 0, 13, 49, 50, 60):
        assert O.searchsortedfirst(v, x) == int(np.searchsorted(v, x, side="left")) + 1


# ------------------------------------------------------------- + and dropzeros
def test_add_random_vs_scipy():
    d = g("add_random.npz")
    m = int(d["m"])
    A = (d["A_rowptr"], d["A_colval"], d["A_nzval"])
    B = (d["B_rowptr"], d["B_colval"], d["B_nzval"])
    C = O.csr_add(m, A, B)
    for a, b in zip(C, (d["C_rowptr"], d["C_colval"], d["C_nzval"])):
        assert np.array_equal(a, b)
    # COO +: column-major sorted result
    Acoo = tuple(d[f"Acoo_{i}"] for i in range(3))
    Bcoo = tuple(d[f"Bcoo_{i}"] for i in range(3))
    for a, b in zip(O.coo_add(m, Acoo, Bcoo), tuple(d[f"Ccoo_{i}"] for i in range(3))):
        assert np.array_equal(a, b)


def test_add_overlap_fixture_and_self_cancel():
    d = g("overlap_add.npz")
    m = int(d["m"])
    A = (d["A_rowptr"], d["A_colval"], d["A_nzval"])
    B = (d["B_rowptr"], d["B_colval"], d["B_nzval"])
    for a, b in zip(O.csr_add(m, A, B), (d["C_rowptr"], d["C_colval"], d["C_nzval"])):
        assert np.array_equal(a, b)
    negA = (A[0], A[1], -A[2])
    rp, cv, v = O.csr_add(m, A, negA)  # A - A has nnz 0 (test/shared/matrix_csr.jl:379-381)
    assert v.size == 0 and np.array_equal(rp, np.ones(m + 1, np.int64))


def test_dropzeros_vs_scipy():
    A = int_csr(40, 30, 0.2, 15, lo=-1, hi=2)
    A.data[::3] = 0.0  # stored zeros
    rp, cv, v = O.csr_dropzeros(40, *arrs(A))
    B = A.copy()
    B.eliminate_zeros()
    for a, b in zip((rp, cv, v), arrs(B)):
        assert np.array_equal(a, b)


# ------------------------------------------------------------- kron ---------
def test_kron_vs_scipy():
    d = g("kron_random.npz")
    mA, nA, mB, nB = (int(d[k]) for k in ("mA", "nA", "mB", "nB"))
    A = (d["A_rowptr"], d["A_colval"], d["A_nzval"])
    B = (d["B_rowptr"], d["B_colval"], d["B_nzval"])
    # COO kron is A-major / B-minor (scipy.sparse.kron(coo, coo) has the same order)
    got = O.kron_coo(mB, nB, O.csr_to_coo(mA, *A), O.csr_to_coo(mB, *B))
    for a, b in zip(got, (d["Kcoo_0"], d["Kcoo_1"], d["Kcoo_2"])):
        assert np.array_equal(a, b)
    for a, b in zip(O.kron_csr(mA, nA, mB, nB, A, B), (d["C_rowptr"], d["C_colval"], d["C_nzval"])):
        assert np.array_equal(a, b)


# ------------------------------------------------------------- SpGEMM -------
def test_spgemm_vs_scipy():
    d = g("spgemm_random.npz")
    m, n = int(d["m"]), int(d["n"])
    A = (d["A_rowptr"], d["A_colval"], d["A_nzval"])
    B = (d["B_rowptr"], d["B_colval"], d["B_nzval"])
    got = O.spgemm_csr(m, n, A, B)
    for a, b in zip(got, (d["C_rowptr"], d["C_colval"], d["C_nzval"])):
        assert np.array_equal(a, b)


def test_spgemm_keeps_cancellation_zero():
    # A = [1 1], B = [1; -1]: C = [0] stored explicitly (matrix_csr_kernels.jl:329-341)
    A = (np.array([1, 3]), np.array([1, 2]), np.array([1.0, 1.0]))
    B = (np.array([1, 2, 3]), np.array([1, 1]), np.array([1.0, -1.0]))
    rp, cv, v = O.spgemm_csr(1, 1, A, B)
    assert np.array_equal(rp, [1, 2]) and np.array_equal(cv, [1]) and v[0] == 0.0


def test_spgemm_integer_exact_vs_dense():
    A = int_csr(30, 25, 0.2, 16)
    B = int_csr(25, 20, 0.2, 17)
    rp, cv, v = O.spgemm_csr(30, 20, arrs(A), arrs(B))
    D = to_sp(30, 20, rp, cv, v).toarray()
    assert np.array_equal(D, A.toarray() @ B.toarray())
    # pattern = structural product (every (i,j) with some A_ik B_kj stored)
    S = (sp.csr_matrix((np.ones_like(A.data), A.indices, A.indptr), shape=A.shape)
         @ sp.csr_matrix((np.ones_like(B.data), B.indices, B.indptr), shape=B.shape)).tocsr()
    S.sort_indices()
    assert np.array_equal(rp - 1, S.indptr) and np.array_equal(cv - 1, S.indices)


# ------------------------------------------------------------- reductions --
def _ptr_arrays(A, csc=False):
    B = sp.csc_matrix(A) if csc else sp.csr_matrix(A)
    B.sort_indices()
    return B.indptr.astype(np.int64) + 1, B.indices.astype(np.int64) + 1, B.data.copy()


@pytest.mark.parametrize("csc", [False, True])
def test_trace_and_sum_integer_exact(csc):
    A = int_csr(90, 90, 0.08, 21)
    A = A + sp.diags(np.arange(1, 91, dtype=np.float64) % 5 - 2)  # some diagonal entries, some zeros
    A = sp.csr_matrix(A)
    p, i, v = _ptr_arrays(A, csc)
    s, a = O.trace_ptr(90, p, i, v)
    assert s == A.diagonal().sum()
    assert a == np.abs(A.diagonal()).sum()
    # COO with duplicates on and off the diagonal: every stored diagonal entry counts
    r = np.array([1, 2, 2, 3, 1, 2], np.int64)
    c = np.array([1, 2, 2, 1, 1, 3], np.int64)
    vals = np.array([1.0, 2.0, -5.0, 7.0, 4.0, 8.0])
    s2, a2 = O.trace_coo(r, c, vals)
    D = sp.coo_matrix((vals, (r - 1, c - 1)), shape=(3, 3)).tocsr()  # sums duplicates
    assert s2 == D.diagonal().sum() == 2.0 and a2 == 12.0
    # and in ptr form: rows [1: (1, 1.0), (1, 4.0)], [2: (2, 2.0), (2, -5.0), (3, 8.0)], [3: (1, 7.0)]
    s3, _ = O.trace_ptr(3, np.array([1, 3, 6, 7]), np.array([1, 1, 2, 2, 3, 1]),
                        np.array([1.0, 4.0, 2.0, -5.0, 8.0, 7.0]))
    assert s3 == 2.0


@pytest.mark.parametrize("fmt", ["csr", "csc", "coo"])
@pytest.mark.parametrize("trans", [False, True])
def test_dot3_integer_exact_vs_dense(fmt, trans):
    m, n = (100, 80)
    A = int_csr(m, n, 0.1, 22)
    rng = np.random.default_rng(23)
    Ad = A.toarray()
    opA = Ad.T if trans else Ad
    x = rng.integers(-3, 4, size=opA.shape[0]).astype(np.float64)
    y = rng.integers(-3, 4, size=opA.shape[1]).astype(np.float64)
    exp = float(x @ opA @ y)
    if fmt == "coo":
        C = A.tocoo()
        got, _ = O.dot3_coo(C.row + 1, C.col + 1, C.data, x, y, trans)
    elif fmt == "csr":
        p, i, v = _ptr_arrays(A)
        got, _ = O.dot3_ptr(m, p, i, v, x, y, major_left=not trans)
    else:
        p, i, v = _ptr_arrays(A, csc=True)
        got, _ = O.dot3_ptr(n, p, i, v, x, y, major_left=trans)
    assert got == exp


def test_dot3_float_close_to_dense():
    rng = np.random.default_rng(24)
    A = sp.random(200, 150, density=0.05, random_state=rng, format="csr")
    x, y = rng.standard_normal(200), rng.standard_normal(150)
    p, i, v = _ptr_arrays(A)
    got, bound = O.dot3_ptr(200, p, i, v, x, y, True)
    assert abs(got - x @ (A @ y)) <= 1e-14 * bound
    # Float32 terms are rounded in Float32: x*a*y of each entry
    got32, _ = O.dot3_ptr(200, p, i, v.astype(np.float32), x.astype(np.float32),
                          y.astype(np.float32), True)
    C = A.tocoo()
    terms = (x.astype(np.float32)[C.row] * C.data.astype(np.float32)) * y.astype(np.float32)[C.col]
    assert abs(got32 - terms.astype(np.float64).sum()) <= 1e-12 * np.abs(terms).sum()


def test_symmetric_vs_scipy_and_known_cases():
    rng = np.random.default_rng(25)
    for seed in range(6):
        A = sp.random(50, 50, density=0.1, random_state=np.random.default_rng(seed), format="csr")
        S = sp.csr_matrix(A + A.T)
        for M, expect in ((A, (abs(A - A.T)).nnz == 0), (S, True)):
            p, i, v = _ptr_arrays(M)
            assert O.symmetric_ptr(50, p, i, v) == expect
            p, i, v = _ptr_arrays(M, csc=True)
            assert O.symmetric_ptr(50, p, i, v) == expect
    # the reference test's non-symmetric pair (test/shared/matrix_coo.jl:96-120): A[1,2] = 1, A[2,1] = 2
    N = sp.csr_matrix(([1.0, 2.0], ([0, 1], [1, 0])), shape=(2, 2))
    assert not O.symmetric_ptr(2, *_ptr_arrays(N))
    # a value differs, a transpose entry is missing, a lone diagonal, an empty matrix
    S = sp.csr_matrix(([1.0, 2.0, 2.0, 5.0], ([0, 0, 2, 1], [0, 2, 0, 1])), shape=(3, 3))
    assert O.symmetric_ptr(3, *_ptr_arrays(S))
    S2 = S.copy(); S2.data[S2.data == 2.0] = [2.0, 2.5]
    assert not O.symmetric_ptr(3, *_ptr_arrays(S2))
    S3 = sp.csr_matrix(([1.0, 2.0], ([0, 0], [0, 2])), shape=(3, 3))
    assert not O.symmetric_ptr(3, *_ptr_arrays(S3))
    assert O.symmetric_ptr(3, np.ones(4, np.int64), np.zeros(0, np.int64), np.zeros(0))
    # a long symmetric row exercises the search: arrow matrix
    n = 300
    rows = np.r_[np.zeros(n - 1, int), np.arange(1, n), np.arange(n)]
    cols = np.r_[np.arange(1, n), np.zeros(n - 1, int), np.arange(n)]
    vals = np.r_[np.arange(1, n), np.arange(1, n), np.ones(n)].astype(np.float64)
    Arw = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    assert O.symmetric_ptr(n, *_ptr_arrays(Arw))
    Arw2 = Arw.tolil(); Arw2[0, 150] = 7.0
    assert not O.symmetric_ptr(n, *_ptr_arrays(sp.csr_matrix(Arw2)))


def test_add_diag_golden_and_scipy():
    d = g("diag_ops.npz")
    p, i, v = O.add_diag(3, d["N_0"], d["N_1"], d["N_2"], 1.0)
    assert np.array_equal(to_sp(3, 3, p, i, v).toarray(), d["E_nodiag"])
    assert p.tolist() == [1, 3, 5, 6] and i.tolist() == [1, 2, 1, 2, 3]  # λ before/after
    p, i, v = O.add_diag(20, d["A4_rowptr"], d["A4_colval"], d["A4_nzval"], 3.0)
    assert np.array_equal(p, d["A4p3_rowptr"]) and np.array_equal(i, d["A4p3_colval"])
    assert np.array_equal(v, d["A4p3_nzval"])
    # CSR + λI keeps a diagonal entry that cancels to zero (no dropzeros, matrix_csr.jl:357-406)
    Z = sp.csr_matrix(([-2.0, 1.0], ([0, 0], [0, 1])), shape=(2, 2))
    p, i, v = O.add_diag(2, *_ptr_arrays(Z), 2.0)
    assert p.tolist() == [1, 3, 4] and i.tolist() == [1, 2, 2] and v.tolist() == [0.0, 1.0, 2.0]
    # COO + λI sums then drops exact zeros (matrix_coo.jl:386-457), column-major order
    C = Z.tocoo()
    r, c, v = O.coo_add_diag(2, (C.row + 1, C.col + 1, C.data), 2.0)
    assert r.tolist() == [1, 2] and c.tolist() == [2, 2] and v.tolist() == [1.0, 2.0]
    A4 =to_sp(20, 20, d["A4_rowptr"], d["A4_colval"], d["A4_nzval"]).tocoo()
    r, c, v = O.coo_add_diag(20, (A4.row + 1, A4.col + 1, A4.data), 3.0)
    exp = sp.csc_matrix(to_sp(20, 20, d["A4p3_rowptr"], d["A4p3_colval"], d["A4p3_nzval"]))
    exp.sort_indices()
    ec = exp.tocoo()
    o = np.lexsort((ec.row, ec.col))
    assert np.array_equal(r, ec.row[o] + 1) and np.array_equal(c, ec.col[o] + 1)
    assert np.array_equal(v, ec.data[o])


def test_kron_diagonal_golden():
    d = g("diag_ops.npz")
    B = (d["B5coo_0"], d["B5coo_1"], d["B5coo_2"])
    A = (d["A5coo_0"], d["A5coo_1"], d["A5coo_2"])
    r, c, v = O.kron_diag_coo(20, 15, d["d5"], B)
    assert np.array_equal(r, d["KD_0"]) and np.array_equal(c, d["KD_1"])
    assert np.array_equal(v, d["KD_2"])
    r, c, v = O.kron_coo_diag(A, d["d5"])
    assert np.array_equal(r, d["KA_0"]) and np.array_equal(c, d["KA_1"])
    assert np.array_equal(v, d["KA_2"])


# ------------------------------------------- complex and integer value types --
# The reference runs every family on Int32, Int64, ComplexF32 and ComplexF64
# values too (test/runtests.jl:34-68).  The oracle's instantiations for them are
# pinned here against exact dense arithmetic: integer-valued complex entries
# make every product and sum exact in any order, and integer values wrap modulo
# 2^64 / 2^32 as numpy's int64 / int32 arithmetic does.
VALUE_KINDS = [np.complex128, np.complex64, np.int64, np.int32]


def typed_csr(m, n, density, seed, dtype, big=False):
    """A random CSR matrix with integer-valued entries of `dtype` (complex: both
    parts; `big`: integers large enough that products wrap)."""
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, random_state=rng, format="csr")
    k = A.data.size
    if np.issubdtype(dtype, np.complexfloating):
        vals = rng.integers(-4, 5, k) + 1j * rng.integers(-4, 5, k)
        vals[vals == 0] = 2 - 1j
    else:
        info = np.iinfo(dtype)
        vals = (rng.integers(info.min // 4, info.max // 4, k) if big
                else rng.integers(-4, 5, k))
        vals[vals == 0] = 3
    A.data = vals.astype(dtype)
    A.sort_indices()
    return A


def typed_dense(shape, seed, dtype, big=False):
    rng = np.random.default_rng(seed)
    if np.issubdtype(dtype, np.complexfloating):
        return (rng.integers(-4, 5, shape) + 1j * rng.integers(-4, 5, shape)).astype(dtype)
    info = np.iinfo(dtype)
    if big:
        return rng.integers(info.min // 4, info.max // 4, shape).astype(dtype)
    return rng.integers(-4, 5, shape).astype(dtype)


def dense_matmul(A, B):
    """A @ B on dense arrays in A's type: exact for integer-valued complex data,
    wrapping for integers (numpy's integer arithmetic wraps)."""
    with np.errstate(over="ignore"):
        return np.asarray(A) @ np.asarray(B)


@pytest.mark.parametrize("dtype", VALUE_KINDS)
def test_products_all_value_types_exact(dtype):
    big = np.issubdtype(dtype, np.integer)
    A = typed_csr(60, 45, 0.15, 601, dtype, big=big)
    Ad = A.toarray()
    rp, cv, nz = arrs(A)
    x = typed_dense(45, 602, dtype, big=big)
    X = typed_dense((45, 3), 603, dtype, big=big)
    assert np.array_equal(O.spmm_csr_N(60, rp, cv, nz, x), dense_matmul(Ad, x))
    assert np.array_equal(O.spmm_csr_N(60, rp, cv, nz, X), dense_matmul(Ad, X))
    xt = typed_dense(60, 604, dtype, big=big)
    assert np.array_equal(O.spmm_csr_T(60, 45, rp, cv, nz, xt), dense_matmul(Ad.T, xt))
    if np.issubdtype(dtype, np.complexfloating):
        # adjoint(A) * x conjugates the stored values (CONJA)
        assert np.array_equal(O.spmm_csr_T(60, 45, rp, cv, nz, xt, conj_a=True),
                              dense_matmul(Ad.conj().T, xt))
        r, c = np.nonzero(Ad)
        got = O.spmm_coo(True, 45, r + 1, c + 1, Ad[r, c], xt, conj_a=True)
        assert np.array_equal(got, dense_matmul(Ad.conj().T, xt))
    # alpha, beta: C = beta*C0 + alpha*A*x
    c0 = typed_dense(60, 605, dtype)
    if np.issubdtype(dtype, np.complexfloating):
        alpha, beta = dtype(2 - 1j), dtype(3 + 0j)
    else:
        alpha, beta = dtype(3), dtype(-2)
    got = O.spmm_csr_N(60, rp, cv, nz, x, C=c0, alpha=alpha, beta=beta)
    with np.errstate(over="ignore"):
        want = (beta * c0 + dense_matmul(Ad, x) * alpha).astype(dtype)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", VALUE_KINDS)
def test_structural_all_value_types_vs_scipy(dtype):
    A = typed_csr(40, 30, 0.2, 611, dtype)
    B = typed_csr(40, 30, 0.2, 612, dtype)
    # + with dropped exact zeros: integer values make A + (-A) cancel exactly
    ra, ca, va = arrs(A)
    rb, cb, vb = arrs(B)
    nnz = O.csr_add(40, (ra, ca, va), (rb, cb, vb))
    S = (A + B).tocsr()
    S.eliminate_zeros()
    S.sort_indices()
    want = arrs(S)
    for got, w in zip(nnz, want):
        assert np.array_equal(got, w)
    neg = O.csr_add(40, (ra, ca, va), (ra, ca, (-va).astype(dtype)))
    assert neg[1].size == 0
    # kron: values bit-exact (one multiply each), pattern of scipy's
    K = O.kron_csr(40, 30, 40, 30, (ra, ca, va), (rb, cb, vb))
    Ks = sp.kron(A, B, format="csr")
    Ks.sort_indices()
    assert np.array_equal(K[0], Ks.indptr + 1) and np.array_equal(K[1], Ks.indices + 1)
    assert np.array_equal(K[2], Ks.data.astype(dtype))
    # SpGEMM: values equal the dense product, pattern the structural product
    # (cancellation zeros are kept, as the reference keeps them)
    Bt = typed_csr(30, 25, 0.2, 613, dtype)
    rt, ct, vt = arrs(Bt)
    rp, cv, nz = O.spgemm_csr(40, 25, (ra, ca, va), (rt, ct, vt))
    C = sp.csr_matrix((nz, cv - 1, rp - 1), shape=(40, 25))
    assert np.array_equal(C.toarray(), dense_matmul(A.toarray(), Bt.toarray()))
    pattern = ((np.abs(A.toarray()) > 0).astype(np.int64) @
               (np.abs(Bt.toarray()) > 0).astype(np.int64)) > 0
    got_pat = np.zeros((40, 25), dtype=bool)
    got_pat[np.repeat(np.arange(40), np.diff(rp)), cv - 1] = True
    assert np.array_equal(got_pat, pattern) and C.nnz == pattern.sum()
    # conversions move values unchanged
    r, c, v = O.csr_to_coo(40, ra, ca, va)
    assert v.dtype == np.dtype(dtype) and np.array_equal(v, va)
    cp, rv, vv = O.csr_to_csc(40, 30, ra, ca, va)
    Ac = A.tocsc()
    Ac.sort_indices()
    assert np.array_equal(vv, Ac.data) and np.array_equal(rv, Ac.indices + 1)


def test_complex_dropzeros_and_zero_kinds():
    """iszero: both parts zero, -0.0 included; (0, 1) is kept."""
    v = np.array([0 + 0j, -0.0 + 0j, 0 + 1j, 1 + 0j, complex(0.0, -0.0)], dtype=np.complex128)
    rp = np.array([1, 4, 6])
    cv = np.array([1, 2, 3, 1, 2])
    p, c, vv = O.csr_dropzeros(2, rp, cv, v)
    assert np.array_equal(p, [1, 2, 3]) and np.array_equal(c, [3, 1])
    assert np.array_equal(vv, [1j, 1 + 0j])
