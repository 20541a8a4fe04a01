"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

Sources, all independent of the oracle under test:
  * literal known answers held by the reference's own tests and docs
    (docs/src/index.md:47-61; test/shared/matrix_csr.jl:35-38,49-59,386-392;
    test/shared/conversions.jl:128-152);
  * scipy.sparse results on seeded inputs (canonical CSR/CSC, sums with
    explicit zeros eliminated, products on cancellation-free inputs, kron);
  * dense numpy arithmetic on integer-valued matrices (exact in floating point).
Every index array is stored ONE-BASED, as the reference stores it.
"""
import os

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def csr_arrays(A):
    B = sp.csr_matrix(A)
    B.sum_duplicates()
    B.sort_indices()
    return (B.indptr.astype(np.int64) + 1, B.indices.astype(np.int64) + 1, B.data.copy())


def csc_arrays(A):
    B = sp.csc_matrix(A)
    B.sum_duplicates()
    B.sort_indices()
    return (B.indptr.astype(np.int64) + 1, B.indices.astype(np.int64) + 1, B.data.copy())


def coo_colmajor(A):
    """findnz order of a SparseMatrixCSC (column-major)."""
    C = sp.csc_matrix(A)
    C.sum_duplicates()
    C.sort_indices()
    C = C.tocoo()
    o = np.lexsort((C.row, C.col))
    return C.row[o].astype(np.int64) + 1, C.col[o].astype(np.int64) + 1, C.data[o].copy()


def coo_rowmajor(A):
    C = sp.csr_matrix(A)
    C.sum_duplicates()
    C.sort_indices()
    C = C.tocoo()
    o = np.lexsort((C.col, C.row))
    return C.row[o].astype(np.int64) + 1, C.col[o].astype(np.int64) + 1, C.data[o].copy()


def save(name, **arrays):
    np.savez_compressed(os.path.join(HERE, name), **arrays)


def main():
    # docs/src/index.md:47-61 — sparse([1,2,1,3],[1,1,2,3],[2,3,4,5],3,3) * [1,2,3] == [10,3,15]
    A = sp.coo_matrix(([2.0, 3.0, 4.0, 5.0], ([0, 1, 0, 2], [0, 0, 1, 2])), shape=(3, 3))
    rp, cv, nz = csr_arrays(A)
    save("doc_spmv.npz", m=3, n=3, rowptr=rp, colval=cv, nzval=nz,
         x=np.array([1.0, 2.0, 3.0]), y=np.array([10.0, 3.0, 15.0]))

    # test/shared/matrix_csr.jl:35-38,49-59 — B = sparse([1,2,1],[1,1,2],[1,2,3],2,2);
    # its CSR storage equals SparseMatrixCSC(transpose(B))'s (colptr, rowval, nzval).
    B = sp.coo_matrix(([1.0, 2.0, 3.0], ([0, 1, 0], [0, 0, 1])), shape=(2, 2))
    Bt = sp.csc_matrix(B.T)
    Bt.sort_indices()
    save("csr_storage_2x2.npz", m=2, n=2,
         rowptr=Bt.indptr.astype(np.int64) + 1, colval=Bt.indices.astype(np.int64) + 1,
         nzval=Bt.data.copy(),
         coo_rows=np.array([1, 2, 1]), coo_cols=np.array([1, 1, 2]), coo_vals=np.array([1.0, 2.0, 3.0]))

    # test/shared/matrix_csr.jl:386-392 — overlapping diagonals, 50 x 40
    Ao = sp.coo_matrix(([1.0, 2.0, 3.0], ([0, 1, 2], [0, 1, 2])), shape=(50, 40))
    Bo = sp.coo_matrix(([4.0, 5.0, 6.0], ([0, 1, 3], [0, 1, 3])), shape=(50, 40))
    Co = (Ao + Bo).tocsr()
    save("overlap_add.npz", m=50, n=40,
         **{f"A_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(Ao))},
         **{f"B_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(Bo))},
         **{f"C_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(Co))})

    # test/shared/conversions.jl:128-152 — single element, full row, full column (3 x 4)
    single = sp.coo_matrix(([5.0], ([0], [1])), shape=(3, 4))
    fullrow = sp.coo_matrix(([1.0, 2.0, 3.0, 4.0], ([0, 0, 0, 0], [0, 1, 2, 3])), shape=(3, 4))
    fullcol = sp.coo_matrix(([1.0, 2.0, 3.0], ([0, 1, 2], [1, 1, 1])), shape=(3, 4))
    out = {}
    for name, M in (("single", single), ("fullrow", fullrow), ("fullcol", fullcol)):
        for tag, arrs in (("csr", csr_arrays(M)), ("csc", csc_arrays(M)),
                          ("coo_colmajor", coo_colmajor(M)), ("coo_rowmajor", coo_rowmajor(M))):
            for i, a in enumerate(arrs):
                out[f"{name}_{tag}_{i}"] = a
        out[f"{name}_dense"] = M.toarray()
    save("conv_edges.npz", m=3, n=4, **out)

    # test/shared/conversions.jl:11 — sprand(100, 200, 0.05), canonical forms by scipy
    rng = np.random.default_rng(11)
    R = sp.random(100, 200, density=0.05, random_state=rng, format="csc")
    out = {}
    for tag, arrs in (("csr", csr_arrays(R)), ("csc", csc_arrays(R)),
                      ("coo_colmajor", coo_colmajor(R)), ("coo_rowmajor", coo_rowmajor(R))):
        for i, a in enumerate(arrs):
            out[f"{tag}_{i}"] = a
    save("conv_random.npz", m=100, n=200, **out)

    # sparse + sparse on small integer values (cancellations produce exact zeros)
    rng = np.random.default_rng(12)
    def intmat(m, n, d):
        M = sp.random(m, n, density=d, random_state=rng, format="csr")
        M.data = rng.integers(-3, 4, size=M.data.size).astype(np.float64)
        M.data[M.data == 0] = 1.0
        return M
    A1, B1 = intmat(50, 40, 0.1), intmat(50, 40, 0.15)
    B1 = B1 + sp.csr_matrix((-A1.data[:20], A1.indices[:20], np.minimum(A1.indptr, 20)),
                            shape=A1.shape)  # force exact cancellations
    B1 = sp.csr_matrix(B1)
    B1.eliminate_zeros()
    C1 = (A1 + B1).tocsr()
    C1.eliminate_zeros()
    save("add_random.npz", m=50, n=40,
         **{f"A_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(A1))},
         **{f"B_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(B1))},
         **{f"C_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(C1))},
         **{f"Ccoo_{i}": a for i, a in enumerate(coo_colmajor(C1))},
         **{f"Acoo_{i}": a for i, a in enumerate(coo_colmajor(A1))},
         **{f"Bcoo_{i}": a for i, a in enumerate(coo_colmajor(B1))})

    # SpGEMM on cancellation-free (positive) values: 50x40 @0.1 times 40x30 @0.15
    rng = np.random.default_rng(13)
    A2 = sp.random(50, 40, density=0.1, random_state=rng, format="csr")
    B2 = sp.random(40, 30, density=0.15, random_state=rng, format="csr")
    C2 = (A2 @ B2).tocsr()
    C2.sort_indices()
    save("spgemm_random.npz", m=50, k=40, n=30,
         **{f"A_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(A2))},
         **{f"B_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(B2))},
         **{f"C_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(C2))})

    # kron: test/shared/matrix_csr.jl:449-452 shapes (30x25 @0.1, 20x15 @0.1)
    rng = np.random.default_rng(14)
    A3 = sp.random(30, 25, density=0.1, random_state=rng, format="csr")
    B3 = sp.random(20, 15, density=0.1, random_state=rng, format="csr")
    ra, ca, va = coo_rowmajor(A3)
    rb, cb, vb = coo_rowmajor(B3)
    K = sp.kron(sp.coo_matrix((va, (ra - 1, ca - 1)), shape=A3.shape),
                sp.coo_matrix((vb, (rb - 1, cb - 1)), shape=B3.shape), format="coo")
    save("kron_random.npz", mA=30, nA=25, mB=20, nB=15,
         **{f"A_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(A3))},
         **{f"B_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(B3))},
         **{f"C_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(K))},
         Kcoo_0=K.row.astype(np.int64) + 1, Kcoo_1=K.col.astype(np.int64) + 1, Kcoo_2=K.data)

    # A + I on a matrix with no diagonal entries, test/shared/matrix_coo.jl:277-281:
    # sparse([1, 2], [2, 1], [1, 2], 3, 3) + I == [1 1 0; 2 1 0; 0 0 1]
    N = sp.coo_matrix(([1.0, 2.0], ([0, 1], [1, 0])), shape=(3, 3))
    E = np.array([[1.0, 1.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    # A + 3I on test/shared/matrix_coo.jl:237-256's shape (20 x 20 @ 0.2), values in (0, 1)
    rng = np.random.default_rng(15)
    A4 = sp.random(20, 20, density=0.2, random_state=rng, format="csr")
    A4p3 = (A4 + 3.0 * sp.identity(20, format="csr")).tocsr()
    A4p3.sort_indices()
    # kron with a Diagonal, test/shared/matrix_coo.jl:437-467 (30x25, 20x15 @ 0.1, D = rand(4)):
    # scipy's kron of COO operands lists entries A-major / B-minor, the reference's order
    rng = np.random.default_rng(16)
    A5 = sp.random(30, 25, density=0.1, random_state=rng, format="csr")
    B5 = sp.random(20, 15, density=0.1, random_state=rng, format="csr")
    d5 = rng.random(4)
    Dc = sp.coo_matrix((d5, (np.arange(4), np.arange(4))), shape=(4, 4))
    rb, cb, vb = coo_colmajor(B5)
    ra, ca, va = coo_colmajor(A5)
    Bc = sp.coo_matrix((vb, (rb - 1, cb - 1)), shape=B5.shape)
    Ac = sp.coo_matrix((va, (ra - 1, ca - 1)), shape=A5.shape)
    KD = sp.kron(Dc, Bc, format="coo")
    KA = sp.kron(Ac, Dc, format="coo")
    save("diag_ops.npz", E_nodiag=E,
         **{f"N_{i}": a for i, a in enumerate(csr_arrays(N))},
         **{f"A4_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(A4))},
         **{f"A4p3_{k}": v for k, v in zip(("rowptr", "colval", "nzval"), csr_arrays(A4p3))},
         d5=d5, **{f"B5coo_{i}": a for i, a in enumerate((rb, cb, vb))},
         **{f"A5coo_{i}": a for i, a in enumerate((ra, ca, va))},
         KD_0=KD.row.astype(np.int64) + 1, KD_1=KD.col.astype(np.int64) + 1, KD_2=KD.data,
         KA_0=KA.row.astype(np.int64) + 1, KA_1=KA.col.astype(np.int64) + 1, KA_2=KA.data)


if __name__ == "__main__":
    main()
