/*
 * gsa_oracle.c — CPU restatement of GenericSparseArrays.jl v0.1.2's sparse
 * kernel layer.  TEST INFRASTRUCTURE ONLY: the parity checker for the HIP
 * library.  Only tests/, the smoke check and bench.py's CPU-baseline leg load
 * it; no product code path calls it.
 *
 * Conventions follow the reference exactly: indices are one-based int64,
 * dense operands column-major, loops run in storage order, products and sums
 * are separate IEEE operations (built with -ffp-contract=off, no FMA), and the
 * index primitives are Base semantics (stable sortperm, inclusive cumsum,
 * searchsortedfirst) as in ext/GenericSparseArraysJLArraysExt.jl:6-9.
 *
 * Scatter-shaped products (CSR-T, CSC-N, COO) use @atomic in the reference, so
 * their summation order there is not defined; here they run in storage order.
 *
 * Each function cites the reference file:line it restates.
 */
#include <complex.h>
#undef I /* complex.h's imaginary unit: I is the index type here */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t I;

void ora_sortperm(I n, const I* keys, I* perm);
I ora_searchsortedfirst(const I* v, I n, I x);
void ora_sort_i64(I* v, I n);

/* ---------------------------------------------------------------------------
 * beta handling: `β != one(β) && LinearAlgebra._rmul_or_fill!(C, β)`
 * (src/matrix_csr/matrix_csr.jl:192): β == 0 fills zeros (clearing NaN),
 * β == 1 leaves C, otherwise C .*= β.
 * ------------------------------------------------------------------------- */
#define DEFINE_BETA(T, S)                                                           \
static void beta_apply_##S(I rows, I k, T* C, I ldc, T beta) {                      \
    if (beta == (T)1) return;                                                       \
    for (I j = 0; j < k; ++j)                                                       \
        for (I i = 0; i < rows; ++i) {                                              \
            if (beta == (T)0) C[i + j * ldc] = (T)0;                                \
            else C[i + j * ldc] = SCAL_##S(C[i + j * ldc], beta);                   \
        }                                                                           \
}

/* B accessor for opB: N reads B[i + j*ldb]; T (1) and C (2) read B[j + i*ldb],
 * C conjugating it (TRANSB / CONJB, src/matrix_csr/matrix_csr_kernels.jl:16-18). */
#define BAT(B, i, j, ldb, opB) ((opB) ? (B)[(j) + (i) * (ldb)] : (B)[(i) + (j) * (ldb)])

/* Value-type helpers.  SCAL: x * s for a scalar s (alpha, beta); a complex
 * scalar with a zero imaginary part multiplies the parts separately, as
 * Julia's Complex * Real does (C's complex * real is also by parts).  CONJA
 * conjugates a stored value when the current product is an Adjoint
 * (ora_conj_a, set by ora_set_conj_a); CONJB conjugates B under opB == 2.
 * Integer values use unsigned arithmetic of their width: two's-complement
 * wrapping, as Julia's Int32/Int64 arithmetic wraps. */
typedef double _Complex zc128;
typedef float _Complex zc64;
static int ora_conj_a = 0;
void ora_set_conj_a(int c) { ora_conj_a = c; }
#define SCAL_f64(x, s) ((x) * (s))
#define SCAL_f32(x, s) ((x) * (s))
#define SCAL_i64(x, s) ((x) * (s))
#define SCAL_i32(x, s) ((x) * (s))
#define SCAL_c128(x, s) (cimag(s) == 0 ? (x) * creal(s) : (x) * (s))
#define SCAL_c64(x, s) (cimagf(s) == 0 ? (x) * crealf(s) : (x) * (s))
#define CONJA_f64(v) (v)
#define CONJA_f32(v) (v)
#define CONJA_i64(v) (v)
#define CONJA_i32(v) (v)
#define CONJA_c128(v) (ora_conj_a ? conj(v) : (v))
#define CONJA_c64(v) (ora_conj_a ? conjf(v) : (v))
#define CONJB_f64(v, op) (v)
#define CONJB_f32(v, op) (v)
#define CONJB_i64(v, op) (v)
#define CONJB_i32(v, op) (v)
#define CONJB_c128(v, op) ((op) == 2 ? conj(v) : (v))
#define CONJB_c64(v, op) ((op) == 2 ? conjf(v) : (v))

#define DEFINE_OPS(T, S)                                                            \
DEFINE_BETA(T, S)                                                                   \
/* kernel_spmatmul_csr_N! (src/matrix_csr/matrix_csr_kernels.jl:1-22): per        \
 * (row, k): tmp = 0; tmp += nzval[j]*B[colval[j],k] in storage order;             \
 * C[row,k] += tmp*α.  Rows are independent, so OpenMP over rows keeps every       \
 * result bit-identical to the sequential loop. */                                 \
void ora_spmm_csr_N_##S(I m, I k, const I* rowptr, const I* colval, const T* nzval, \
                        const T* B, I ldb, int opB, T* C, I ldc, const T* alpha_p,  \
                        const T* beta_p, int threads) {                             \
    const T alpha = *alpha_p, beta = *beta_p;                                       \
    beta_apply_##S(m, k, C, ldc, beta);                                             \
    (void)threads;                                                                  \
    _Pragma("omp parallel for schedule(dynamic, 1024) num_threads(threads)")        \
    for (I row = 0; row < m; ++row) {                                               \
        for (I kk = 0; kk < k; ++kk) {                                              \
            T tmp = (T)0;                                                           \
            for (I j = rowptr[row] - 1; j < rowptr[row + 1] - 1; ++j) {             \
                T p = CONJA_##S(nzval[j]) * CONJB_##S(BAT(B, colval[j] - 1, kk, ldb, opB), opB); \
                tmp = tmp + p;                                                      \
            }                                                                       \
            T t = SCAL_##S(tmp, alpha);                                             \
            C[row + kk * ldc] = C[row + kk * ldc] + t;                              \
        }                                                                           \
    }                                                                               \
}                                                                                   \
/* kernel_spmatmul_csr_T! (matrix_csr_kernels.jl:24-45): axj = B[row,k]*α;         \
 * C[colval[j],k] += nzval[j]*axj (atomic there; storage order here).              \
 * Also kernel_spmatmul_csc_N! (src/matrix_csc/matrix_csc_kernels.jl:1-22) on      \
 * (colptr, rowval). */                                                             \
void ora_spmm_csr_T_##S(I m, I n, I k, const I* rowptr, const I* colval,            \
                        const T* nzval, const T* B, I ldb, int opB, T* C, I ldc,    \
                        const T* alpha_p, const T* beta_p) {                        \
    const T alpha = *alpha_p, beta = *beta_p;                                       \
    beta_apply_##S(n, k, C, ldc, beta);                                             \
    for (I kk = 0; kk < k; ++kk)                                                    \
        for (I row = 0; row < m; ++row) {                                           \
            T axj = SCAL_##S(CONJB_##S(BAT(B, row, kk, ldb, opB), opB), alpha);     \
            for (I j = rowptr[row] - 1; j < rowptr[row + 1] - 1; ++j) {             \
                T p = CONJA_##S(nzval[j]) * axj;                                    \
                C[(colval[j] - 1) + kk * ldc] = C[(colval[j] - 1) + kk * ldc] + p;  \
            }                                                                       \
        }                                                                           \
}                                                                                   \
/* kernel_spmatmul_coo_N!/T! (src/matrix_coo/matrix_coo_kernels.jl:1-43): one      \
 * update per stored entry, C[row,k] += nzval*(B[col,k]*α) (T swaps row/col). */    \
void ora_spmm_coo_##S(int trans, I rows_out, I k, I nnz, const I* rowind,           \
                      const I* colind, const T* nzval, const T* B, I ldb, int opB,  \
                      T* C, I ldc, const T* alpha_p, const T* beta_p) {             \
    const T alpha = *alpha_p, beta = *beta_p;                                       \
    beta_apply_##S(rows_out, k, C, ldc, beta);                                      \
    for (I kk = 0; kk < k; ++kk)                                                    \
        for (I i = 0; i < nnz; ++i) {                                               \
            I r = trans ? colind[i] - 1 : rowind[i] - 1;                            \
            I c = trans ? rowind[i] - 1 : colind[i] - 1;                            \
            T axj = SCAL_##S(CONJB_##S(BAT(B, c, kk, ldb, opB), opB), alpha);       \
            T p = CONJA_##S(nzval[i]) * axj;                                        \
            C[r + kk * ldc] = C[r + kk * ldc] + p;                                  \
        }                                                                           \
}                                                                                   \
/* kernel_csr_to_coo! (src/conversions/conversion_kernels.jl:20-35). */             \
void ora_csr_to_coo_##S(I m, const I* rowptr, const I* colval, const T* nzval,      \
                        I* rowind, I* colind, T* nzout) {                           \
    for (I row = 0; row < m; ++row)                                                 \
        for (I j = rowptr[row] - 1; j < rowptr[row + 1] - 1; ++j) {                 \
            rowind[j] = row + 1; colind[j] = colval[j]; nzout[j] = nzval[j];        \
        }                                                                           \
}                                                                                   \
/* GenericSparseMatrixCSR(::COO) (src/conversions/conversions.jl:282-323): keys    \
 * rowind*n + colind (kernel_make_csr_keys!, conversion_kernels.jl:49-57),         \
 * stable sortperm, gather, rowptr[1:m] = searchsortedfirst(rowind_sorted, 1:m),   \
 * rowptr[m+1] = nnz+1; nnz == 0 gives all ones (conversions.jl:289-295). */       \
void ora_coo_to_csr_##S(I m, I n, I nnz, const I* rowind, const I* colind,          \
                        const T* nzval, I* rowptr, I* colval, T* nzout) {           \
    if (nnz == 0) { for (I r = 0; r <= m; ++r) rowptr[r] = 1; return; }             \
    I* keys = (I*)malloc(sizeof(I) * nnz);                                          \
    I* perm = (I*)malloc(sizeof(I) * nnz);                                          \
    for (I i = 0; i < nnz; ++i) keys[i] = rowind[i] * n + colind[i];                \
    ora_sortperm(nnz, keys, perm);                                                  \
    I* rs = (I*)malloc(sizeof(I) * nnz);                                            \
    for (I i = 0; i < nnz; ++i) {                                                   \
        rs[i] = rowind[perm[i]]; colval[i] = colind[perm[i]]; nzout[i] = nzval[perm[i]]; \
    }                                                                               \
    for (I r = 1; r <= m; ++r) rowptr[r - 1] = ora_searchsortedfirst(rs, nnz, r);   \
    rowptr[m] = nnz + 1;                                                            \
    free(rs); free(keys); free(perm);                                               \
}                                                                                   \
/* Base.:+(::CSR, ::CSR) (src/matrix_csr/matrix_csr.jl:290-355): two-pointer       \
 * merge per row (kernel_count_nnz_per_row_csr!/kernel_merge_csr!,                 \
 * matrix_csr_kernels.jl:148-248), equal columns summed vA+vB, then dropzeros      \
 * (matrix_csr.jl:723-745, kernels :410-443).  Returns nnz of the result; with     \
 * rowptr_out == NULL only counts. */                                               \
I ora_csr_add_##S(I m, const I* rpA, const I* cA, const T* vA,                      \
                  const I* rpB, const I* cB, const T* vB,                           \
                  I* rowptr_out, I* col_out, T* val_out) {                          \
    I pos = 0;                                                                      \
    if (rowptr_out) rowptr_out[0] = 1;                                              \
    for (I row = 0; row < m; ++row) {                                               \
        I a = rpA[row] - 1, ea = rpA[row + 1] - 1, b = rpB[row] - 1, eb = rpB[row + 1] - 1; \
        while (a < ea || b < eb) {                                                  \
            I c; T v;                                                               \
            if (a < ea && (b >= eb || cA[a] < cB[b])) { c = cA[a]; v = vA[a]; ++a; } \
            else if (b < eb && (a >= ea || cA[a] > cB[b])) { c = cB[b]; v = vB[b]; ++b; } \
            else { c = cA[a]; v = vA[a] + vB[b]; ++a; ++b; }                        \
            if (v != (T)0) {                                                        \
                if (col_out) { col_out[pos] = c; val_out[pos] = v; }                \
                ++pos;                                                              \
            }                                                                       \
        }                                                                           \
        if (rowptr_out) rowptr_out[row + 1] = pos + 1;                              \
    }                                                                               \
    return pos;                                                                     \
}                                                                                   \
/* SparseArrays.dropzeros(::CSR) (src/matrix_csr/matrix_csr.jl:650-683,723-745). */ \
I ora_csr_dropzeros_##S(I m, const I* rowptr, const I* colval, const T* nzval,      \
                        I* rowptr_out, I* col_out, T* val_out) {                    \
    I pos = 0;                                                                      \
    if (rowptr_out) rowptr_out[0] = 1;                                              \
    for (I row = 0; row < m; ++row) {                                               \
        for (I j = rowptr[row] - 1; j < rowptr[row + 1] - 1; ++j)                   \
            if (nzval[j] != (T)0) {                                                 \
                if (col_out) { col_out[pos] = colval[j]; val_out[pos] = nzval[j]; } \
                ++pos;                                                              \
            }                                                                       \
        if (rowptr_out) rowptr_out[row + 1] = pos + 1;                              \
    }                                                                               \
    return pos;                                                                     \
}                                                                                   \
/* Base.:+(::COO, ::COO) (src/matrix_coo/matrix_coo.jl:300-384): concatenate A     \
 * then B, keys colind*m + rowind (column-major, kernel_make_csc_keys!), stable    \
 * sortperm, gather, first-of-run mask (kernel_mark_unique_coo!,                   \
 * matrix_coo_kernels.jl:260-275), each run summed left to right                   \
 * (kernel_compact_coo!, :278-310), then dropzeros (matrix_coo.jl:909-935).        \
 * Returns nnz of the result; outputs may be NULL to count only. */                \
I ora_coo_add_##S(I m, I nA, const I* rA, const I* cA, const T* vA,                 \
                  I nB, const I* rB, const I* cB, const T* vB,                      \
                  I* r_out, I* c_out, T* v_out) {                                   \
    I nc = nA + nB;                                                                 \
    if (nc == 0) return 0;                                                          \
    I* r = (I*)malloc(sizeof(I) * nc); I* c = (I*)malloc(sizeof(I) * nc);           \
    T* v = (T*)malloc(sizeof(T) * nc); I* keys = (I*)malloc(sizeof(I) * nc);        \
    I* perm = (I*)malloc(sizeof(I) * nc);                                           \
    for (I i = 0; i < nA; ++i) { r[i] = rA[i]; c[i] = cA[i]; v[i] = vA[i]; }        \
    for (I i = 0; i < nB; ++i) { r[nA + i] = rB[i]; c[nA + i] = cB[i]; v[nA + i] = vB[i]; } \
    for (I i = 0; i < nc; ++i) keys[i] = c[i] * m + r[i];                           \
    ora_sortperm(nc, keys, perm);                                                   \
    I pos = 0;                                                                      \
    I i = 0;                                                                        \
    while (i < nc) {                                                                \
        I ri = r[perm[i]], ci = c[perm[i]];                                         \
        T s = v[perm[i]];                                                           \
        I j = i + 1;                                                                \
        while (j < nc && r[perm[j]] == ri && c[perm[j]] == ci) { s = s + v[perm[j]]; ++j; } \
        if (s != (T)0) {                                                            \
            if (r_out) { r_out[pos] = ri; c_out[pos] = ci; v_out[pos] = s; }        \
            ++pos;                                                                  \
        }                                                                           \
        i = j;                                                                      \
    }                                                                               \
    free(r); free(c); free(v); free(keys); free(perm);                              \
    return pos;                                                                     \
}                                                                                   \
/* SparseArrays.dropzeros(::COO) (src/matrix_coo/matrix_coo.jl:909-935):          \
 * mask nzval .!= 0, filter in order. */                                            \
I ora_coo_dropzeros_##S(I nnz, const I* r, const I* c, const T* v,                  \
                        I* r_out, I* c_out, T* v_out) {                             \
    I pos = 0;                                                                      \
    for (I i = 0; i < nnz; ++i)                                                     \
        if (v[i] != (T)0) {                                                         \
            if (r_out) { r_out[pos] = r[i]; c_out[pos] = c[i]; v_out[pos] = v[i]; } \
            ++pos;                                                                  \
        }                                                                           \
    return pos;                                                                     \
}                                                                                   \
/* kernel_kron_coo! (src/matrix_coo/matrix_coo_kernels.jl:144-183): idx -> idx_A = \
 * (idx-1)÷nnzB+1, idx_B = (idx-1)%nnzB+1; row (iA-1)*mB+iB, col (jA-1)*nB+jB,     \
 * value vA*vB.  A-major / B-minor order. */                                        \
void ora_kron_coo_##S(I mB, I nB, I nA_, const I* rA, const I* cA, const T* vA,     \
                      I nB_, const I* rB, const I* cB, const T* vB,                 \
                      I* r_out, I* c_out, T* v_out) {                               \
    for (I a = 0; a < nA_; ++a)                                                     \
        for (I b = 0; b < nB_; ++b) {                                               \
            I idx = a * nB_ + b;                                                    \
            r_out[idx] = (rA[a] - 1) * mB + rB[b];                                  \
            c_out[idx] = (cA[a] - 1) * nB + cB[b];                                  \
            v_out[idx] = vA[a] * vB[b];                                             \
        }                                                                           \
}                                                                                   \
/* Base.:*(::CSR, ::CSR) (src/matrix_csr/matrix_csr.jl:498-572) with               \
 * kernel_count_nnz_spgemm_csr!/kernel_spgemm_csr! (matrix_csr_kernels.jl:254-343):\
 * per row, acc[j] += a_ik*b_kj in (A-row order, B-row order), flags set whatever  \
 * the value (cancellation zeros kept), columns emitted ascending.  The dense      \
 * per-row workspace of the reference is replaced by a sparse accumulator with the \
 * same arithmetic.  Returns nnz(C); with rowptr_out == NULL only counts. */        \
I ora_spgemm_csr_##S(I m, I n, const I* rpA, const I* cA, const T* vA,              \
                     const I* rpB, const I* cB, const T* vB,                        \
                     I* rowptr_out, I* col_out, T* val_out) {                       \
    T* acc = (T*)calloc((size_t)(n > 0 ? n : 1), sizeof(T));                        \
    unsigned char* flag = (unsigned char*)calloc((size_t)(n > 0 ? n : 1), 1);       \
    I* list = (I*)malloc(sizeof(I) * (size_t)(n > 0 ? n : 1));                      \
    I pos = 0;                                                                      \
    if (rowptr_out) rowptr_out[0] = 1;                                              \
    for (I row = 0; row < m; ++row) {                                               \
        I cnt = 0;                                                                  \
        for (I ia = rpA[row] - 1; ia < rpA[row + 1] - 1; ++ia) {                    \
            I kk = cA[ia] - 1;                                                      \
            T a = vA[ia];                                                           \
            for (I ib = rpB[kk] - 1; ib < rpB[kk + 1] - 1; ++ib) {                  \
                I j = cB[ib] - 1;                                                   \
                T p = a * vB[ib];                                                   \
                if (!flag[j]) { flag[j] = 1; acc[j] = (T)0; list[cnt++] = j; }      \
                acc[j] = acc[j] + p;                                                \
            }                                                                       \
        }                                                                           \
        ora_sort_i64(list, cnt);                                                    \
        for (I t = 0; t < cnt; ++t) {                                               \
            I j = list[t];                                                          \
            if (col_out) { col_out[pos] = j + 1; val_out[pos] = acc[j]; }           \
            ++pos;                                                                  \
            flag[j] = 0;                                                            \
        }                                                                           \
        if (rowptr_out) rowptr_out[row + 1] = pos + 1;                              \
    }                                                                               \
    free(acc); free(flag); free(list);                                              \
    return pos;                                                                     \
}

/* ------------------------- index primitives ---------------------------------
 * Base.sortperm on integer keys is stable; here a bottom-up merge sort of the
 * permutation (ties keep input order).  helpers.jl:2 (_sortperm_AK) with the
 * Base semantics of ext/GenericSparseArraysJLArraysExt.jl:6. */
void ora_sortperm(I n, const I* keys, I* perm) {
    I* tmp = (I*)malloc(sizeof(I) * (n > 0 ? n : 1));
    for (I i = 0; i < n; ++i) perm[i] = i;
    for (I width = 1; width < n; width *= 2) {
        for (I lo = 0; lo < n; lo += 2 * width) {
            I mid = lo + width < n ? lo + width : n;
            I hi = lo + 2 * width < n ? lo + 2 * width : n;
            I a = lo, b = mid, o = lo;
            while (a < mid && b < hi) {
                if (keys[perm[b]] < keys[perm[a]]) tmp[o++] = perm[b++];
                else tmp[o++] = perm[a++];
            }
            while (a < mid) tmp[o++] = perm[a++];
            while (b < hi) tmp[o++] = perm[b++];
        }
        memcpy(perm, tmp, sizeof(I) * n);
    }
    free(tmp);
}

/* searchsortedfirst(v, x): first one-based index i with v[i] >= x, or n+1
 * (helpers.jl:4). */
I ora_searchsortedfirst(const I* v, I n, I x) {
    I lo = 0, hi = n;
    while (lo < hi) {
        I mid = lo + (hi - lo) / 2;
        if (v[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo + 1;
}

static int cmp_i64(const void* a, const void* b) {
    I x = *(const I*)a, y = *(const I*)b;
    return (x > y) - (x < y);
}
void ora_sort_i64(I* v, I n) { qsort(v, (size_t)n, sizeof(I), cmp_i64); }

/* Inclusive cumsum (helpers.jl:3). */
void ora_cumsum(I n, const I* x, I* out) {
    I s = 0;
    for (I i = 0; i < n; ++i) { s += x[i]; out[i] = s; }
}

DEFINE_OPS(double, f64)
DEFINE_OPS(float, f32)
DEFINE_OPS(zc128, c128)
DEFINE_OPS(zc64, c64)
DEFINE_OPS(uint64_t, i64)
DEFINE_OPS(uint32_t, i32)

/* ---------------------------------------------------------------------------
 * Reductions over the stored entries, A + λI.  "ptr form": CSR as (rows,
 * rowptr, colval), CSC as (columns, colptr, rowval).  The reference adds the
 * terms of tr with @atomic and those of dot in per-work-item partials, so its
 * summation order is undefined; here terms are added in storage order in
 * double, and *abs_out (if not NULL) gets the sum of |term| for the error bound.
 * ------------------------------------------------------------------------- */
#define DEFINE_REDUCE(T, S)                                                         \
/* kernel_tr (src/matrix_csr/matrix_csr.jl:118-142, matrix_csc.jl:117-140):        \
 * entries whose index equals their major, duplicates included. */                 \
double ora_trace_ptr_##S(I n, const I* ptr, const I* idx, const T* val, double* abs_out) { \
    double s = 0.0, a = 0.0;                                                        \
    for (I r = 0; r < n; ++r)                                                       \
        for (I j = ptr[r] - 1; j < ptr[r + 1] - 1; ++j)                             \
            if (idx[j] == r + 1) { s += (double)val[j]; a += val[j] < 0 ? -(double)val[j] : (double)val[j]; } \
    if (abs_out) *abs_out = a;                                                      \
    return s;                                                                       \
}                                                                                   \
/* tr(::COO) (src/matrix_coo/matrix_coo.jl:131-150): rowind == colind. */           \
double ora_trace_coo_##S(I nnz, const I* r, const I* c, const T* val, double* abs_out) { \
    double s = 0.0, a = 0.0;                                                        \
    for (I j = 0; j < nnz; ++j)                                                     \
        if (r[j] == c[j]) { s += (double)val[j]; a += val[j] < 0 ? -(double)val[j] : (double)val[j]; } \
    if (abs_out) *abs_out = a;                                                      \
    return s;                                                                       \
}                                                                                   \
/* dot(x, A, y) (kernel_workgroup_dot_csr_N!/T!, src/matrix_csr/matrix_csr_kernels.jl:47-131;\
 * csc :47-131): term = (x[p]*a)*y[q] rounded in T (dot(x,a,y) = conj(x)*a*y),      \
 * (p, q) = (major, index) when major_left, else (index, major). */                \
double ora_dot3_ptr_##S(I nmaj, const I* ptr, const I* idx, const T* val, const T* x, \
                        const T* y, int major_left, double* abs_out) {              \
    double s = 0.0, a = 0.0;                                                        \
    for (I r = 0; r < nmaj; ++r)                                                    \
        for (I j = ptr[r] - 1; j < ptr[r + 1] - 1; ++j) {                           \
            I i = idx[j] - 1;                                                       \
            T t = major_left ? (x[r] * val[j]) * y[i] : (x[i] * val[j]) * y[r];     \
            s += (double)t; a += t < 0 ? -(double)t : (double)t;                   \
        }                                                                           \
    if (abs_out) *abs_out = a;                                                      \
    return s;                                                                       \
}                                                                                   \
/* dot(x, op(A::COO), y) (kernel_workgroup_dot_coo_N!/T!,                          \
 * src/matrix_coo/matrix_coo_kernels.jl:45-131). */                                 \
double ora_dot3_coo_##S(I nnz, const I* rr, const I* cc, const T* val, const T* x,  \
                        const T* y, int trans, double* abs_out) {                   \
    double s = 0.0, a = 0.0;                                                        \
    for (I j = 0; j < nnz; ++j) {                                                   \
        I r = rr[j] - 1, c = cc[j] - 1;                                             \
        T t = trans ? (x[c] * val[j]) * y[r] : (x[r] * val[j]) * y[c];              \
        s += (double)t; a += t < 0 ? -(double)t : (double)t;                       \
    }                                                                               \
    if (abs_out) *abs_out = a;                                                      \
    return s;                                                                       \
}                                                                                   \
/* kernel_check_symmetry_csr! (src/matrix_csr/matrix_csr_kernels.jl:348-407): for  \
 * each off-diagonal entry (r, c), binary search of major c (lo/hi/mid one-based,  \
 * mid = (lo+hi)÷2) for index r; found and equal, else not symmetric.  Real       \
 * values: the Hermitian test is the same. */                                       \
int ora_symmetric_ptr_##S(I n, const I* ptr, const I* idx, const T* val) {          \
    for (I r = 0; r < n; ++r)                                                       \
        for (I j = ptr[r] - 1; j < ptr[r + 1] - 1; ++j) {                           \
            I c = idx[j];                                                           \
            if (c == r + 1) continue;                                               \
            I lo = ptr[c - 1], hi = ptr[c] - 1;                                     \
            int found = 0;                                                          \
            while (lo <= hi) {                                                      \
                I mid = (lo + hi) / 2;                                              \
                I mc = idx[mid - 1];                                                \
                if (mc == r + 1) {                                                  \
                    if (val[mid - 1] != val[j]) return 0;                           \
                    found = 1;                                                      \
                    break;                                                          \
                } else if (mc < r + 1) lo = mid + 1;                                \
                else hi = mid - 1;                                                  \
            }                                                                       \
            if (!found) return 0;                                                   \
        }                                                                           \
    return 1;                                                                       \
}

DEFINE_REDUCE(double, f64)
DEFINE_REDUCE(float, f32)

#define DEFINE_ADD_DIAG(T, S)                                                       \
/* Base.:+(::CSR, ::UniformScaling) (src/matrix_csr/matrix_csr.jl:357-406):        \
 * kernel_count_diag_entries_csr! then kernel_add_uniformscaling_csr!             \
 * (matrix_csr_kernels.jl:445-509).  Returns nnz; outputs may be NULL to count. */  \
I ora_add_diag_##S(I n, const I* ptr, const I* idx, const T* val, const T* lambda_p, \
                   I* ptr_out, I* idx_out, T* val_out) {                            \
    const T lambda = *lambda_p;                                                     \
    I w = 0;                                                                        \
    if (ptr_out) ptr_out[0] = 1;                                                    \
    for (I r = 0; r < n; ++r) {                                                     \
        int written = 0;                                                            \
        for (I j = ptr[r] - 1; j < ptr[r + 1] - 1; ++j) {                           \
            I c = idx[j];                                                           \
            if (!written && c > r + 1) {                                            \
                if (idx_out) { idx_out[w] = r + 1; val_out[w] = lambda; }           \
                ++w; written = 1;                                                   \
            }                                                                       \
            if (idx_out) {                                                          \
                idx_out[w] = c;                                                     \
                val_out[w] = c == r + 1 ? val[j] + lambda : val[j];                 \
            }                                                                       \
            if (c == r + 1) written = 1;                                            \
            ++w;                                                                    \
        }                                                                           \
        if (!written) {                                                             \
            if (idx_out) { idx_out[w] = r + 1; val_out[w] = lambda; }               \
            ++w;                                                                    \
        }                                                                           \
        if (ptr_out) ptr_out[r + 1] = w + 1;                                        \
    }                                                                               \
    return w;                                                                       \
}

DEFINE_ADD_DIAG(double, f64)
DEFINE_ADD_DIAG(float, f32)
DEFINE_ADD_DIAG(zc128, c128)
DEFINE_ADD_DIAG(zc64, c64)
DEFINE_ADD_DIAG(uint64_t, i64)
DEFINE_ADD_DIAG(uint32_t, i32)

int ora_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
