/*
 * gsa_hip.h — C ABI of libgsa_hip.so, the gfx950 (MI355X) kernel layer behind
 * GenericSparseArrays.jl's sparse operators.
 *
 * Every entry point takes plain device pointers, sizes and enums and returns an
 * int status (GSA_OK = 0, negative on error; the message is in gsa_last_error()).
 * Nothing throws across the ABI.  Index arrays are ONE-BASED, exactly as the
 * reference stores them (Julia), so buffers are shared without copies.
 * Dense operands are column-major with a leading dimension, as Julia's
 * DenseVecOrMat.  Every operation is enqueued on the caller's `stream`
 * (hipStream_t, NULL = legacy default stream) and does not synchronise, except
 * the "count"/"begin" calls that return a size to the host, which synchronise
 * the stream the way the reference's `@allowscalar` read-backs of nnz totals do
 * (src/matrix_csr/matrix_csr.jl:332,543).
 *
 * Reference: GenericSparseArrays.jl v0.1.2 (Project.toml:3).  Each declaration
 * names the reference method it replaces.
 */
#ifndef GSA_HIP_H
#define GSA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define GSA_OK               0
#define GSA_ERR_INVALID     -1  /* bad argument: maps to ArgumentError           */
#define GSA_ERR_DIM         -2  /* dimension mismatch: maps to DimensionMismatch */
#define GSA_ERR_HIP         -3  /* HIP runtime error                             */
#define GSA_ERR_OOM         -4  /* device allocation failed                      */
#define GSA_ERR_UNSUPPORTED -5  /* type/op combination not provided              */

/* ---- element and index types (Tv, Ti of GenericSparseMatrix*{Tv,Ti}) -------- */
#define GSA_F64  0   /* Float64                                                 */
#define GSA_F32  1   /* Float32                                                 */
#define GSA_BF16 2   /* BFloat16 storage with one Float32 scale per row (serving)*/
#define GSA_C64  3   /* ComplexF32: (re, im) Float32 pairs, Julia's layout       */
#define GSA_C128 4   /* ComplexF64: (re, im) Float64 pairs                       */
#define GSA_VI32 5   /* Int32 values (wrapping two's-complement arithmetic)      */
#define GSA_VI64 6   /* Int64 values (wrapping two's-complement arithmetic)      */
/* Index types (Ti). */
#define GSA_I32  0   /* Int32                                                   */
#define GSA_I64  1   /* Int64                                                   */
/* Value types the reference's test suite runs every matrix and vector family on
 * (test/runtests.jl:34-68): Int32, Int64, Float32, Float64, ComplexF32,
 * ComplexF64.  Scalars (alpha, beta, lambda) and reduction results are passed
 * as host pointers to one value of the operand's value type.  Complex
 * products are Julia's (re*re - im*im, re*im + im*re), each operation rounded
 * on its own; a scalar whose imaginary part is zero multiplies both parts
 * separately, as a real scalar does in Julia. */

/* ---- operand wrappers: identity / Transpose / Adjoint ----------------------
 * Mirrors trans_adj_wrappers (src/core.jl:102-113).  Adjoint conjugates the
 * stored values (CONJA) and, on a dense operand, the dense values (CONJB);
 * for real value types it equals Transpose. */
#define GSA_OP_N 0
#define GSA_OP_T 1
#define GSA_OP_C 2

typedef struct gsa_plan_s* gsa_plan;   /* per-matrix cached analysis (row blocks) */
typedef struct gsa_ctx_s*  gsa_ctx;    /* state carried between two-phase calls   */

/* ========================== runtime ========================================= */
/* One-time device set-up (memory-pool policy).  Called implicitly by every op. */
int         gsa_init(void);
/* Message of the last failing call on this host thread ("" if none). */
const char* gsa_last_error(void);
int         gsa_version(void);   /* 10000*major + 100*minor + patch */

/* Device memory/stream helpers for hosts with no array library of their own
 * (the Julia HipVector storage type): hipMalloc/hipFree/hipMemcpyAsync. */
int gsa_malloc(void** ptr, int64_t bytes);
int gsa_free(void* ptr);
/* kind: 0 = host->device, 1 = device->host, 2 = device->device (async on stream) */
int gsa_memcpy(void* dst, const void* src, int64_t bytes, int kind, void* stream);
int gsa_memset(void* dst, int value, int64_t bytes, void* stream);
int gsa_stream_create(void** stream);
/* Waits for the stream's queued work, frees the scratch blocks cached for it,
 * then destroys it. */
int gsa_stream_destroy(void* stream);
int gsa_stream_sync(void* stream);
/* Bytes of scratch cached for `stream` (a host-side lookup; no device call). */
int64_t gsa_scratch_cached(void* stream);
/* Scratch blocks held by live HIP graphs.  A call captured into a graph takes
 * its scratch as plain device memory owned by that graph (released when the
 * graph and its executables are destroyed), so a replay runs no allocation. */
int64_t gsa_scratch_graph_blocks(void);

/* ========================== SpMV / SpMM (mul!) ============================== */
/* Row-block analysis of a CSR (or, on its column pointer, CSC) matrix, computed
 * once and reused by every mul! on the same structure: the first row starting
 * in each block of entries, for the SpMV's 576-entry windows and the SpMM's
 * 2048-entry blocks.  `ptr` has length `m+1`.  Stream-ordered device work; the
 * plan is usable on `stream` immediately.  Once built, a plan (of any kind) is
 * read-only: every product keeps its in-flight state (partial sums of rows
 * that cross window ends, binned products, tile counters) in scratch of its
 * own call, so one plan may serve concurrent calls on any streams and host
 * threads.  A plan describes the index arrays it was built on (matched by
 * address): after an in-place write to those arrays, build a new plan.  No
 * counterpart in the reference, whose mul! analyses nothing
 * (src/matrix_csr/matrix_csr.jl:145-211). */
int gsa_plan_create(void* stream, int itype, int64_t m, int64_t nnz,
                    const void* ptr, gsa_plan* plan);
int gsa_plan_destroy(gsa_plan plan);
/* What a plan holds, as GSA_PLAN_* bits: the SpMV windows of a row or column
 * pointer, their XCD launch order, a row view, row bins, a batch schedule.
 * No counterpart in the reference. */
#define GSA_PLAN_WINDOWS   1
#define GSA_PLAN_XCD_ORDER 2
#define GSA_PLAN_ROW_VIEW  4
#define GSA_PLAN_ROW_BINS  8
#define GSA_PLAN_BATCH     16
#define GSA_PLAN_TRANSPOSE_BINS 32
int gsa_plan_describe(gsa_plan plan, int* flags);
/* Add to a plan the row bins of the transposed product: transpose(A) * x
 * (k = 1) then runs the two streaming passes of the CSC SpMV (no atomics, a
 * fixed summation order) instead of the reference's atomic scatter
 * (kernel_spmatmul_csr_T! src/matrix_csr/matrix_csr_kernels.jl:24-45,
 * kernel_spmatmul_coo_T! src/matrix_coo/matrix_coo_kernels.jl:22-43).
 * format GSA_FMT_CSR: a plan from gsa_plan_create, a = rowptr, b = colval;
 * GSA_FMT_COO: a plan from gsa_plan_create_coo, a = rowind, b = colind.  A
 * (m x n) with n above 4,194,304 columns gets no bins (the scatter stays).
 * About 20 bytes per stored entry.  No counterpart in the reference. */
#define GSA_FMT_CSR 0
#define GSA_FMT_CSC 1
#define GSA_FMT_COO 2
int gsa_plan_add_transpose(void* stream, gsa_plan plan, int format, int itype, int64_t m,
                           int64_t n, int64_t nnz, const void* a, const void* b);
/* Add to a plan the gather view of a scatter-shaped product with several
 * right-hand sides (k > 1): the output rows in order, each entry's row of B and
 * its storage position.  The product then runs the gather SpMM kernel, reading
 * values through the positions, instead of the reference's atomic scatter
 * (kernel_spmatmul_csc_N! src/matrix_csc/matrix_csc_kernels.jl:1-22,
 * kernel_spmatmul_csr_T! src/matrix_csr/matrix_csr_kernels.jl:24-45,
 * kernel_spmatmul_coo_N!/T! src/matrix_coo/matrix_coo_kernels.jl:1-43); each
 * output is then summed in a fixed order.  format GSA_FMT_CSC with trans = 0
 * (a = colptr, b = rowval), GSA_FMT_CSR with trans != 0 (a = rowptr,
 * b = colval), GSA_FMT_COO with either (a = rowind, b = colind).  About 12-16
 * bytes per stored entry; synchronises the stream.  No counterpart in the
 * reference. */
#define GSA_PLAN_GATHER_VIEW 64
int gsa_plan_add_gather_view(void* stream, gsa_plan plan, int format, int trans, int itype,
                             int64_t m, int64_t n, int64_t nnz, const void* a, const void* b);
/* Give a plan the column array it will be used with (`colval`, n columns):
 * the SpMV windows lying inside rows longer than 576 entries (whose
 * ascending columns span a narrow band of x) are then launched on the XCD
 * that gathers that eighth of x, so their gathers hit its L2; with Int64
 * indices and n < 2^32 the plan also keeps a 32-bit copy of the columns
 * (4 bytes per stored entry), which the SpMV then streams instead.  Used
 * while the call's column array is this one (an in-place write to it needs a
 * new plan); results never depend on it.  Synchronises `stream`.  No
 * counterpart in the reference. */
int gsa_plan_set_columns(void* stream, gsa_plan plan, const void* colval, int64_t n);

/* C = beta*C + alpha*opA(A)*opB(B) with A in CSR (m x n, `nnz` stored entries).
 * C is size(opA(A),1) x k column-major (ldc); opB(B) is size(opA(A),2) x k:
 * opB = N reads B[i + j*ldb], opB = T reads B[j + i*ldb], opB = C reads the
 * conjugate of B[j + i*ldb].  Every value type (GSA_F64 ... GSA_VI32); alpha
 * and beta point to host values of that type (NULL: alpha = 1, beta = 0).
 * beta == 0 zero-fills C first, beta == 1 leaves it, else scales it.
 * opA = N is the row gather of kernel_spmatmul_csr_N! and opA = T/C the scatter
 * of kernel_spmatmul_csr_T! (C conjugates A's values, CONJA).  `plan` may be
 * NULL (analysis done inline).
 * Replaces LinearAlgebra.mul!(C, ::W{GenericSparseMatrixCSR}, ::W{DenseVecOrMat}, α, β)
 * — src/matrix_csr/matrix_csr.jl:145-211, kernels src/matrix_csr/matrix_csr_kernels.jl:1-45. */
int gsa_spmm_csr(void* stream, int vtype, int itype, int opA,
                 int64_t m, int64_t n, int64_t k,
                 const void* rowptr, const void* colval, const void* nzval, int64_t nnz,
                 const void* B, int64_t ldb, int opB,
                 void* C, int64_t ldc,
                 const void* alpha, const void* beta, gsa_plan plan);

/* Same for A in CSC (colptr length n+1).  opA = T/C gathers (windows over
 * colptr).  opA = N is the scatter of kernel_spmatmul_csc_N!; with a plan from
 * gsa_plan_create_csc, an SpMV (k = 1) instead runs the plan's row bins (or,
 * for more rows than the bins hold, its row view): no atomics.  Replaces mul!
 * for GenericSparseMatrixCSC —
 * src/matrix_csc/matrix_csc.jl:144-213, kernels src/matrix_csc/matrix_csc_kernels.jl:1-45. */
int gsa_spmm_csc(void* stream, int vtype, int itype, int opA,
                 int64_t m, int64_t n, int64_t k,
                 const void* colptr, const void* rowval, const void* nzval, int64_t nnz,
                 const void* B, int64_t ldb, int opB,
                 void* C, int64_t ldc,
                 const void* alpha, const void* beta, gsa_plan plan);

/* Same for A in COO (any entry order, duplicates summed): atomic scatter, with
 * equal consecutive outputs summed in the wavefront first; with a plan from
 * gsa_plan_create_coo an SpMV (opA = N, k = 1) gathers rows through the row
 * view instead.  `plan` may be NULL.
 * Replaces mul! for GenericSparseMatrixCOO — src/matrix_coo/matrix_coo.jl:155-221,
 * kernels src/matrix_coo/matrix_coo_kernels.jl:1-43. */
int gsa_spmm_coo(void* stream, int vtype, int itype, int opA,
                 int64_t m, int64_t n, int64_t k,
                 const void* rowind, const void* colind, const void* nzval, int64_t nnz,
                 const void* B, int64_t ldb, int opB,
                 void* C, int64_t ldc,
                 const void* alpha, const void* beta, gsa_plan plan);

/* Plans for the SpMV (opA = N, k = 1) of COO and CSC without atomics.
 * - Row bins (CSC, and COO whose entries are not in row order, for m up to
 *   4,194,304 rows): y is cut into bins of 2048 rows; each call streams the
 *   storage once, writing (row in bin, a_ij * x_j) pairs grouped by bin at
 *   offsets the plan computed from the rows, then one workgroup per bin adds
 *   its pairs into Float64 accumulators in LDS and writes its rows of y.  The
 *   plan holds where each entry goes, the binned rows and products (16 bytes
 *   per entry in all) and, for CSC, each entry's column (4 bytes).  The
 *   summation order is fixed, so a call's result does not change from run
 *   to run.
 * - Row view (COO already in row order, or more rows): the stored entries in
 *   row order as a CSR row pointer plus, unless the storage is row-ordered,
 *   each entry's column and storage position (a stable sort by row, once);
 *   the CSR window kernel then sums each row.
 * Values and x are read at every call, so a plan stays valid while only the
 * values change.  Synchronises `stream`.  COO: the rows of (rowind, colind);
 * CSC (m x n): the column blocks of gsa_plan_create over colptr (for opA = T)
 * plus the bins or the row view of (colptr, rowval).  No counterpart in the
 * reference (its COO/CSC mul! use @atomic, src/matrix_coo/matrix_coo_kernels.jl:1-43,
 * src/matrix_csc/matrix_csc_kernels.jl:1-22). */
int gsa_plan_create_coo(void* stream, int itype, int64_t m, int64_t n, int64_t nnz,
                        const void* rowind, const void* colind, gsa_plan* plan);
int gsa_plan_create_csc(void* stream, int itype, int64_t m, int64_t n, int64_t nnz,
                        const void* colptr, const void* rowval, gsa_plan* plan);

/* gsa_spmm_csr with opA = N and a choice of C layout: opC = GSA_OP_N keeps C
 * column-major (ldc >= m), opC = GSA_OP_T stores it row-major, C[r*ldc + j]
 * (ldc >= k; ldc = 1 when k = 1), so the k outputs of a row are written as one
 * contiguous segment instead of k lines ldc apart.  For callers that keep
 * results row-major (the serving loop); no counterpart in the reference, whose
 * C is a column-major DenseVecOrMat. */
int gsa_spmm_csr_ex(void* stream, int vtype, int itype, int64_t m, int64_t n, int64_t k,
                    const void* rowptr, const void* colval, const void* nzval, int64_t nnz,
                    const void* B, int64_t ldb, int opB, void* C, int64_t ldc, int opC,
                    double alpha, double beta, gsa_plan plan);

/* CSR SpMV y = beta*y + alpha*A*x (A m x n, `nnz` stored entries; x and y
 * contiguous) that allocates nothing on the device: the same kernels and
 * results as gsa_spmm_csr with opA = N, k = 1 and no plan, bit for bit, with
 * the per-call window table and partial sums kept in a fixed device array of
 * GSA_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead of
 * allocated scratch.  nnz above GSA_NOALLOC_MAX_NNZ is rejected with
 * GSA_ERR_INVALID before any device work.  The array is shared by every call
 * on one GPU: calls that overlap on one GPU (several streams or host threads
 * at once) race and must be serialised by the caller.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). */
#define GSA_NOALLOC_SCRATCH_BYTES (4 << 20)
#define GSA_NOALLOC_MAX_NNZ 100000000LL
int gsa_spmv_csr_noalloc(void* stream, int vtype, int itype, int64_t m, int64_t n,
                         const void* rowptr, const void* colval, const void* nzval, int64_t nnz,
                         const void* x, void* y, double alpha, double beta);

/* BF16-stored CSR SpMV/SpMM with a Float32 scale per row, Float32 B/C and
 * Float32 accumulation (TwoSum-compensated on rows reduced across lanes,
 * which include every row over 256 entries):
 * C = beta*C + alpha * diag(row_scale) * A_bf16 * opB(B), same kernels and
 * layouts as gsa_spmm_csr_ex (opC as there).  Serving precision variant; no
 * counterpart in the reference (whose Tv is a Julia number type). */
int gsa_spmm_csr_bf16(void* stream, int itype, int64_t m, int64_t n, int64_t k,
                      const void* rowptr, const void* colval, const uint16_t* nzval_bf16,
                      const float* row_scale, int64_t nnz,
                      const float* B, int64_t ldb, int opB, float* C, int64_t ldc, int opC,
                      float alpha, float beta, gsa_plan plan);

/* ========================== format conversion =============================== */
/* out[i] = in[i] + delta, i < n, from index type itype_in to itype_out (Int32
 * or Int64); an Int32 output is range-checked (GSA_ERR_INVALID if a value does
 * not fit; synchronises once); in == out is allowed when the types match.
 * Host ingest and egress: zero-based host index
 * arrays copied as they are, then made one-based on the device (and back), so
 * the host neither shifts nor widens them.  Replaces the index handling of
 * GenericSparseMatrixCSR/CSC(::SparseMatrixCSC) and SparseMatrixCSC(::Generic…),
 * src/conversions/conversions.jl:7-79, where the arrays are one-based on both
 * sides and the host transposes. */
int gsa_index_shift(void* stream, int itype_in, int itype_out, int64_t n, const void* in,
                    void* out, int64_t delta);

/* CSR -> COO (row-major entry order).  Replaces GenericSparseMatrixCOO(::CSR),
 * src/conversions/conversions.jl:264-280, kernel_csr_to_coo!
 * src/conversions/conversion_kernels.jl:20-35.  For CSC -> COO
 * (conversions.jl:153-169, kernel_csc_to_coo! conversion_kernels.jl:2-17) pass
 * (n, colptr, rowval) and swap the two index outputs. */
int gsa_csr_to_coo(void* stream, int vtype, int itype, int64_t m, int64_t nnz,
                   const void* rowptr, const void* colval, const void* nzval,
                   void* rowind, void* colind, void* nzval_out);

/* COO -> CSR: stable sort by key (row, col), gather, row pointer by
 * searchsortedfirst; duplicates are kept in input order.  Keys already in
 * order (a COO made from a CSR) skip the sort: the arrays are copied.  The
 * call synchronises the stream once, to learn which.  Replaces
 * GenericSparseMatrixCSR(::COO), src/conversions/conversions.jl:282-323
 * (keys kernel_make_csr_keys! conversion_kernels.jl:49-57).  For COO -> CSC
 * (conversions.jl:171-212) pass (n, m, colind, rowind) and read the outputs as
 * (colptr, rowval). */
int gsa_coo_to_csr(void* stream, int vtype, int itype, int64_t m, int64_t n, int64_t nnz,
                   const void* rowind, const void* colind, const void* nzval,
                   void* rowptr, void* colval, void* nzval_out);

/* CSR (m x n) -> CSC (m x n): storage transpose by a stable sort on the column
 * index.  Bit-identical to the reference's CSR -> COO -> CSC route,
 * src/conversions/conversions.jl:85-86.  CSC -> CSR (conversions.jl:104-105)
 * is the same call on (n, m, colptr, rowval). */
int gsa_csr_to_csc(void* stream, int vtype, int itype, int64_t m, int64_t n, int64_t nnz,
                   const void* rowptr, const void* colval, const void* nzval,
                   void* colptr, void* rowval, void* nzval_out);

/* ========================== sparse + sparse ================================= */
/* CSR + CSR with exact zeros dropped (merge then dropzeros).  Phase 1 writes the
 * result row pointer (m+1) and returns nnz(C) (synchronises); phase 2 fills.
 * Replaces Base.:+(::CSR, ::CSR) src/matrix_csr/matrix_csr.jl:290-355 with
 * kernels matrix_csr_kernels.jl:148-248 and dropzeros matrix_csr.jl:723-745.
 * CSC + CSC (src/matrix_csc/matrix_csc.jl:292-357) is the same call on the
 * column pointers.  nnzA and nnzB are the operands' stored entries (the lengths
 * of colA/valA and colB/valB, nnz(A) and nnz(B) in Julia); every read is
 * clamped into them. */
int gsa_csr_add_count(void* stream, int vtype, int itype, int64_t m,
                      int64_t nnzA, const void* rowptrA, const void* colA, const void* valA,
                      int64_t nnzB, const void* rowptrB, const void* colB, const void* valB,
                      void* rowptrC, int64_t* nnzC);
int gsa_csr_add_fill(void* stream, int vtype, int itype, int64_t m,
                     int64_t nnzA, const void* rowptrA, const void* colA, const void* valA,
                     int64_t nnzB, const void* rowptrB, const void* colB, const void* valB,
                     const void* rowptrC, void* colC, void* valC);

/* COO + COO: concatenate, stable sort by column-major key, sum duplicate runs in
 * order, drop exact zeros.  Output is column-major sorted COO.  begin returns
 * nnzC and a context; the caller allocates the outputs and calls finish, which
 * releases the context.  The operands must stay unchanged until finish (with
 * both stored in column-major order, finish reads them again).
 * Replaces Base.:+(::COO, ::COO) src/matrix_coo/matrix_coo.jl:300-384
 * (kernel_mark_unique_coo!/kernel_compact_coo! matrix_coo_kernels.jl:260-310). */
int gsa_coo_add_begin(void* stream, int vtype, int itype, int64_t m, int64_t n,
                      int64_t nnzA, const void* rowA, const void* colA, const void* valA,
                      int64_t nnzB, const void* rowB, const void* colB, const void* valB,
                      gsa_ctx* ctx, int64_t* nnzC);
int gsa_coo_add_finish(gsa_ctx ctx, void* rowC, void* colC, void* valC);

/* ========================== dropzeros ====================================== */
/* Replaces SparseArrays.dropzeros(::CSR) src/matrix_csr/matrix_csr.jl:650-683,723-745
 * (kernels matrix_csr_kernels.jl:410-443); CSC is the same call on colptr.
 * Both phases read nnz from rowptr[m] (synchronises).  Phase 1 writes the result
 * row pointer (m+1) and returns the kept count; phase 2 writes each kept entry
 * at the number of kept entries before it (storage order), so rowptr_out, the
 * pointer phase 1 wrote, is taken for the reference's signature and not read. */
int gsa_csr_dropzeros_count(void* stream, int vtype, int itype, int64_t m,
                            const void* rowptr, const void* nzval,
                            void* rowptr_out, int64_t* nnz_out);
int gsa_csr_dropzeros_fill(void* stream, int vtype, int itype, int64_t m,
                           const void* rowptr, const void* colval, const void* nzval,
                           const void* rowptr_out, void* colval_out, void* nzval_out);
/* Replaces SparseArrays.dropzeros(::COO) src/matrix_coo/matrix_coo.jl:824-854,909-935. */
int gsa_coo_dropzeros_count(void* stream, int vtype, int64_t nnz, const void* nzval,
                            int64_t* nnz_out);
int gsa_coo_dropzeros_fill(void* stream, int vtype, int itype, int64_t nnz,
                           const void* rowind, const void* colind, const void* nzval,
                           void* rowind_out, void* colind_out, void* nzval_out);

/* ========================== kron =========================================== */
/* kron(COO, COO): entry (iA, iB) -> ((iA-1)*mB+iB, (jA-1)*nB+jB, vA*vB), A-major /
 * B-minor entry order.  Replaces LinearAlgebra.kron(::COO, ::COO)
 * src/matrix_coo/matrix_coo.jl:641-682, kernel_kron_coo! matrix_coo_kernels.jl:144-183. */
int gsa_kron_coo(void* stream, int vtype, int itype,
                 int64_t mB, int64_t nB,
                 int64_t nnzA, const void* rowA, const void* colA, const void* valA,
                 int64_t nnzB, const void* rowB, const void* colB, const void* valB,
                 void* rowC, void* colC, void* valC);
/* kron(CSR, CSR) -> CSR in canonical order, built directly (no sort) when both
 * inputs have strictly ascending columns per row, else through kron(COO,COO)
 * and the COO -> CSR sort, as the reference does (src/matrix_csr/matrix_csr.jl:442-458).
 * rowptrC has mA*mB+1 entries; nnz(C) = nnzA*nnzB.  CSC is the same call on the
 * column pointers with (nA, nB) in place of (mA, mB). */
int gsa_kron_csr(void* stream, int vtype, int itype,
                 int64_t mA, int64_t nA, int64_t mB, int64_t nB,
                 const void* rowptrA, const void* colA, const void* valA, int64_t nnzA,
                 const void* rowptrB, const void* colB, const void* valB, int64_t nnzB,
                 void* rowptrC, void* colC, void* valC);

/* ========================== SpGEMM ========================================= */
/* C = A*B, A (m x kdim) and B (kdim x n) in CSR.  Row-wise Gustavson semantics of
 * kernel_count_nnz_spgemm_csr!/kernel_spgemm_csr! (src/matrix_csr/matrix_csr_kernels.jl:254-343):
 * per output entry the products are summed in (A-row order, B-row order) from
 * +0.0, columns ascending, cancellation zeros kept.  Two phases: begin counts
 * the distinct columns of every output row (symbolic) and returns nnz(C)
 * (synchronises); finish writes the row pointer (m+1), columns and values
 * (numeric) and releases ctx.  Each output row is sorted by column inside a
 * wavefront or workgroup; a product with a row over 2048 partial products
 * uses a global sort instead.  A and B must stay unchanged until finish.
 * Replaces Base.:*(::CSR, ::CSR) src/matrix_csr/matrix_csr.jl:498-572.
 * CSC * CSC (src/matrix_csc/matrix_csc.jl:500-574) is this call with the
 * operands swapped (B's column arrays first). */
int gsa_spgemm_csr_begin(void* stream, int vtype, int itype,
                         int64_t m, int64_t kdim, int64_t n,
                         const void* rowptrA, const void* colA, const void* valA, int64_t nnzA,
                         const void* rowptrB, const void* colB, const void* valB, int64_t nnzB,
                         gsa_ctx* ctx, int64_t* nnzC);
int gsa_spgemm_csr_finish(gsa_ctx ctx, void* rowptrC, void* colC, void* valC);
/* Release a ctx without finishing (error paths). */
int gsa_ctx_destroy(gsa_ctx ctx);

/* ========================== reductions over stored entries ================= */
/* "ptr form": CSR as (n = m rows, ptr = rowptr, idx = colval); CSC as
 * (n = columns, ptr = colptr, idx = rowval).  The sums are accumulated in a
 * fixed order (the same result on every call) and returned on the host, in
 * `out`: a Float64 for real value types, a ComplexF64 (re, im) for complex ones,
 * an Int64 for integer ones (wrapping; its low 32 bits are the Int32 sum); the
 * call synchronises the stream.  `plan` (may be NULL) is a plan built
 * on this `ptr` (gsa_plan_create, or gsa_plan_create_csc for CSC): its row
 * blocks are reused, else they are rebuilt for the call.  Majors of up to 64
 * entries are read by one lane, longer ones by a whole workgroup.
 *
 * tr: the sum of the entries whose index equals their row (column), duplicates
 * included.  The caller checks squareness.  Replaces LinearAlgebra.tr(::CSR)
 * src/matrix_csr/matrix_csr.jl:118-142 and tr(::CSC) src/matrix_csc/matrix_csc.jl:117-140. */
int gsa_trace_ptr(void* stream, int vtype, int itype, int64_t n, int64_t nnz,
                  const void* ptr, const void* idx, const void* val, gsa_plan plan, void* out);
/* tr(::COO): entries with rowind == colind.  src/matrix_coo/matrix_coo.jl:131-150. */
int gsa_trace_coo(void* stream, int vtype, int itype, int64_t nnz, const void* rowind,
                  const void* colind, const void* val, void* out);
/* sum(A) = sum(nonzeros(A)) for every format; also sum(::GenericSparseVector).
 * src/core.jl:24, src/vector/vector.jl. */
int gsa_sum_values(void* stream, int vtype, int64_t nnz, const void* val, void* out);
/* dot(x, A, y) = sum over stored entries of (conj(x[p])*a)*y[q], each term
 * rounded in the value type, a conjugated when conj_a != 0 (Adjoint, CONJA):
 * (p, q) = (major, index) when major_left != 0, else (index, major).  CSR:
 * major_left = 1 for A, 0 for transpose(A)/adjoint(A); CSC the opposite.  x and
 * y have the lengths the caller checked (size(opA(A))).
 * Replaces LinearAlgebra.dot(x, ::CSR, y) src/matrix_csr/matrix_csr.jl:214-277
 * (kernels matrix_csr_kernels.jl:47-131) and dot(x, ::CSC, y)
 * src/matrix_csc/matrix_csc.jl:216-279 (kernels matrix_csc_kernels.jl:47-131). */
int gsa_dot3_ptr(void* stream, int vtype, int itype, int64_t nmaj, int64_t nidx, int64_t nnz,
                 const void* ptr, const void* idx, const void* val, const void* x, const void* y,
                 int major_left, int conj_a, gsa_plan plan, void* out);
/* dot(x, op(A), y) for COO (m x n): (conj(x[row])*a)*y[col], or
 * (conj(x[col])*a)*y[row] when trans != 0; a conjugated when conj_a != 0.
 * src/matrix_coo/matrix_coo.jl:224-285 (kernels matrix_coo_kernels.jl:45-131). */
int gsa_dot3_coo(void* stream, int vtype, int itype, int64_t m, int64_t n, int64_t nnz,
                 const void* rowind, const void* colind, const void* val, const void* x,
                 const void* y, int trans, int conj_a, void* out);
/* issymmetric / ishermitian for a square matrix in ptr form: *out = 1 when
 * every off-diagonal entry (r, c) finds, by binary search of major c, an entry
 * at index r with an equal value (hermitian != 0: equal to its conjugate, and
 * every diagonal value equal to its own conjugate); the indices of each major
 * sorted, as the reference's search assumes.
 * Replaces issymmetric/ishermitian(::CSR) src/matrix_csr/matrix_csr.jl:611-647
 * (kernel_check_symmetry_csr! matrix_csr_kernels.jl:348-407) and (::CSC)
 * src/matrix_csc/matrix_csc.jl:616-652.  COO converts to CSC first
 * (src/matrix_coo/matrix_coo.jl:811-821). */
int gsa_symmetric_ptr(void* stream, int vtype, int itype, int64_t n, int64_t nnz,
                      const void* ptr, const void* idx, const void* val, int hermitian,
                      gsa_plan plan, int* out);

/* ========================== A + λI ========================================= */
/* CSR / CSC (ptr form, n x n) + UniformScaling λ != 0, two phases.  count:
 * ptr_out (n+1) gets the new pointer, one more entry for each major with no
 * diagonal entry; *nnz_out its total (synchronises).  fill: the entries in
 * storage order with λ inserted before the first index past the diagonal (or
 * appended), and λ added to each stored diagonal entry; `plan` as for tr.
 * The indices of each major ascend (the reference's merge assumes it).  Replaces
 * Base.:+(::CSR, ::UniformScaling) src/matrix_csr/matrix_csr.jl:357-406
 * (kernels matrix_csr_kernels.jl:445-509) and (::CSC)
 * src/matrix_csc/matrix_csc.jl:359-408.  COO + λI is gsa_coo_add_begin with
 * the n diagonal entries as the second operand (src/matrix_coo/matrix_coo.jl:386-457). */
int gsa_add_diag_count(void* stream, int itype, int64_t n, int64_t nnz, const void* ptr,
                       const void* idx, gsa_plan plan, void* ptr_out, int64_t* nnz_out);
int gsa_add_diag_fill(void* stream, int vtype, int itype, int64_t n, int64_t nnz,
                      const void* ptr, const void* idx, const void* val, const void* lambda,
                      gsa_plan plan, const void* ptr_out, void* idx_out, void* val_out);

/* The n diagonal entries (i, i, d_i), i = 1..n, of an n x n COO matrix: d_i =
 * dvals[i] (a Diagonal's values, in the value type) or, when dvals is null,
 * *lambda (a host value of the value type).  The second operand of COO + λI (kernel_fill_diagonal_coo!
 * src/matrix_coo/matrix_coo_kernels.jl:312-325) and of kron with a Diagonal,
 * which is kron(COO, COO) with this operand (kernel_kron_diag_coo! /
 * kernel_kron_coo_diag! matrix_coo_kernels.jl:185-258 index the same entries;
 * src/matrix_coo/matrix_coo.jl:704-809). */
int gsa_diag_coo(void* stream, int vtype, int itype, int64_t n, const void* dvals,
                 const void* lambda, void* row, void* col, void* val);

/* Sparse + dense: C[i, j] += a_ij for every stored entry, into the column-major
 * m x n dense C (leading dimension ldc) that already holds a copy of the dense
 * operand.  format GSA_FMT_CSR: a = rowptr (m+1), b = colval; GSA_FMT_CSC: a =
 * colptr (n+1), b = rowval (also transpose/adjoint(CSR) with the CSR arrays);
 * GSA_FMT_COO: a = rowind, b = colind.  conj != 0 adds conj(a_ij) (an
 * adjoint).  Without duplicate indices every element gets at most one add, so
 * C = B + a_ij rounded once in any order; duplicates add in hardware order.
 * Replaces Base.:+(::AbstractGenericSparseMatrix, ::DenseMatrix) src/core.jl:67-87
 * (copy(B), then _add_sparse_to_dense!: kernel_add_sparse_to_dense_csr!
 * src/matrix_csr/matrix_csr_kernels.jl:134-145, _csc! matrix_csc_kernels.jl:134-145,
 * _coo! matrix_coo_kernels.jl:132-142). */
int gsa_add_to_dense(void* stream, int vtype, int itype, int format, int conj, int64_t m,
                     int64_t n, int64_t nnz, const void* a, const void* b, const void* val,
                     void* C, int64_t ldc);

/* ========================== GenericSparseVector ============================ */
/* dot(x, y) for x a sparse vector (length n, nnz stored entries, one-based
 * nzind) and y a dense device vector: sum of conj(nzval[i]) * y[nzind[i]] in a
 * fixed order, returned in `out` as the reductions above return theirs; no
 * atomics.  Synchronises.  Replaces LinearAlgebra.dot(::GenericSparseVector,
 * ::DenseVector) src/vector/vector.jl:102-139 (dot(y, x) is its conjugate,
 * :140-145). */
int gsa_spvec_dot(void* stream, int vtype, int itype, int64_t n, int64_t nnz,
                  const void* nzind, const void* nzval, const void* y, void* out);
/* norm(v, p) of the stored values (nnz of them): p = 2, 1, Inf, -Inf, 0 (the
 * count of nonzeros) or any other real p, as the largest |v| times the p-norm
 * of v / max|v| (no overflow or underflow of the powers); complex values by
 * their modulus.  *out is a Float64.  Synchronises.  Replaces
 * LinearAlgebra.norm(::GenericSparseVector, p) src/vector/vector.jl:91-93
 * (norm(nonzeros(V), p)). */
int gsa_norm_values(void* stream, int vtype, int64_t nnz, const void* val, double p,
                    double* out);

/* ========================== batched uneven CSR SpMM ======================== */
/* Many independent CSR matrices in ONE persistent launch.  Each instance i has
 * its own (m_i, n_i, k_i) and pointers; the descriptor table lives in device
 * memory (gsa_batch_desc, 72 bytes each).  Row work is split into tiles by a
 * per-batch schedule built on the device (gsa_batch_plan_create), so the
 * launch is graph-capturable (no host work, no allocation at run time).
 * No single counterpart in the reference: it replaces a loop of
 * mul!(C_i, A_i, B_i) calls (src/matrix_csr/matrix_csr.jl:145-211). */
typedef struct {
    const void* rowptr;   /* m+1 one-based                                  */
    const void* colval;   /* nnz one-based                                  */
    const void* nzval;    /* nnz values (vtype)                             */
    const void* row_scale;/* BF16 only: m scales, else NULL                 */
    const void* B;        /* n x k, column-major (ld = ldb) unless flagged   */
    void*       C;        /* m x k, column-major (ld = ldc) unless flagged   */
    int32_t m, n, k, ldb, ldc;
    int32_t flags;        /* GSA_BATCH_B_ROWMAJOR | GSA_BATCH_C_ROWMAJOR      */
} gsa_batch_desc;
#define GSA_BATCH_B_ROWMAJOR 1  /* B[i,j] at B[i*ldb + j] (Julia: transpose(Bt)) */
#define GSA_BATCH_C_ROWMAJOR 2  /* C[i,j] at C[i*ldc + j]                        */

/* Tiles: instance i is cut into row tiles of max(16, min(1024,
 * rows_per_tile / lanes_i)) rows, lanes_i = next_pow2(k_i / w_i), w_i the
 * elements one lane loads per stored entry (up to 16 bytes of a row-major B,
 * else 1).  The instances are dealt to 8 tile lists, one per XCD, balanced by
 * their stored entries and k (synchronises once).  vtype GSA_BF16 needs
 * row_scale. */
int gsa_batch_plan_create(void* stream, int vtype, int itype, int64_t count,
                          const gsa_batch_desc* host_desc, int rows_per_tile,
                          gsa_plan* plan);
/* C_i = alpha*A_i*B_i + beta*C_i for all i: a counter-reset kernel and one
 * persistent launch on `stream`; allocates nothing, so it can be captured. */
int gsa_batch_spmm(void* stream, gsa_plan plan, float alpha, float beta);
/* The same launch on caller-built device tables: d_tiles holds 3 int32 per
 * tile (instance, first row, end row), d_counter one int of scratch.  For
 * serving loops that regroup instances per launch.  grid <= 0: 8 per CU. */
int gsa_batch_spmm_tables(void* stream, int vtype, int itype, const gsa_batch_desc* d_desc,
                          const int32_t* d_tiles, int64_t ntiles, int* d_counter,
                          float alpha, float beta, int grid);
/* Capture gsa_batch_spmm into an executable HIP graph; replay with
 * gsa_graph_launch, release with gsa_graph_destroy. */
int gsa_batch_graph_create(void* stream, gsa_plan plan, float alpha, float beta,
                           void** graph_exec);
int gsa_graph_launch(void* graph_exec, void* stream);
int gsa_graph_destroy(void* graph_exec);

/* BF16 storage of CSR values with one power-of-two Float32 scale per row
 * (scale_i >= max_j |a_ij|, so a_ij / scale_i is exact before rounding to
 * BF16).  Input Float64 or Float32.  Serving precision variant; no
 * counterpart in the reference. */
int gsa_csr_to_bf16(void* stream, int vtype, int itype, int64_t m, const void* rowptr,
                    const void* nzval, uint16_t* nzval_bf16, float* row_scale);

#ifdef __cplusplus
}
#endif
#endif /* GSA_HIP_H */
