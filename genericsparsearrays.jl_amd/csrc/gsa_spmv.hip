// gsa_spmv.hip — mul! for CSR / CSC / COO on gfx950.
//
// CSR gather (CSR-N, CSC-T): "row-block streaming".  The nnz range is cut into
// windows of SPMV_NB entries; block b owns the rows whose first entry lies in
// window b (the plan: one lower bound per window, built in O(m) on the device).
// A 256-lane workgroup
//   1. streams its rows' colval/nzval with fully coalesced loads and stores the
//      products nzval[j]*x[colval[j]] in LDS,
//   2. sums each short row (<= LONG_T entries) on one lane, left to right — the
//      summation order of kernel_spmatmul_csr_N! (src/matrix_csr/matrix_csr_kernels.jl:15-20),
//      so these rows are bit-identical to the reference,
//   3. reduces longer staged rows with one wavefront each (strided, then a
//      butterfly), and
//   4. reduces its last row, if it spills past the LDS capacity, with the whole
//      workgroup straight from HBM.
// Every output element is written by exactly one lane: no atomics, and the
// result does not depend on scheduling.
//
// SpMM (k > 1) uses the same row blocks; the dense operand is read as a
// row-major n x k array (transposed once into scratch when it is column-major),
// so the k values a stored entry needs are one contiguous segment, and each
// (row, k) output is summed on one lane in storage order.
//
// Scatter-shaped products (CSR-T, CSC-N, COO) add into C with hardware float
// atomics, as the reference's @atomic kernels do (matrix_csr_kernels.jl:24-45,
// src/matrix_csc/matrix_csc_kernels.jl:1-22, src/matrix_coo/matrix_coo_kernels.jl:1-43).
#include "gsa_common.h"

namespace gsa {

constexpr int WG = 256;
constexpr int SPMV_CAP = SPMV_NB + 256;  // LDS staging capacity (entries)
constexpr int LONG_T = 48;               // rows longer than this: one wavefront each
constexpr int MAX_LONG = SPMV_CAP / (LONG_T + 1) + 1;

__device__ __forceinline__ int64_t clamp64(int64_t v, int64_t lo, int64_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

// out = (β==0 ? 0 : β==1 ? y : y*β) + acc*α — the two roundings of
// `_rmul_or_fill!(C, β)` followed by `C[row,k] += tmp*α` (matrix_csr.jl:192,
// matrix_csr_kernels.jl:21).
template <typename T>
__device__ __forceinline__ void store_out(T* __restrict__ y, int64_t idx, T acc, T alpha,
                                          T beta, int mode) {
    T t = acc * alpha;
    T base;
    if (mode == 0) base = T(0);
    else if (mode == 1) base = y[idx];
    else base = y[idx] * beta;
    y[idx] = base + t;
}

// ------------------------------------------------------------ plan kernel --
template <typename I>
__global__ void k_csr_plan(int64_t m, int64_t nnz, const I* __restrict__ ptr,
                           int64_t* __restrict__ blk_row, int64_t nblocks) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r > m) return;
    int64_t cur = (r == m) ? nnz : clamp64((int64_t)ptr[r] - 1, 0, nnz);
    int64_t bstart = 0;
    if (r > 0) bstart = clamp64((int64_t)ptr[r - 1] - 1, 0, nnz) / SPMV_NB + 1;
    int64_t bend = cur / SPMV_NB;
    for (int64_t b = bstart; b <= bend && b < nblocks; ++b) blk_row[b] = r;
    if (r == m) blk_row[nblocks] = m;
}

int build_csr_plan(hipStream_t s, int itype, int64_t m, int64_t nnz, const void* ptr,
                   int64_t* blk_row, int64_t nblocks) {
    GSA_HIP(hipMemsetAsync(blk_row, 0, sizeof(int64_t) * (nblocks + 1), s));
    return dispatch_i(itype, [&](auto ti) {
        using I = decltype(ti);
        unsigned g = (unsigned)cdiv(m + 1, 256);
        hipLaunchKernelGGL(k_csr_plan<I>, dim3(g), dim3(256), 0, s, m, nnz,
                           (const I*)ptr, blk_row, nblocks);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    });
}

// ------------------------------------------------------------ SpMV kernel --
template <typename T, typename I>
__global__ __launch_bounds__(WG) void k_csr_spmv(
    int64_t m, int64_t n, int64_t nnz, const I* __restrict__ rowptr,
    const I* __restrict__ colval, const T* __restrict__ nzval, const T* __restrict__ x,
    T* __restrict__ y, T alpha, T beta, int mode, const int64_t* __restrict__ blk_row) {
    __shared__ T prod[SPMV_CAP];
    __shared__ int64_t long_rows[MAX_LONG];
    __shared__ int nlong;
    __shared__ T red[WG / 64];

    const int tid = threadIdx.x;
    const int64_t b = blockIdx.x;
    const int64_t r0 = clamp64(blk_row[b], 0, m);
    const int64_t r1 = clamp64(blk_row[b + 1], r0, m);
    if (r0 >= r1) return;

    const int64_t P0 = clamp64((int64_t)rowptr[r0] - 1, 0, nnz);
    const int64_t Pend = clamp64((int64_t)rowptr[r1] - 1, P0, nnz);
    int64_t rs_end, Ps;
    if (Pend - P0 <= SPMV_CAP) {
        rs_end = r1;
        Ps = Pend;
    } else {
        rs_end = r1 - 1;
        Ps = clamp64((int64_t)rowptr[r1 - 1] - 1, P0, P0 + SPMV_CAP);
    }
    const int cnt = (int)(Ps - P0);
    if (tid == 0) nlong = 0;

    // 1. stage products (coalesced index/value loads, 8 in flight per lane)
    constexpr int U = 8;
    for (int base = 0; base < cnt; base += WG * U) {
        I ci[U];
        T vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int i = base + u * WG + tid;
            if (i < cnt) { ci[u] = colval[P0 + i]; vv[u] = nzval[P0 + i]; }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int i = base + u * WG + tid;
            if (i < cnt) {
                int64_t c = clamp64((int64_t)ci[u] - 1, 0, n - 1);
                prod[i] = vv[u] * x[c];
            }
        }
    }
    __syncthreads();

    // 2. one lane per short row, left-to-right sum
    for (int64_t r = r0 + tid; r < rs_end; r += WG) {
        int s = (int)(clamp64((int64_t)rowptr[r] - 1, P0, Ps) - P0);
        int e = (int)(clamp64((int64_t)rowptr[r + 1] - 1, P0, Ps) - P0);
        if (e - s <= LONG_T) {
            T acc = T(0);
            for (int j = s; j < e; ++j) acc = acc + prod[j];
            store_out(y, r, acc, alpha, beta, mode);
        } else {
            int slot = atomicAdd(&nlong, 1);
            if (slot < MAX_LONG) long_rows[slot] = r;
        }
    }
    __syncthreads();

    // 3. one wavefront per long staged row
    const int wave = tid >> 6, lane = tid & 63;
    const int nl = nlong < MAX_LONG ? nlong : MAX_LONG;
    for (int q = wave; q < nl; q += WG / 64) {
        int64_t r = long_rows[q];
        int s = (int)(clamp64((int64_t)rowptr[r] - 1, P0, Ps) - P0);
        int e = (int)(clamp64((int64_t)rowptr[r + 1] - 1, P0, Ps) - P0);
        T acc = T(0);
        for (int j = s + lane; j < e; j += 64) acc = acc + prod[j];
        acc = wave_sum(acc);
        if (lane == 0) store_out(y, r, acc, alpha, beta, mode);
    }

    // 4. spill row: whole workgroup from HBM
    if (rs_end < r1) {
        const int64_t r = r1 - 1;
        T acc = T(0);
        int64_t j = Ps + tid;
        for (; j + 3 * WG < Pend; j += 4 * WG) {
            I c0 = colval[j], c1 = colval[j + WG], c2 = colval[j + 2 * WG], c3 = colval[j + 3 * WG];
            T v0 = nzval[j], v1 = nzval[j + WG], v2 = nzval[j + 2 * WG], v3 = nzval[j + 3 * WG];
            T p0 = v0 * x[clamp64((int64_t)c0 - 1, 0, n - 1)];
            T p1 = v1 * x[clamp64((int64_t)c1 - 1, 0, n - 1)];
            T p2 = v2 * x[clamp64((int64_t)c2 - 1, 0, n - 1)];
            T p3 = v3 * x[clamp64((int64_t)c3 - 1, 0, n - 1)];
            acc = acc + p0; acc = acc + p1; acc = acc + p2; acc = acc + p3;
        }
        for (; j < Pend; j += WG) acc = acc + nzval[j] * x[clamp64((int64_t)colval[j] - 1, 0, n - 1)];
        acc = wave_sum(acc);
        if (lane == 0) red[wave] = acc;
        __syncthreads();
        if (tid == 0) {
            T s = red[0];
            for (int w = 1; w < WG / 64; ++w) s = s + red[w];
            store_out(y, r, s, alpha, beta, mode);
        }
    }
}

// ------------------------------------------------------------ SpMM kernel --
// Dense operand read as BT[col*ldbt + kk] (row-major).  LPR lanes per row.
template <typename T, typename I, int LPR>
__global__ __launch_bounds__(WG) void k_csr_spmm(
    int64_t m, int64_t n, int64_t k, int64_t nnz, const I* __restrict__ rowptr,
    const I* __restrict__ colval, const T* __restrict__ nzval, const T* __restrict__ BT,
    int64_t ldbt, T* __restrict__ C, int64_t ldc, T alpha, T beta, int mode,
    const int64_t* __restrict__ blk_row) {
    __shared__ I s_col[SPMV_CAP];
    __shared__ T s_val[SPMV_CAP];
    __shared__ T red[WG];

    const int tid = threadIdx.x;
    const int64_t b = blockIdx.x;
    const int64_t r0 = clamp64(blk_row[b], 0, m);
    const int64_t r1 = clamp64(blk_row[b + 1], r0, m);
    if (r0 >= r1) return;

    const int64_t P0 = clamp64((int64_t)rowptr[r0] - 1, 0, nnz);
    const int64_t Pend = clamp64((int64_t)rowptr[r1] - 1, P0, nnz);
    int64_t rs_end, Ps;
    if (Pend - P0 <= SPMV_CAP) { rs_end = r1; Ps = Pend; }
    else { rs_end = r1 - 1; Ps = clamp64((int64_t)rowptr[r1 - 1] - 1, P0, P0 + SPMV_CAP); }
    const int cnt = (int)(Ps - P0);

    for (int i = tid; i < cnt; i += WG) {
        s_col[i] = (I)clamp64((int64_t)colval[P0 + i] - 1, 0, n - 1);
        s_val[i] = nzval[P0 + i];
    }
    __syncthreads();

    constexpr int GROUPS = WG / LPR;
    const int g = tid / LPR, lane = tid % LPR;
    for (int64_t r = r0 + g; r < rs_end; r += GROUPS) {
        int s = (int)(clamp64((int64_t)rowptr[r] - 1, P0, Ps) - P0);
        int e = (int)(clamp64((int64_t)rowptr[r + 1] - 1, P0, Ps) - P0);
        for (int64_t kk = lane; kk < k; kk += LPR) {
            T acc = T(0);
            int j = s;
            for (; j + 1 < e; j += 2) {
                T b0 = BT[(int64_t)s_col[j] * ldbt + kk];
                T b1 = BT[(int64_t)s_col[j + 1] * ldbt + kk];
                T p0 = s_val[j] * b0;
                T p1 = s_val[j + 1] * b1;
                acc = acc + p0;
                acc = acc + p1;
            }
            if (j < e) acc = acc + s_val[j] * BT[(int64_t)s_col[j] * ldbt + kk];
            store_out(C, r + kk * ldc, acc, alpha, beta, mode);
        }
    }

    if (rs_end < r1) {  // spill row: groups take strided entries, LDS combine
        const int64_t r = r1 - 1;
        for (int64_t k0 = 0; k0 < k; k0 += LPR) {
            const int64_t kk = k0 + lane;
            T acc = T(0);
            if (kk < k) {
                for (int64_t j = Ps + g; j < Pend; j += GROUPS) {
                    int64_t c = clamp64((int64_t)colval[j] - 1, 0, n - 1);
                    acc = acc + nzval[j] * BT[c * ldbt + kk];
                }
            }
            __syncthreads();
            red[tid] = acc;
            __syncthreads();
            if (g == 0 && kk < k) {
                T s = red[lane];
                for (int q = 1; q < GROUPS; ++q) s = s + red[q * LPR + lane];
                store_out(C, r + kk * ldc, s, alpha, beta, mode);
            }
        }
    }
}

// ------------------------------------------------------ helpers / scatter --
template <typename T>
__global__ void k_beta(int64_t rows, int64_t k, T* __restrict__ C, int64_t ldc, T beta,
                       int mode) {
    int64_t total = rows * k;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = t % rows, j = t / rows;
        T* p = C + i + j * ldc;
        *p = (mode == 0) ? T(0) : (*p) * beta;
    }
}

// B (rows x cols, column-major, ld) -> BT row-major (ld = cols).  64-row tiles.
template <typename T>
__global__ __launch_bounds__(256) void k_to_rowmajor(int64_t rows, int64_t cols,
                                                     const T* __restrict__ B, int64_t ldb,
                                                     T* __restrict__ BT) {
    __shared__ T tile[64][65];
    const int64_t r0 = blockIdx.x * 64;
    for (int64_t c0 = 0; c0 < cols; c0 += 64) {
        const int cw = (int)((cols - c0) < 64 ? (cols - c0) : 64);
        for (int t = threadIdx.x; t < 64 * cw; t += 256) {
            int rr = t & 63, cc = t >> 6;
            if (r0 + rr < rows) tile[rr][cc] = B[(r0 + rr) + (c0 + cc) * ldb];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < 64 * cw; t += 256) {
            int rr = t / cw, cc = t % cw;
            if (r0 + rr < rows) BT[(r0 + rr) * cols + c0 + cc] = tile[rr][cc];
        }
        __syncthreads();
    }
}

template <typename T>
__device__ __forceinline__ void atomic_add(T* p, T v) {
    unsafeAtomicAdd(p, v);
}

// Scatter: for every stored (row, col, v) of the CSR arrays,
// C[col, kk] += v * (B[row, kk] * α)  (kernel_spmatmul_csr_T!).
template <typename T, typename I>
__global__ __launch_bounds__(WG) void k_csr_scatter(
    int64_t m, int64_t n, int64_t k, int64_t nnz, const I* __restrict__ rowptr,
    const I* __restrict__ colval, const T* __restrict__ nzval, const T* __restrict__ B,
    int64_t ldb, int opB, T* __restrict__ C, int64_t ldc, T alpha,
    const int64_t* __restrict__ blk_row) {
    const int tid = threadIdx.x;
    const int64_t b = blockIdx.x;
    const int64_t r0 = clamp64(blk_row[b], 0, m);
    const int64_t r1 = clamp64(blk_row[b + 1], r0, m);
    if (r0 >= r1) return;
    // rows [r0, r1-1): one lane per row; last row: whole workgroup
    for (int64_t kk = 0; kk < k; ++kk) {
        for (int64_t r = r0 + tid; r < r1 - 1; r += WG) {
            T xb = opB ? B[kk + r * ldb] : B[r + kk * ldb];
            T axj = xb * alpha;
            int64_t s = clamp64((int64_t)rowptr[r] - 1, 0, nnz);
            int64_t e = clamp64((int64_t)rowptr[r + 1] - 1, s, nnz);
            for (int64_t j = s; j < e; ++j) {
                int64_t c = clamp64((int64_t)colval[j] - 1, 0, n - 1);
                atomic_add(&C[c + kk * ldc], nzval[j] * axj);
            }
        }
        const int64_t r = r1 - 1;
        T xb = opB ? B[kk + r * ldb] : B[r + kk * ldb];
        T axj = xb * alpha;
        int64_t s = clamp64((int64_t)rowptr[r] - 1, 0, nnz);
        int64_t e = clamp64((int64_t)rowptr[r + 1] - 1, s, nnz);
        for (int64_t j = s + tid; j < e; j += WG) {
            int64_t c = clamp64((int64_t)colval[j] - 1, 0, n - 1);
            atomic_add(&C[c + kk * ldc], nzval[j] * axj);
        }
    }
}

// COO scatter: C[r, kk] += v * (B[c, kk] * α); trans swaps the roles of r and c.
template <typename T, typename I>
__global__ void k_coo_scatter(int64_t rows_out, int64_t cols_in, int64_t k, int64_t nnz,
                              const I* __restrict__ rowind, const I* __restrict__ colind,
                              const T* __restrict__ nzval, int trans,
                              const T* __restrict__ B, int64_t ldb, int opB,
                              T* __restrict__ C, int64_t ldc, T alpha) {
    int64_t total = nnz * k;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = t % nnz, kk = t / nnz;
        int64_t r = (int64_t)(trans ? colind[i] : rowind[i]) - 1;
        int64_t c = (int64_t)(trans ? rowind[i] : colind[i]) - 1;
        r = clamp64(r, 0, rows_out - 1);
        c = clamp64(c, 0, cols_in - 1);
        T xb = opB ? B[kk + c * ldb] : B[c + kk * ldb];
        T axj = xb * alpha;
        atomic_add(&C[r + kk * ldc], nzval[i] * axj);
    }
}

// ---------------------------------------------------------------- drivers --
static int beta_mode(double beta) { return beta == 0.0 ? 0 : (beta == 1.0 ? 1 : 2); }

template <typename T>
static int apply_beta(hipStream_t s, int64_t rows, int64_t k, T* C, int64_t ldc, double beta) {
    int mode = beta_mode(beta);
    if (mode == 1 || rows == 0 || k == 0) return GSA_OK;
    if (ldc == rows && mode == 0) {
        GSA_HIP(hipMemsetAsync(C, 0, sizeof(T) * rows * k, s));
        return GSA_OK;
    }
    hipLaunchKernelGGL(k_beta<T>, dim3(grid_for(rows * k, 256)), dim3(256), 0, s, rows, k, C,
                       ldc, (T)beta, mode);
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

// Row-block plan for one call: use the caller's plan when it matches, else build
// a temporary one in scratch.
struct PlanUse {
    const int64_t* blk_row = nullptr;
    int64_t nblocks = 0;
    Scratch tmp;
    explicit PlanUse(hipStream_t s) : tmp(s) {}
    int get(hipStream_t s, gsa_plan plan, int itype, int64_t m, int64_t nnz, const void* ptr) {
        if (plan && plan->kind == 0 && plan->csr.m == m && plan->csr.nnz == nnz &&
            plan->csr.itype == itype && plan->csr.ptr == ptr) {
            blk_row = plan->csr.blk_row;
            nblocks = plan->csr.nblocks;
            return GSA_OK;
        }
        nblocks = nnz / SPMV_NB + 1;
        GSA_TRY(tmp.alloc(sizeof(int64_t) * (nblocks + 1)));
        GSA_TRY(build_csr_plan(s, itype, m, nnz, ptr, tmp.as<int64_t>(), nblocks));
        blk_row = tmp.as<int64_t>();
        return GSA_OK;
    }
};

// Gather-shaped product on CSR arrays: C (rows m) = β C + α A B, B has n rows.
template <typename T, typename I>
static int gather_csr(hipStream_t s, int64_t m, int64_t n, int64_t k, const I* rowptr,
                      const I* colval, const T* nzval, int64_t nnz, const T* B, int64_t ldb,
                      int opB, T* C, int64_t ldc, double alpha, double beta, gsa_plan plan,
                      int itype) {
    if (m == 0 || k == 0) return GSA_OK;
    if (n == 0 || nnz == 0) {  // A*B contributes nothing: only β applies
        return apply_beta<T>(s, m, k, C, ldc, beta);
    }
    PlanUse pu(s);
    GSA_TRY(pu.get(s, plan, itype, m, nnz, rowptr));
    const int mode = beta_mode(beta);
    const unsigned grid = (unsigned)pu.nblocks;
    if (k == 1) {
        const T* x = B;
        Scratch xs(s);
        if (opB && ldb != 1) {  // strided vector: compact it first
            GSA_TRY(xs.alloc(sizeof(T) * n));
            GSA_HIP(hipMemcpy2DAsync(xs.p, sizeof(T), B, sizeof(T) * ldb, sizeof(T), n,
                                     hipMemcpyDeviceToDevice, s));
            x = xs.as<T>();
        }
        hipLaunchKernelGGL((k_csr_spmv<T, I>), dim3(grid), dim3(WG), 0, s, m, n, nnz, rowptr,
                           colval, nzval, x, C, (T)alpha, (T)beta, mode, pu.blk_row);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    }
    const T* BT = B;
    int64_t ldbt = ldb;
    Scratch bt(s);
    if (!opB) {
        GSA_TRY(bt.alloc(sizeof(T) * n * k));
        hipLaunchKernelGGL(k_to_rowmajor<T>, dim3((unsigned)cdiv(n, 64)), dim3(256), 0, s, n, k,
                           B, ldb, bt.as<T>());
        GSA_HIP(hipGetLastError());
        BT = bt.as<T>();
        ldbt = k;
    }
#define GSA_SPMM_LAUNCH(L)                                                                  \
    hipLaunchKernelGGL((k_csr_spmm<T, I, L>), dim3(grid), dim3(WG), 0, s, m, n, k, nnz, rowptr, \
                       colval, nzval, BT, ldbt, C, ldc, (T)alpha, (T)beta, mode, pu.blk_row)
    if (k <= 2) GSA_SPMM_LAUNCH(2);
    else if (k <= 4) GSA_SPMM_LAUNCH(4);
    else if (k <= 8) GSA_SPMM_LAUNCH(8);
    else if (k <= 16) GSA_SPMM_LAUNCH(16);
    else if (k <= 32) GSA_SPMM_LAUNCH(32);
    else GSA_SPMM_LAUNCH(64);
#undef GSA_SPMM_LAUNCH
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

// Scatter-shaped product on CSR arrays (m rows stored): C (n rows) += α Aᵀ B.
template <typename T, typename I>
static int scatter_csr(hipStream_t s, int64_t m, int64_t n, int64_t k, const I* rowptr,
                       const I* colval, const T* nzval, int64_t nnz, const T* B, int64_t ldb,
                       int opB, T* C, int64_t ldc, double alpha, double beta, gsa_plan plan,
                       int itype) {
    if (n == 0 || k == 0) return GSA_OK;
    GSA_TRY(apply_beta<T>(s, n, k, C, ldc, beta));
    if (m == 0 || nnz == 0) return GSA_OK;
    PlanUse pu(s);
    GSA_TRY(pu.get(s, plan, itype, m, nnz, rowptr));
    hipLaunchKernelGGL((k_csr_scatter<T, I>), dim3((unsigned)pu.nblocks), dim3(WG), 0, s, m, n, k,
                       nnz, rowptr, colval, nzval, B, ldb, opB, C, ldc, (T)alpha, pu.blk_row);
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

static int check_common(int opA, int opB, int64_t m, int64_t n, int64_t k, int64_t nnz,
                        int64_t ldb, int64_t ldc) {
    GSA_CHECK(opA >= GSA_OP_N && opA <= GSA_OP_C, GSA_ERR_INVALID, "invalid opA %d", opA);
    GSA_CHECK(opB >= GSA_OP_N && opB <= GSA_OP_C, GSA_ERR_INVALID, "invalid opB %d", opB);
    GSA_CHECK(m >= 0 && n >= 0 && k >= 0 && nnz >= 0, GSA_ERR_INVALID,
              "negative dimension (m=%lld n=%lld k=%lld nnz=%lld)", (long long)m, (long long)n,
              (long long)k, (long long)nnz);
    const int64_t rows_b = (opA == GSA_OP_N) ? n : m;
    const int64_t rows_c = (opA == GSA_OP_N) ? m : n;
    if (opB == GSA_OP_N) {
        GSA_CHECK(ldb >= (rows_b > 1 ? rows_b : 1) || k <= 1, GSA_ERR_INVALID,
                  "ldb %lld < rows of B %lld", (long long)ldb, (long long)rows_b);
    } else {
        GSA_CHECK(ldb >= (k > 1 ? k : 1), GSA_ERR_INVALID, "ldb %lld < k %lld for transposed B",
                  (long long)ldb, (long long)k);
    }
    GSA_CHECK(ldc >= (rows_c > 1 ? rows_c : 1) || k <= 1, GSA_ERR_INVALID,
              "ldc %lld < rows of C %lld", (long long)ldc, (long long)rows_c);
    return GSA_OK;
}

}  // namespace gsa

using namespace gsa;

extern "C" {

int gsa_plan_create(void* stream, int itype, int64_t m, int64_t nnz, const void* ptr,
                    gsa_plan* plan) {
    GSA_TRY(ensure_init());
    GSA_CHECK(plan && m >= 0 && nnz >= 0 && (ptr || m == 0), GSA_ERR_INVALID,
              "gsa_plan_create: bad arguments");
    GSA_CHECK(itype == GSA_I32 || itype == GSA_I64, GSA_ERR_UNSUPPORTED, "bad itype %d", itype);
    hipStream_t s = as_stream(stream);
    auto* p = new gsa_plan_s();
    p->kind = 0;
    p->csr.itype = itype;
    p->csr.m = m;
    p->csr.nnz = nnz;
    p->csr.ptr = ptr;
    p->csr.nblocks = nnz / SPMV_NB + 1;
    hipError_t e = hipMalloc(&p->csr.blk_row, sizeof(int64_t) * (p->csr.nblocks + 1));
    if (e != hipSuccess) { delete p; return hip_fail(e, "hipMalloc(plan)"); }
    int r = build_csr_plan(s, itype, m, nnz, ptr, p->csr.blk_row, p->csr.nblocks);
    if (r != GSA_OK) { (void)hipFree(p->csr.blk_row); delete p; return r; }
    *plan = p;
    return GSA_OK;
}

int gsa_plan_destroy(gsa_plan plan) {
    if (!plan) return GSA_OK;
    if (plan->kind == 0 && plan->csr.blk_row) {
        // the plan may still be in use by queued work on any stream
        GSA_HIP(hipDeviceSynchronize());
        GSA_HIP(hipFree(plan->csr.blk_row));
    }
    if (plan->kind == 1) batch_plan_release(plan);
    delete plan;
    return GSA_OK;
}

int gsa_spmm_csr(void* stream, int vtype, int itype, int opA, int64_t m, int64_t n, int64_t k,
                 const void* rowptr, const void* colval, const void* nzval, int64_t nnz,
                 const void* B, int64_t ldb, int opB, void* C, int64_t ldc, double alpha,
                 double beta, gsa_plan plan) {
    GSA_TRY(ensure_init());
    GSA_TRY(check_common(opA, opB, m, n, k, nnz, ldb, ldc));
    hipStream_t s = as_stream(stream);
    return dispatch_vi(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        if (opA == GSA_OP_N)
            return gather_csr<T, I>(s, m, n, k, (const I*)rowptr, (const I*)colval,
                                    (const T*)nzval, nnz, (const T*)B, ldb, opB != GSA_OP_N,
                                    (T*)C, ldc, alpha, beta, plan, itype);
        return scatter_csr<T, I>(s, m, n, k, (const I*)rowptr, (const I*)colval, (const T*)nzval,
                                 nnz, (const T*)B, ldb, opB != GSA_OP_N, (T*)C, ldc, alpha, beta,
                                 plan, itype);
    });
}

int gsa_spmm_csc(void* stream, int vtype, int itype, int opA, int64_t m, int64_t n, int64_t k,
                 const void* colptr, const void* rowval, const void* nzval, int64_t nnz,
                 const void* B, int64_t ldb, int opB, void* C, int64_t ldc, double alpha,
                 double beta, gsa_plan plan) {
    GSA_TRY(ensure_init());
    GSA_TRY(check_common(opA, opB, m, n, k, nnz, ldb, ldc));
    hipStream_t s = as_stream(stream);
    // A (m x n) in CSC is Aᵀ (n x m) in CSR: A x is the scatter, Aᵀ x the gather.
    return dispatch_vi(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        if (opA == GSA_OP_N)
            return scatter_csr<T, I>(s, n, m, k, (const I*)colptr, (const I*)rowval,
                                     (const T*)nzval, nnz, (const T*)B, ldb, opB != GSA_OP_N,
                                     (T*)C, ldc, alpha, beta, plan, itype);
        return gather_csr<T, I>(s, n, m, k, (const I*)colptr, (const I*)rowval, (const T*)nzval,
                                nnz, (const T*)B, ldb, opB != GSA_OP_N, (T*)C, ldc, alpha, beta,
                                plan, itype);
    });
}

int gsa_spmm_coo(void* stream, int vtype, int itype, int opA, int64_t m, int64_t n, int64_t k,
                 const void* rowind, const void* colind, const void* nzval, int64_t nnz,
                 const void* B, int64_t ldb, int opB, void* C, int64_t ldc, double alpha,
                 double beta) {
    GSA_TRY(ensure_init());
    GSA_TRY(check_common(opA, opB, m, n, k, nnz, ldb, ldc));
    hipStream_t s = as_stream(stream);
    const int trans = opA != GSA_OP_N;
    const int64_t rows_out = trans ? n : m, cols_in = trans ? m : n;
    return dispatch_vi(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        if (rows_out == 0 || k == 0) return (int)GSA_OK;
        GSA_TRY(apply_beta<T>(s, rows_out, k, (T*)C, ldc, beta));
        if (nnz == 0 || cols_in == 0) return (int)GSA_OK;
        hipLaunchKernelGGL((k_coo_scatter<T, I>), dim3(grid_for(nnz * k, 256)), dim3(256), 0, s,
                           rows_out, cols_in, k, nnz, (const I*)rowind, (const I*)colind,
                           (const T*)nzval, trans, (const T*)B, ldb, opB != GSA_OP_N, (T*)C, ldc,
                           (T)alpha);
        GSA_HIP(hipGetLastError());
        return (int)GSA_OK;
    });
}

}  // extern "C"
