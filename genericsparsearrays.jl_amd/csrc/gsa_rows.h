// gsa_rows.h — device helpers shared by the row-structured kernels.
#pragma once
#include "gsa_common.h"

namespace gsa {

constexpr int ROW_WG = 256;
constexpr int ROW_SHORT = 64;                         // rows up to this length: one lane
constexpr int ROW_MAX_LONG = SPMV_NB / (ROW_SHORT + 1) + 2;

__device__ __forceinline__ int64_t dclamp(int64_t v, int64_t lo, int64_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

__device__ __forceinline__ void atomic_inc(int64_t* p) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), 1ull);
}
__device__ __forceinline__ void atomic_inc(int32_t* p) { atomicAdd(reinterpret_cast<int*>(p), 1); }

// out = (β==0 ? 0 : β==1 ? y : y*β) + acc*α — the two roundings of
// `_rmul_or_fill!(C, β)` followed by `C[row,k] += tmp*α` (matrix_csr.jl:192,
// matrix_csr_kernels.jl:21).
template <typename T>
__device__ __forceinline__ void store_out(T* __restrict__ y, int64_t idx, T acc, T alpha,
                                          T beta, int mode) {
    T t = mul_scalar(acc, alpha);
    T base;
    if (mode == 0) base = T(0);
    else if (mode == 1) base = y[idx];
    else base = mul_scalar(y[idx], beta);
    y[idx] = base + t;
}

// Exclusive prefix sum over the 256 lanes of a workgroup; *total gets the sum.
// Every lane of the workgroup must call it.
template <typename T>
__device__ __forceinline__ T wg_exclusive_sum(T v, T* lds4, T* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds4[wave] = x;
    __syncthreads();
    T pre = 0, all = 0;
#pragma unroll
    for (int w = 0; w < ROW_WG / 64; ++w) {
        T s = lds4[w];
        if (w < wave) pre += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return pre + x - v;
}

// Visit every row owned by this workgroup under a row-block plan
// (blk_row, built by build_csr_plan over `ptr`).  Rows of length <= ROW_SHORT
// go to short_fn(r, s, e) on one lane; longer rows go to long_fn(r, s, e),
// called by ALL lanes of the workgroup together (it may synchronise).
// s, e are 0-based entry positions.
template <typename I, typename ShortFn, typename LongFn>
__device__ __forceinline__ void visit_owned_rows(int64_t m, int64_t nnz, const I* __restrict__ ptr,
                                                 const int64_t* __restrict__ blk_row,
                                                 ShortFn short_fn, LongFn long_fn) {
    __shared__ int64_t longs[ROW_MAX_LONG];
    __shared__ int nlong;
    const int64_t b = blockIdx.x;
    const int64_t r0 = dclamp(blk_row[b], 0, m);
    const int64_t r1 = dclamp(blk_row[b + 1], r0, m);
    if (r0 >= r1) return;
    if (threadIdx.x == 0) nlong = 0;
    __syncthreads();
    for (int64_t r = r0 + threadIdx.x; r < r1; r += ROW_WG) {
        int64_t s = dclamp((int64_t)ptr[r] - 1, 0, nnz);
        int64_t e = dclamp((int64_t)ptr[r + 1] - 1, s, nnz);
        if (e - s <= ROW_SHORT) {
            short_fn(r, s, e);
        } else {
            int slot = atomicAdd(&nlong, 1);
            if (slot < ROW_MAX_LONG) longs[slot] = r;
            else short_fn(r, s, e);  // unreachable for a valid plan
        }
    }
    __syncthreads();
    const int nl = nlong < ROW_MAX_LONG ? nlong : ROW_MAX_LONG;
    for (int q = 0; q < nl; ++q) {
        const int64_t r = longs[q];
        int64_t s = dclamp((int64_t)ptr[r] - 1, 0, nnz);
        int64_t e = dclamp((int64_t)ptr[r + 1] - 1, s, nnz);
        long_fn(r, s, e);
        __syncthreads();
    }
}

// Entries staged per chunk by visit_owned_entries (their rows, in LDS).
constexpr int ENT_CAP = 2048;

// Visit the rows and entries owned by this workgroup under a row-block plan,
// entry-parallel: row_fn(r, s, e) once per owned row of length <= ROW_SHORT
// (one lane each); then entry_fn(r, j) for every entry j of those rows, the
// lanes taking consecutive entries (coalesced reads of the index and value
// arrays; each chunk of ENT_CAP entries gets its rows written to LDS first);
// then long_fn(r, s, e) for each longer row, called by ALL lanes together.
// Writes made by row_fn are visible to entry_fn, and both to long_fn.
template <typename I, typename RowFn, typename EntryFn, typename LongFn>
__device__ __forceinline__ void visit_owned_entries(int64_t m, int64_t nnz,
                                                    const I* __restrict__ ptr,
                                                    const int64_t* __restrict__ blk_row,
                                                    RowFn row_fn, EntryFn entry_fn,
                                                    LongFn long_fn) {
    __shared__ int64_t rid[ENT_CAP];
    __shared__ int64_t longs[ROW_MAX_LONG];
    __shared__ int nlong;
    const int64_t b = blockIdx.x;
    const int64_t r0 = dclamp(blk_row[b], 0, m);
    const int64_t r1 = dclamp(blk_row[b + 1], r0, m);
    if (r0 >= r1) return;
    const int64_t P0 = dclamp((int64_t)ptr[r0] - 1, 0, nnz);
    const int64_t P1 = dclamp((int64_t)ptr[r1] - 1, P0, nnz);
    if (threadIdx.x == 0) nlong = 0;
    __syncthreads();
    for (int64_t r = r0 + threadIdx.x; r < r1; r += ROW_WG) {
        const int64_t s = dclamp((int64_t)ptr[r] - 1, 0, nnz);
        const int64_t e = dclamp((int64_t)ptr[r + 1] - 1, s, nnz);
        if (e - s <= ROW_SHORT) {
            row_fn(r, s, e);
        } else {
            const int slot = atomicAdd(&nlong, 1);
            if (slot < ROW_MAX_LONG) longs[slot] = r;
        }
    }
    __syncthreads();
    for (int64_t c0 = P0; c0 < P1; c0 += ENT_CAP) {
        const int64_t c1 = c0 + ENT_CAP < P1 ? c0 + ENT_CAP : P1;
        for (int64_t k = threadIdx.x; k < c1 - c0; k += ROW_WG) rid[k] = -1;
        __syncthreads();
        for (int64_t r = r0 + threadIdx.x; r < r1; r += ROW_WG) {
            const int64_t s = dclamp((int64_t)ptr[r] - 1, 0, nnz);
            const int64_t e = dclamp((int64_t)ptr[r + 1] - 1, s, nnz);
            if (e - s > ROW_SHORT) continue;
            const int64_t lo = s > c0 ? s : c0, hi = e < c1 ? e : c1;
            for (int64_t j = lo; j < hi; ++j) rid[j - c0] = r;
        }
        __syncthreads();
        for (int64_t k = threadIdx.x; k < c1 - c0; k += ROW_WG) {
            const int64_t r = rid[k];
            if (r >= 0) entry_fn(r, c0 + k);
        }
        __syncthreads();
    }
    const int nl = nlong < ROW_MAX_LONG ? nlong : ROW_MAX_LONG;
    for (int q = 0; q < nl; ++q) {
        const int64_t r = longs[q];
        const int64_t s = dclamp((int64_t)ptr[r] - 1, 0, nnz);
        const int64_t e = dclamp((int64_t)ptr[r + 1] - 1, s, nnz);
        long_fn(r, s, e);
        __syncthreads();
    }
}

}  // namespace gsa
