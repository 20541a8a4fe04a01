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

// Entries per workgroup of visit_owned_entries: the row-block plan's block size.
constexpr int ENT_CAP = SPMV_NB;
constexpr int ENT_PER_LANE = ENT_CAP / ROW_WG;
static_assert(ENT_CAP % ROW_WG == 0, "a block's entries split evenly over its lanes");

// One lane's share of a batch of entries, in rounds whose loads are all in
// flight before any of them is used: pre[u] holds what was read through the
// entry's position alone (read by visit_owned_entries before the rows were
// known); L[u] = load_fn(r, j, pre[u]) adds what is read through the row (it
// must read through r only: an entry with ok[u] false gets the row rs, known
// to exist, and any j); then fetch_fn(r, j, L[u]) (reads that depend on the
// previous rounds, such as a gather through the index), then use_fn(r, j,
// L[u]).  Entries with ok[u] false are neither fetched nor used.
template <int EB, typename P, typename LoadFn, typename FetchFn, typename UseFn>
__device__ __forceinline__ void lane_batch(const int64_t* r, const int64_t* j, const bool* ok,
                                           const P* pre, int64_t rs, LoadFn load_fn,
                                           FetchFn fetch_fn, UseFn use_fn) {
    using L = decltype(load_fn(int64_t(0), int64_t(0), pre[0]));
    L ld[EB];
#pragma unroll
    for (int u = 0; u < EB; ++u) ld[u] = load_fn(ok[u] ? r[u] : rs, j[u], pre[u]);
#pragma unroll
    for (int u = 0; u < EB; ++u)
        if (ok[u]) fetch_fn(r[u], j[u], ld[u]);
#pragma unroll
    for (int u = 0; u < EB; ++u)
        if (ok[u]) use_fn(r[u], j[u], ld[u]);
}

// lane_batch over (load_fn, fetch_fn, use_fn) in batches of EB entries, as
// the batch_fn of visit_owned_entries.
template <int EB, typename LoadFn, typename FetchFn, typename UseFn>
__device__ __forceinline__ auto lane_rounds(LoadFn load_fn, FetchFn fetch_fn, UseFn use_fn) {
    return [=](const int64_t* r, const int64_t* j, const bool* ok, const auto* pre, int64_t rs) {
        static_assert(ENT_PER_LANE % EB == 0, "a batch divides a lane's share");
#pragma unroll
        for (int g = 0; g < ENT_PER_LANE; g += EB)
            lane_batch<EB>(r + g, j + g, ok + g, pre + g, rs, load_fn, fetch_fn, use_fn);
    };
}

// Entry-parallel visit of a CSR-shaped array under a row-block plan (blk_row,
// built by build_csr_plan over `ptr` with blocks of ENT_CAP entries).
// Workgroup b takes the entries [b * ENT_CAP, (b + 1) * ENT_CAP) whatever rows
// they fall in, ENT_PER_LANE per lane at positions j = b * ENT_CAP +
// threadIdx.x + u * ROW_WG, so that:
//  * pre[u] = pre_fn(j) (the reads through the position: index, value) is
//    issued at once, with the plan's row bounds, before any row is known;
//  * row_fn(r, s, e) runs once per row the plan gives the workgroup (rows
//    starting in its range; zero-based entries [s, e)), one lane per row;
//  * each entry's row is then written to LDS (rid): the rows starting in the
//    range and the one row running into it from the left; a row with up to
//    ROW_SHORT entries in the range by its own lane, a longer one by the
//    whole workgroup (no more than ROW_MAX_LONG such rows fit the range);
//  * batch_fn(r[], j[], ok[], pre[], rs) then runs once per lane over its
//    ENT_PER_LANE entries: entry u is entry j[u] of row r[u] where ok[u], and
//    rs is a row known to exist.  Writes made by row_fn are visible to it.
// A long row costs no more rounds of loads than a short one: it is split
// over the workgroups its entries fall in.  rid[k] = -1 marks an entry no row
// claims (malformed pointers), which is skipped.
template <typename I, typename RowFn, typename PreFn, typename BatchFn>
__device__ __forceinline__ void visit_owned_entries(int64_t m, int64_t nnz,
                                                    const I* __restrict__ ptr,
                                                    const int64_t* __restrict__ blk_row,
                                                    RowFn row_fn, PreFn pre_fn,
                                                    BatchFn batch_fn) {
    __shared__ int64_t rid[ENT_CAP];
    __shared__ int64_t longs[ROW_MAX_LONG][3];  // row, first and end entry in the range
    __shared__ int nlong;
    using P = decltype(pre_fn(int64_t(0)));
    const int64_t b = blockIdx.x;
    const int64_t lo = b * (int64_t)ENT_CAP;
    const int64_t hi = lo + ENT_CAP < nnz ? lo + ENT_CAP : nnz;
    P pre[ENT_PER_LANE];
    int64_t jj[ENT_PER_LANE];
    if (lo < hi) {
#pragma unroll
        for (int u = 0; u < ENT_PER_LANE; ++u) {
            jj[u] = lo + threadIdx.x + (int64_t)u * ROW_WG;
            pre[u] = pre_fn(jj[u] < hi ? jj[u] : lo);
        }
    }
    const int64_t r0 = dclamp(blk_row[b], 0, m);
    const int64_t r1 = dclamp(blk_row[b + 1], r0, m);
    const int64_t ra = r0 > 0 ? r0 - 1 : 0;  // may run into the range from the left
#pragma unroll
    for (int u = 0; u < ENT_PER_LANE; ++u) rid[threadIdx.x + u * ROW_WG] = -1;
    if (threadIdx.x == 0) nlong = 0;
    __syncthreads();
    for (int64_t r = ra + threadIdx.x; r < r1; r += ROW_WG) {
        const int64_t s = dclamp((int64_t)ptr[r] - 1, 0, nnz);
        const int64_t e = dclamp((int64_t)ptr[r + 1] - 1, s, nnz);
        if (r >= r0) row_fn(r, s, e);
        const int64_t a = s > lo ? s : lo, z = e < hi ? e : hi;
        if (z - a > ROW_SHORT) {
            const int slot = atomicAdd(&nlong, 1);
            if (slot < ROW_MAX_LONG) {
                longs[slot][0] = r;
                longs[slot][1] = a;
                longs[slot][2] = z;
                continue;
            }
        }
        for (int64_t j = a; j < z; ++j) rid[j - lo] = r;
    }
    __syncthreads();
    const int nl = nlong < ROW_MAX_LONG ? nlong : ROW_MAX_LONG;
    if (nl > 0) {
        for (int q = 0; q < nl; ++q) {  // disjoint ranges: no barrier between rows
            const int64_t r = longs[q][0], z = longs[q][2];
            for (int64_t j = longs[q][1] + threadIdx.x; j < z; j += ROW_WG) rid[j - lo] = r;
        }
        __syncthreads();
    }
    if (lo >= hi) return;
    int64_t rr[ENT_PER_LANE];
    bool ok[ENT_PER_LANE];
#pragma unroll
    for (int u = 0; u < ENT_PER_LANE; ++u) {
        rr[u] = jj[u] < hi ? rid[threadIdx.x + u * ROW_WG] : -1;
        ok[u] = rr[u] >= 0 && rr[u] < m;
    }
    batch_fn(rr, jj, ok, pre, ra < m ? ra : m - 1);
}

}  // namespace gsa
