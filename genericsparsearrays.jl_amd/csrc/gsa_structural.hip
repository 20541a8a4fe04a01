// gsa_structural.hip — dropzeros, sparse +, kron and SpGEMM on gfx950.
//
// Variable-size results follow the reference's two-phase shape (count, prefix
// sum, host reads the total, caller allocates, fill):
//  * CSR dropzeros: per-row count of nonzeros, pointer scan, order-preserving
//    copy (kernel_count_nonzeros_csr!/kernel_dropzeros_csr!,
//    src/matrix_csr/matrix_csr_kernels.jl:410-443).  Long rows are compacted by
//    a whole workgroup with a prefix sum of keep flags.
//  * CSR +: the two-pointer merge per row of kernel_count_nnz_per_row_csr! and
//    kernel_merge_csr! (matrix_csr_kernels.jl:148-248) with the following
//    dropzeros folded into the same two passes: an entry is kept when its
//    (summed) value is nonzero.
//  * COO +: concatenation, stable radix sort on column-major keys, left-to-right
//    sums of equal-key runs, zero sums dropped (src/matrix_coo/matrix_coo.jl:300-384).
//  * kron COO: entry (a, b) -> a*nnzB + b, as kernel_kron_coo!
//    (src/matrix_coo/matrix_coo_kernels.jl:144-183), one workgroup per A entry.
//  * kron CSR: for canonical operands (strictly ascending columns per row) the
//    canonical result is written directly: row (iA, iB) starts at
//    (rowptrA[iA]-1)*nnzB + lenA(iA)*(rowptrB[iB]-1) and lists jA-major, jB-minor,
//    which is exactly the reference's sorted COO -> CSR result.  Other operands
//    take the reference's route: kron COO then the keyed sort.
//  * SpGEMM: every product a_ik*b_kj is numbered in the reference's order
//    (A-row order, then B-row order) and each output entry is the left-to-right
//    sum from +0.0 of its products — the accumulation order and arithmetic of
//    kernel_spgemm_csr! (src/matrix_csr/matrix_csr_kernels.jl:293-343), with
//    cancellation zeros kept.  Rows of at most 2048 products are sorted by
//    (column, number) inside a lane group or a workgroup, in a symbolic (count)
//    and a numeric (write) phase; a product with a longer row takes the global
//    expand-sort-compress path (keyed (row, col), stable radix sort, run sums).
#include "gsa_scan.h"
#include "gsa_rows.h"
#include "gsa_merge.h"
#include "gsa_radix_pairs.h"

namespace gsa {

int expand_rows(hipStream_t s, int itype, int64_t m, int64_t nnz, const void* ptr, void* out);

// ------------------------------------------------------------ dropzeros ----
// Order-preserving stream compaction over the stored entries
// (kernel_dropzeros_count / _fill, src/matrix_csr/matrix_csr.jl:723-745): rows
// are contiguous in storage, so a kept entry's place in the result is the
// number of kept entries before it, and a row's new pointer is that number at
// the row's start.  Workgroup b takes the entries [2048b, 2048b + 2048) (the
// row-block plan's blocks), ENT_PER_LANE per lane at j = 2048b + threadIdx.x +
// u * ROW_WG.

// Each lane's ENT_PER_LANE keep flags, and each one's rank among the kept
// entries of the workgroup's block (entry order), from wavefront ballots and
// one barrier; *total gets the block's kept count.
template <typename T>
__device__ __forceinline__ void dz_block_ranks(const T* __restrict__ val, int64_t lo, int64_t hi,
                                               bool (&keep)[ENT_PER_LANE],
                                               int (&rank)[ENT_PER_LANE], int* total) {
    __shared__ int wcnt[ENT_PER_LANE][ROW_WG / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T v[ENT_PER_LANE];
#pragma unroll
    for (int u = 0; u < ENT_PER_LANE; ++u) {
        const int64_t j = lo + threadIdx.x + (int64_t)u * ROW_WG;
        v[u] = val[j < hi ? j : lo];
    }
    int below[ENT_PER_LANE];
#pragma unroll
    for (int u = 0; u < ENT_PER_LANE; ++u) {
        const int64_t j = lo + threadIdx.x + (int64_t)u * ROW_WG;
        keep[u] = j < hi && v[u] != T(0);
        const uint64_t bal = __ballot(keep[u]);
        below[u] = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wcnt[u][wave] = __popcll(bal);
    }
    __syncthreads();
    int run = 0;
#pragma unroll
    for (int u = 0; u < ENT_PER_LANE; ++u) {
        int before = 0;
#pragma unroll
        for (int w = 0; w < ROW_WG / 64; ++w) before += w < wave ? wcnt[u][w] : 0;
        rank[u] = run + before + below[u];
#pragma unroll
        for (int w = 0; w < ROW_WG / 64; ++w) run += wcnt[u][w];
    }
    *total = run;
}

// Kept entries of each block.
template <typename T>
__global__ __launch_bounds__(ROW_WG) void k_dz_block_sums(int64_t nnz, const T* __restrict__ val,
                                                          int64_t* __restrict__ bsum) {
    const int64_t lo = blockIdx.x * (int64_t)ENT_CAP;
    const int64_t hi = lo + ENT_CAP < nnz ? lo + ENT_CAP : nnz;
    bool keep[ENT_PER_LANE];
    int rank[ENT_PER_LANE], total = 0;
    if (lo < hi) dz_block_ranks<T>(val, lo, hi, keep, rank, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// The new row pointer of every row starting in the block (rows at the end of
// the entries, s = nnz, belong to the last block): 1 + the kept entries before
// the row's start.  incl is the inclusive scan of the block sums.
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_dz_rowptr(int64_t m, int64_t nnz,
                                                      const I* __restrict__ ptr,
                                                      const T* __restrict__ val,
                                                      const int64_t* __restrict__ blk_row,
                                                      const int64_t* __restrict__ incl,
                                                      I* __restrict__ ptr_out) {
    __shared__ int pos[ENT_CAP + 1];
    const int64_t b = blockIdx.x;
    const int64_t lo = b * (int64_t)ENT_CAP;
    const int64_t hi = lo + ENT_CAP < nnz ? lo + ENT_CAP : nnz;
    const int64_t base = b > 0 ? incl[b - 1] : 0;
    bool keep[ENT_PER_LANE];
    int rank[ENT_PER_LANE], total = 0;
    if (lo < hi) dz_block_ranks<T>(val, lo, hi, keep, rank, &total);
#pragma unroll
    for (int u = 0; u < ENT_PER_LANE; ++u) {
        // the kept entries of the block before entry k: the rank of a kept
        // entry, or of the next kept one
        const int k = threadIdx.x + u * ROW_WG;
        pos[k] = lo < hi ? rank[u] : 0;
    }
    if (threadIdx.x == 0) pos[ENT_CAP] = total;
    __syncthreads();
    const int64_t r0 = dclamp(blk_row[b], 0, m);
    const int64_t r1 = dclamp(blk_row[b + 1], r0, m);
    for (int64_t r = r0 + threadIdx.x; r < r1; r += ROW_WG) {
        const int64_t k = dclamp(dclamp((int64_t)ptr[r] - 1, 0, nnz) - lo, 0, ENT_CAP);
        ptr_out[r] = (I)(1 + base + pos[k]);
    }
    if (b == (int64_t)gridDim.x - 1 && threadIdx.x == 0) ptr_out[m] = (I)(1 + base + total);
}

// The kept entries, each to its block's offset plus its rank in the block.
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_dz_compact(int64_t nnz, const I* __restrict__ idx,
                                                       const T* __restrict__ val,
                                                       const int64_t* __restrict__ incl,
                                                       I* __restrict__ idx_out,
                                                       T* __restrict__ val_out) {
    const int64_t b = blockIdx.x;
    const int64_t lo = b * (int64_t)ENT_CAP;
    const int64_t hi = lo + ENT_CAP < nnz ? lo + ENT_CAP : nnz;
    if (lo >= hi) return;
    const int64_t base = b > 0 ? incl[b - 1] : 0;
    bool keep[ENT_PER_LANE];
    int rank[ENT_PER_LANE], total = 0;
    dz_block_ranks<T>(val, lo, hi, keep, rank, &total);
#pragma unroll
    for (int u = 0; u < ENT_PER_LANE; ++u) {
        if (!keep[u]) continue;
        const int64_t j = lo + threadIdx.x + (int64_t)u * ROW_WG;
        idx_out[base + rank[u]] = idx[j];
        val_out[base + rank[u]] = val[j];
    }
}

// Kept entries per 2048-entry block, inclusively scanned, into incl[0 .. nb).
template <typename T>
static int dz_block_prefix(hipStream_t s, int64_t nnz, const T* val, int64_t nb, int64_t* incl) {
    hipLaunchKernelGGL((k_dz_block_sums<T>), dim3((unsigned)nb), dim3(ROW_WG), 0, s, nnz, val, incl);
    GSA_HIP(hipGetLastError());
    return inclusive_scan<int64_t>(s, incl, incl, nb, 0);
}

// ------------------------------------------------------------ COO helpers --
template <typename T>
__global__ void k_nonzero_flags(int64_t n, const T* __restrict__ v, int64_t* __restrict__ f) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        f[i] = v[i] != T(0);
}

template <typename T, typename I>
__global__ void k_compact(int64_t n, const int64_t* __restrict__ incl, const I* __restrict__ r,
                          const I* __restrict__ c, const T* __restrict__ v, I* __restrict__ ro,
                          I* __restrict__ co, T* __restrict__ vo) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t prev = i ? incl[i - 1] : 0;
        if (incl[i] != prev) { ro[prev] = r[i]; co[prev] = c[i]; vo[prev] = v[i]; }
    }
}

// concatenated (A then B) column-major keys with positions
template <typename I, typename V>
__global__ void k_concat_keys(int64_t nA, const I* __restrict__ rA, const I* __restrict__ cA,
                              int64_t nB, const I* __restrict__ rB, const I* __restrict__ cB,
                              int64_t m, int64_t n, uint64_t* __restrict__ keys,
                              V* __restrict__ pos) {
    const int64_t tot = nA + nB;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = i < nA ? (int64_t)rA[i] : (int64_t)rB[i - nA];
        int64_t c = i < nA ? (int64_t)cA[i] : (int64_t)cB[i - nA];
        r = dclamp(r - 1, 0, m - 1);
        c = dclamp(c - 1, 0, n - 1);
        keys[i] = (uint64_t)c * (uint64_t)m + (uint64_t)r;
        pos[i] = (V)i;
    }
}

// key / m and key % m of a column-major key, for m > 0: a double-precision
// estimate of the quotient corrected to the exact one (keys below 2^53, which
// covers m * n < 2^53), integer division otherwise.
__device__ __forceinline__ void key_split(uint64_t key, uint64_t m, double inv_m, uint64_t& q,
                                          uint64_t& r) {
    if (key < (1ull << 53)) {
        int64_t qq = (int64_t)((double)key * inv_m);
        int64_t rr = (int64_t)key - qq * (int64_t)m;
        while (rr < 0) { --qq; rr += (int64_t)m; }
        while (rr >= (int64_t)m) { ++qq; rr -= (int64_t)m; }
        q = (uint64_t)qq;
        r = (uint64_t)rr;
    } else {
        q = key / m;
        r = key - q * m;
    }
}

// Order of the two operands' storage, from their one-based indices (clamped
// into range): bit 0 (A) / bit 2 (B) when (col, row) decreases somewhere in
// the part, bit 1 / bit 3 when (row, col) does.  One update per workgroup.
template <typename I>
__global__ void k_parts_order(int64_t nA, int64_t tot, const I* __restrict__ rA,
                              const I* __restrict__ cA, const I* __restrict__ rB,
                              const I* __restrict__ cB, int64_t m, int64_t n,
                              int* __restrict__ flag) {
    int f = 0;
    for (int64_t i = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i == nA) continue;
        const bool inA = i < nA;
        const I* rr = inA ? rA : rB;
        const I* cc = inA ? cA : cB;
        const int64_t k = inA ? i : i - nA;
        const int64_t r0 = dclamp((int64_t)rr[k - 1] - 1, 0, m - 1);
        const int64_t r1 = dclamp((int64_t)rr[k] - 1, 0, m - 1);
        const int64_t c0 = dclamp((int64_t)cc[k - 1] - 1, 0, n - 1);
        const int64_t c1 = dclamp((int64_t)cc[k] - 1, 0, n - 1);
        const int sh = inA ? 0 : 2;
        if (c1 < c0 || (c1 == c0 && r1 < r0)) f |= 1 << sh;
        if (r1 < r0 || (r1 == r0 && c1 < c0)) f |= 2 << sh;
    }
    int g = 0;
    for (int b = 0; b < 4; ++b) g |= __syncthreads_or(f & (1 << b)) ? 1 << b : 0;
    if (threadIdx.x == 0 && (g & ~*(volatile int*)flag)) atomicOr(flag, g);
}

// One-based row pointer of entries whose one-based rows never decrease:
// position i writes the pointers of the rows in (row[i-1], row[i]].
template <typename I>
__global__ void k_rows_ptr(int64_t nnz, const I* __restrict__ rows, int64_t m,
                           I* __restrict__ ptr) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t prev = -1, cur = m;
        if (i > 0) prev = dclamp((int64_t)rows[i - 1] - 1, 0, m - 1);
        if (i < nnz) cur = dclamp((int64_t)rows[i] - 1, 0, m - 1);
        for (int64_t r = prev + 1; r <= cur; ++r) ptr[r] = (I)(i + 1);
    }
}

// One-based indices from zero-based sorted keys.
template <typename I, typename K>
__global__ void k_keys_to_idx(int64_t nnz, const K* __restrict__ keys, I* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (I)((int64_t)keys[i] + 1);
}

// The column of each entry of one part (its key / m) and its position in the part.
__global__ void k_part_cols(int64_t cnt, const uint64_t* __restrict__ keys, uint64_t m,
                            uint32_t* __restrict__ col, uint32_t* __restrict__ pos) {
    const double inv_m = 1.0 / (double)m;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cnt;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t c, r;
        key_split(keys[i], m, inv_m, c, r);
        col[i] = (uint32_t)c;
        pos[i] = (uint32_t)i;
    }
}

// A part put in the order of `pos` (positions within the part): its keys, and
// each entry's position in the concatenation (offset + pos).
template <typename V>
__global__ void k_part_apply(int64_t cnt, int64_t offset, const uint32_t* __restrict__ pos,
                             const uint64_t* __restrict__ keys, uint64_t* __restrict__ keys_out,
                             V* __restrict__ cat_pos) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cnt;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t p = pos[i];
        keys_out[i] = keys[p];
        cat_pos[i] = (V)(offset + p);
    }
}

// perm[i] = part_pos[merged[i]]: merge positions back to concatenation positions.
template <typename V>
__global__ void k_compose(int64_t tot, const V* __restrict__ part_pos, V* __restrict__ perm) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
         i += (int64_t)gridDim.x * blockDim.x)
        perm[i] = part_pos[(int64_t)perm[i]];
}

// Stable merge of the two sorted parts keys[0, nA) and keys[nA, tot) into
// keys_out / pos_out (pos = index into the concatenation).  Merge path: output
// position d of the merge holds, before it, the first i entries of part A and
// the first d - i of part B, with i found by a binary search on that diagonal;
// on equal keys part A's entry comes first, so the order is that of a stable
// sort of the concatenation.  Each workgroup takes MERGE_TILE outputs: it finds
// its two diagonals, loads the A and B ranges between them into LDS, merges
// MERGE_CH outputs per lane from LDS and writes the tile coalesced.
constexpr int MERGE_CH = 8;
constexpr int MERGE_WG = 256;
constexpr int MERGE_TILE = MERGE_CH * MERGE_WG;

__device__ __forceinline__ int64_t merge_split(const uint64_t* __restrict__ A, int64_t nA,
                                               const uint64_t* __restrict__ B, int64_t nB,
                                               int64_t d) {
    int64_t lo = d - nB > 0 ? d - nB : 0, hi = d < nA ? d : nA;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (A[mid] <= B[d - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The merge-path split of every tile boundary d = t * MERGE_TILE (t = 0..ntiles),
// all searched at once (the tiles then start without a chain of global loads).
__global__ void k_merge_splits(int64_t nA, int64_t tot, const uint64_t* __restrict__ keys,
                               int64_t ntiles, int64_t* __restrict__ split) {
    const int64_t nB = tot - nA;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t <= ntiles;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = t * MERGE_TILE < tot ? t * MERGE_TILE : tot;
        split[t] = merge_split(keys, nA, keys + nA, nB, d);
    }
}

template <typename V>
__global__ __launch_bounds__(MERGE_WG) void k_merge_parts(int64_t nA, int64_t tot,
                                                          const uint64_t* __restrict__ keys,
                                                          const int64_t* __restrict__ split,
                                                          uint64_t* __restrict__ keys_out,
                                                          V* __restrict__ pos_out) {
    __shared__ uint64_t s_in[MERGE_TILE];   // A's range, then B's
    __shared__ uint64_t s_key[MERGE_TILE];
    __shared__ V s_pos[MERGE_TILE];
    const uint64_t* __restrict__ B = keys + nA;
    const int64_t d0 = (int64_t)blockIdx.x * MERGE_TILE;
    if (d0 >= tot) return;
    const int64_t d1 = d0 + MERGE_TILE < tot ? d0 + MERGE_TILE : tot;
    const int64_t a0 = split[blockIdx.x], a1 = split[blockIdx.x + 1];
    const int64_t b0 = d0 - a0, b1 = d1 - a1;
    const int na = (int)(a1 - a0), nb = (int)(b1 - b0);
    for (int t = threadIdx.x; t < na + nb; t += MERGE_WG)
        s_in[t] = t < na ? keys[a0 + t] : B[b0 + (t - na)];
    __syncthreads();
    const uint64_t* sA = s_in;
    const uint64_t* sB = s_in + na;
    const int dl = threadIdx.x * MERGE_CH;
    if (dl < na + nb) {
        int lo = dl - nb > 0 ? dl - nb : 0, hi = dl < na ? dl : na;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sA[mid] <= sB[dl - 1 - mid]) lo = mid + 1;
            else hi = mid;
        }
        int i = lo, j = dl - lo;
        const int e = dl + MERGE_CH < na + nb ? dl + MERGE_CH : na + nb;
        for (int q = dl; q < e; ++q) {
            const bool take_a = i < na && (j >= nb || sA[i] <= sB[j]);
            if (take_a) {
                s_key[q] = sA[i];
                s_pos[q] = (V)(a0 + i);
                ++i;
            } else {
                s_key[q] = sB[j];
                s_pos[q] = (V)(nA + b0 + j);
                ++j;
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < na + nb; t += MERGE_WG) {
        keys_out[d0 + t] = s_key[t];
        pos_out[d0 + t] = s_pos[t];
    }
}

// keys[0, nA) and keys[nA, tot), each sorted: their stable merge into
// keys_out, with each output's position in the concatenation in pos_out.
template <typename V>
static int merge_sorted_parts(hipStream_t s, int64_t nA, int64_t tot, const uint64_t* keys,
                              uint64_t* keys_out, V* pos_out) {
    const int64_t ntiles = cdiv(tot, MERGE_TILE);
    Scratch split(s);
    GSA_TRY(split.alloc(sizeof(int64_t) * (ntiles + 1)));
    hipLaunchKernelGGL(k_merge_splits, dim3(grid_for(ntiles + 1, 256)), dim3(256), 0, s, nA, tot,
                       keys, ntiles, split.as<int64_t>());
    hipLaunchKernelGGL((k_merge_parts<V>), dim3((unsigned)ntiles), dim3(MERGE_WG), 0, s, nA, tot,
                       keys, (const int64_t*)split.as<int64_t>(), keys_out, pos_out);
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

// The concatenation in sorted order: row and column from the sorted key
// (c*m + r), the value gathered through the permutation (GATHER_U positions
// per thread in flight).
template <typename T, typename I, typename V>
__global__ void k_sorted_entries(int64_t tot, const V* __restrict__ perm,
                                 const uint64_t* __restrict__ skeys, uint64_t m, int64_t nA,
                                 const T* __restrict__ vA, const T* __restrict__ vB,
                                 I* __restrict__ ro, I* __restrict__ co, T* __restrict__ vo) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const double inv_m = 1.0 / (double)m;
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; base < tot;
         base += stride * GATHER_U) {
        T v[GATHER_U];
#pragma unroll
        for (int u = 0; u < GATHER_U; ++u) {
            const int64_t i = base + u * stride;
            if (i < tot) {
                const int64_t p = (int64_t)perm[i];
                v[u] = p < nA ? vA[p] : vB[p - nA];
            }
        }
#pragma unroll
        for (int u = 0; u < GATHER_U; ++u) {
            const int64_t i = base + u * stride;
            if (i < tot) {
                uint64_t c, r;
                key_split(skeys[i], m, inv_m, c, r);
                ro[i] = (I)(r + 1);
                co[i] = (I)(c + 1);
                vo[i] = v[u];
            }
        }
    }
}

// Run heads of sorted keys: sum each run left to right starting from its first
// value (kernel_compact_coo!), flag it kept when the sum is nonzero.
template <typename T>
__global__ void k_run_sums(int64_t n, const uint64_t* __restrict__ keys, const T* __restrict__ v,
                           T* __restrict__ sums, int64_t* __restrict__ keep) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        bool head = (i == 0) || keys[i] != keys[i - 1];
        int64_t k = 0;
        if (head) {
            T s = v[i];
            int64_t j = i + 1;
            while (j < n && keys[j] == keys[i]) { s = s + v[j]; ++j; }
            sums[i] = s;
            k = s != T(0);
        }
        keep[i] = k;
    }
}

// ------------------------------------------------------------ kron --------
template <typename T, typename I>
__global__ __launch_bounds__(256) void k_kron_coo(int64_t nnzA, int64_t nnzB, const I* __restrict__ rA,
                                                  const I* __restrict__ cA, const T* __restrict__ vA,
                                                  const I* __restrict__ rB, const I* __restrict__ cB,
                                                  const T* __restrict__ vB, int64_t mB, int64_t nB,
                                                  I* __restrict__ rC, I* __restrict__ cC,
                                                  T* __restrict__ vC) {
    for (int64_t a = blockIdx.x; a < nnzA; a += gridDim.x) {
        const int64_t ra = (int64_t)rA[a] - 1, ca = (int64_t)cA[a] - 1;
        const T va = vA[a];
        const int64_t base = a * nnzB;
        for (int64_t b = threadIdx.x; b < nnzB; b += blockDim.x) {
            rC[base + b] = (I)(ra * mB + (int64_t)rB[b]);
            cC[base + b] = (I)(ca * nB + (int64_t)cB[b]);
            vC[base + b] = va * vB[b];
        }
    }
}

// flag[0] = 1 if some row has a non-increasing column pair
template <typename I>
__global__ void k_check_canonical(int64_t m, const I* __restrict__ p, const I* __restrict__ c,
                                  int* __restrict__ flag) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        for (int64_t j = (int64_t)p[r]; j < (int64_t)p[r + 1] - 1; ++j)
            if (c[j] <= c[j - 1]) { *flag = 1; break; }
    }
}

template <typename I>
__global__ void k_kron_rowptr(int64_t mA, int64_t mB, int64_t nnzB, const I* __restrict__ pA,
                              const I* __restrict__ pB, I* __restrict__ pC) {
    const int64_t total = mA * mB;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= total;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i == total) {
            pC[i] = (I)((int64_t)(pA[mA] - 1) * nnzB + 1);
            continue;
        }
        int64_t iA = i / mB, iB = i - iA * mB;
        int64_t lenA = (int64_t)pA[iA + 1] - pA[iA];
        pC[i] = (I)(((int64_t)pA[iA] - 1) * nnzB + lenA * ((int64_t)pB[iB] - 1) + 1);
    }
}

// One wavefront per output row (iA, iB): its lenA*lenB entries are contiguous
// at (pA[iA]-1)*nnzB + lenA*(pB[iB]-1), listed jA-major, jB-minor.
template <typename T, typename I>
__global__ __launch_bounds__(256) void k_kron_csr_fill(
    int64_t mA, int64_t mB, int64_t nB, int64_t nnzB, const I* __restrict__ pA,
    const I* __restrict__ cA, const T* __restrict__ vA, const I* __restrict__ pB,
    const I* __restrict__ cB, const T* __restrict__ vB, I* __restrict__ cC, T* __restrict__ vC) {
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t rows = mA * mB;
    for (int64_t i = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; i < rows;
         i += nwaves) {
        const int64_t iA = i / mB, iB = i - iA * mB;
        const int64_t a0 = (int64_t)pA[iA] - 1, lenA = (int64_t)pA[iA + 1] - 1 - a0;
        const int64_t b0 = (int64_t)pB[iB] - 1, lenB = (int64_t)pB[iB + 1] - 1 - b0;
        if (lenA <= 0 || lenB <= 0) continue;
        const int64_t cnt = lenA * lenB;
        const int64_t base = a0 * nnzB + lenA * b0;
        auto put = [&](int64_t o, int64_t aL, int64_t bL) {
            const int64_t a = a0 + aL, b = b0 + bL;
            cC[base + o] = (I)(((int64_t)cA[a] - 1) * nB + (int64_t)cB[b]);
            vC[base + o] = vA[a] * vB[b];
        };
        if (cnt <= (int64_t)UINT32_MAX) {
            const uint32_t lb = (uint32_t)lenB;
            for (uint32_t o = lane; o < (uint32_t)cnt; o += 64) {
                const uint32_t aL = o / lb;
                put(o, aL, o - aL * lb);
            }
        } else {
            for (int64_t o = lane; o < cnt; o += 64) {
                const int64_t aL = o / lenB;
                put(o, aL, o - aL * lenB);
            }
        }
    }
}

// ------------------------------------------------------------ SpGEMM ------

template <typename T, typename I, typename V>
__global__ void k_sg_expand(int64_t m, int64_t n, int64_t kdim, const I* __restrict__ pA,
                            const I* __restrict__ cA, const T* __restrict__ vA,
                            const I* __restrict__ pB, const I* __restrict__ cB,
                            const T* __restrict__ vB, const int64_t* __restrict__ pptr,
                            uint64_t* __restrict__ keys, T* __restrict__ prods,
                            V* __restrict__ pos) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        int64_t w = pptr[r] - 1;
        for (int64_t a = (int64_t)pA[r] - 1; a < (int64_t)pA[r + 1] - 1; ++a) {
            int64_t k = dclamp((int64_t)cA[a] - 1, 0, kdim - 1);
            T va = vA[a];
            for (int64_t b = (int64_t)pB[k] - 1; b < (int64_t)pB[k + 1] - 1; ++b) {
                int64_t c = dclamp((int64_t)cB[b] - 1, 0, n - 1);
                keys[w] = (uint64_t)r * (uint64_t)n + (uint64_t)c;
                prods[w] = va * vB[b];
                pos[w] = (V)w;
                ++w;
            }
        }
    }
}

template <typename T, typename V>
__global__ void k_gather_vals(int64_t n, const V* __restrict__ perm, const T* __restrict__ in,
                              T* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; base < n;
         base += stride * GATHER_U) {
        T v[GATHER_U];
#pragma unroll
        for (int u = 0; u < GATHER_U; ++u) {
            const int64_t i = base + u * stride;
            if (i < n) v[u] = in[(int64_t)perm[i]];
        }
#pragma unroll
        for (int u = 0; u < GATHER_U; ++u) {
            const int64_t i = base + u * stride;
            if (i < n) out[i] = v[u];
        }
    }
}

template <typename T>
__global__ void k_sg_heads(int64_t P, const uint64_t* __restrict__ keys, int64_t* __restrict__ f) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P;
         i += (int64_t)gridDim.x * blockDim.x)
        f[i] = (i == 0) || keys[i] != keys[i - 1];
}

// per output entry: sum its run from +0.0 in generation order; per-row counts
template <typename T, typename I>
__global__ void k_sg_sum(int64_t P, int64_t n, const uint64_t* __restrict__ keys,
                         const T* __restrict__ v, const int64_t* __restrict__ incl,
                         uint64_t* __restrict__ okeys, T* __restrict__ ovals,
                         I* __restrict__ rowcnt /* ptr+1 */) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i != 0 && keys[i] == keys[i - 1]) continue;
        T s = T(0);
        int64_t j = i;
        while (j < P && keys[j] == keys[i]) { s = s + v[j]; ++j; }
        int64_t o = incl[i] - 1;
        okeys[o] = keys[i];
        ovals[o] = s;
        atomic_inc(&rowcnt[keys[i] / (uint64_t)n]);
    }
}

template <typename T, typename I>
__global__ void k_sg_write(int64_t nnzC, int64_t n, const uint64_t* __restrict__ okeys,
                           const T* __restrict__ ovals, I* __restrict__ cC, T* __restrict__ vC) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnzC;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t k = okeys[i];
        cC[i] = (I)(k % (uint64_t)n + 1);
        vC[i] = ovals[i];
    }
}

// ---- row-local SpGEMM ------------------------------------------------------
// Two phases over the same per-row work.  Output row r's products are
// generated in the reference's order (A-row entries in order, each B row in
// order: kernel_spgemm_csr!, matrix_csr_kernels.jl:293-343) and numbered p;
// the key (column << 24 | p) is sorted inside a wavefront (rows of at most
// SG_WAVE_P products, bitonic across lanes) or a workgroup (at most
// SG_BLOCK_P, bitonic in LDS).  Equal columns then lie together in generation
// order, so each output entry is the left-to-right sum from +0.0 of its
// products, cancellation zeros kept.  The symbolic phase counts the distinct
// columns of every row; the numeric phase writes columns and values at the row
// pointer.  A product with a row over SG_BLOCK_P products uses the global
// expand-sort-compress path below instead.
constexpr int SG_WAVE_P = 64;
constexpr int SG_BLOCK_P = 2048;
constexpr int SG_WAVES = 4;
constexpr uint64_t SG_NONE = ~0ull;

__device__ __forceinline__ uint64_t sg_key(int64_t col, int p) {
    return ((uint64_t)col << 24) | (uint64_t)p;
}
__device__ __forceinline__ int64_t sg_col(uint64_t k) { return (int64_t)(k >> 24); }


// Products of output row r: one-based exclusive prefix pp (m+1 entries).
__device__ __forceinline__ int64_t sg_products(const int64_t* __restrict__ pp, int64_t r) {
    return pp[r + 1] - pp[r];
}

// Rows with at most G products (and more than G/2, or any for the smallest
// group): groups of G lanes, 64/G rows per wavefront.  Scans, the bitonic sort
// and the run sums stay inside a group.
// STAGE (with NUM): one pass that also counts.  Row r's distinct columns and
// sums go to a staging area at its product offset pp[r] - 1 (the row's output
// is never longer than its products), and rowcnt[r + 1] gets their number;
// k_sg_compact later moves them to their final place.
template <bool NUM, int G, typename T, typename I, bool STAGE = false>
__global__ __launch_bounds__(64 * SG_WAVES) void k_sg_row_wave(
    int64_t m, int64_t n, int64_t kdim, int64_t nnzA, int64_t nnzB, const I* __restrict__ pA,
    const I* __restrict__ cA, const T* __restrict__ vA, const I* __restrict__ pB,
    const I* __restrict__ cB, const T* __restrict__ vB, const int64_t* __restrict__ pp,
    int min_p, I* __restrict__ rowcnt, const I* __restrict__ rowptrC, I* __restrict__ cC,
    T* __restrict__ vC) {
    constexpr int RPW = 64 / G;  // rows per wavefront
    __shared__ T s_val[SG_WAVES][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int gl = lane & (G - 1);        // lane inside the group
    const int gbase = lane & ~(G - 1);    // first lane of the group
    const int64_t r = ((int64_t)blockIdx.x * SG_WAVES + wv) * RPW + (lane / G);
    // a group with no row of its size class sits out the collective steps
    int64_t P = 0;
    if (r < m) P = sg_products(pp, r);
    const bool mine = r < m && P > min_p && P <= G;
    if (__ballot(mine) == 0) return;
    uint64_t key = SG_NONE;
    T val = T(0);
    if (mine) {
        const int64_t a0 = dclamp((int64_t)pA[r] - 1, 0, nnzA);
        const int64_t a1 = dclamp((int64_t)pA[r + 1] - 1, a0, nnzA);
        int base = 0;  // products generated by earlier chunks of A entries
        for (int64_t ac = a0; ac < a1 && base < P; ac += G) {
            const int64_t a = ac + gl;
            const bool va = a < a1;
            int64_t bs = 0, bl = 0;
            if (va) {
                const int64_t k = dclamp((int64_t)cA[a] - 1, 0, kdim - 1);
                bs = dclamp((int64_t)pB[k] - 1, 0, nnzB);
                bl = dclamp((int64_t)pB[k + 1] - 1, bs, nnzB) - bs;
            }
            // inclusive sum over the group (every source lane lies in the group, and
            // a group is active or inactive as a whole)
            const int incl = group_incl_sum<G>((int)bl);
            const int chunkP = __shfl(incl, G - 1, G);
            const int q = gl - base;  // this lane's product inside the chunk
            int lo = 0;
#pragma unroll
            for (int step = G / 2; step >= 1; step >>= 1) {
                const int im = __shfl(incl, lo + step - 1, G);
                if (im <= q) lo += step;
            }
            const int e = lo < G - 1 ? lo : G - 1;
            const int64_t ea = __shfl(a, e, G);
            const int64_t ebs = __shfl(bs, e, G);
            const int eex = __shfl(incl, e, G) - (int)__shfl(bl, e, G);
            if (q >= 0 && q < chunkP) {
                const int64_t b = dclamp(ebs + (q - eex), 0, nnzB - 1);
                const int64_t col = dclamp((int64_t)cB[b] - 1, 0, n - 1);
                key = sg_key(col, gl);
                if (NUM) val = vA[dclamp(ea, 0, nnzA - 1)] * vB[b];
            }
            base += chunkP;
        }
    }
    // bitonic sort of (key, val) inside each group, ascending
#pragma unroll
    for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t ok = lane_xor(key, j);
            const T ov = lane_xor(val, j);
            const bool up = (gl & k) == 0;
            const bool lower = (gl & j) == 0;
            const bool take = (lower == up) ? (ok < key) : (ok > key);
            if (take) { key = ok; val = ov; }
        }
    }
    const bool valid = mine && gl < P;
    const int64_t col = sg_col(key);
    const int64_t pcol = __shfl_up(col, 1, G);
    const bool head = valid && (gl == 0 || pcol != col);
    const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << gbase);
    const uint64_t heads = __ballot(head) & gmask;
    if (!NUM || STAGE) {
        if (mine && gl == 0) rowcnt[r + 1] = (I)__popcll(heads);
        if (!NUM) return;
    }
    s_val[wv][lane] = val;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (head) {
        const uint64_t later = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
        const int stop = later ? (__ffsll((long long)later) - 1) : gbase + (int)P;
        T acc = T(0);
        for (int j = lane; j < stop; ++j) acc = acc + s_val[wv][j];
        const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const int64_t o = (STAGE ? pp[r] : (int64_t)rowptrC[r]) - 1 + __popcll(heads & lt);
        cC[o] = (I)(col + 1);
        vC[o] = acc;
    }
}

// Staged rows (k_sg_row_wave / k_sg_row_block with STAGE) to their final place:
// row r's entries from offset pp[r] - 1 to rowptrC[r] - 1, 16 lanes per row.
template <typename T, typename I>
__global__ void k_sg_compact(int64_t m, const int64_t* __restrict__ pp,
                             const I* __restrict__ rowptrC, const I* __restrict__ sc,
                             const T* __restrict__ sv, I* __restrict__ cC, T* __restrict__ vC) {
    const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m * 16; t += nthreads) {
        const int64_t r = t >> 4;
        const int gl = (int)(t & 15);
        const int64_t src = pp[r] - 1;
        const int64_t dst = (int64_t)rowptrC[r] - 1;
        const int64_t cnt = (int64_t)rowptrC[r + 1] - 1 - dst;
        for (int64_t j = gl; j < cnt; j += 16) {
            cC[dst + j] = sc[src + j];
            vC[dst + j] = sv[src + j];
        }
    }
}

// Launch the grouped row kernels for the size classes (0,16], (16,32], (32,64].
// Launch the classes whose bit is set in `classes` (k_sg_count_flags): each launch
// covers every row, so an empty class would still cost a wave per row group.
template <bool NUM, typename T, typename I, bool STAGE = false>
static void sg_rows_small(hipStream_t s, int64_t m, int64_t n, int64_t kdim, int64_t nnzA,
                          int64_t nnzB, const I* pA, const I* cA, const T* vA, const I* pB,
                          const I* cB, const T* vB, const int64_t* pp, I* rowcnt,
                          const I* rowptrC, I* cC, T* vC, unsigned classes) {
    if (classes & 1u)
        hipLaunchKernelGGL((k_sg_row_wave<NUM, 16, T, I, STAGE>), dim3((unsigned)cdiv(m, SG_WAVES * 4)),
                           dim3(64 * SG_WAVES), 0, s, m, n, kdim, nnzA, nnzB, pA, cA, vA, pB, cB,
                           vB, pp, 0, rowcnt, rowptrC, cC, vC);
    if (classes & 2u)
        hipLaunchKernelGGL((k_sg_row_wave<NUM, 32, T, I, STAGE>), dim3((unsigned)cdiv(m, SG_WAVES * 2)),
                           dim3(64 * SG_WAVES), 0, s, m, n, kdim, nnzA, nnzB, pA, cA, vA, pB, cB,
                           vB, pp, 16, rowcnt, rowptrC, cC, vC);
    if (classes & 4u)
        hipLaunchKernelGGL((k_sg_row_wave<NUM, 64, T, I, STAGE>), dim3((unsigned)cdiv(m, SG_WAVES)),
                           dim3(64 * SG_WAVES), 0, s, m, n, kdim, nnzA, nnzB, pA, cA, vA, pB, cB,
                           vB, pp, 32, rowcnt, rowptrC, cC, vC);
}

template <bool NUM, typename T, typename I, bool STAGE = false>
__global__ __launch_bounds__(ROW_WG) void k_sg_row_block(
    int64_t n, int64_t kdim, int64_t nnzA, int64_t nnzB, const I* __restrict__ pA,
    const I* __restrict__ cA, const T* __restrict__ vA, const I* __restrict__ pB,
    const I* __restrict__ cB, const T* __restrict__ vB, const int64_t* __restrict__ pp,
    const int64_t* __restrict__ rows, I* __restrict__ rowcnt, const I* __restrict__ rowptrC,
    I* __restrict__ cC, T* __restrict__ vC) {
    __shared__ uint64_t sk[SG_BLOCK_P];
    __shared__ T sv[SG_BLOCK_P];
    __shared__ int64_t lds4[ROW_WG / 64];
    const int tid = threadIdx.x;
    const int64_t r = rows[blockIdx.x];
    const int P = (int)sg_products(pp, r);  // 64 < P <= SG_BLOCK_P (the row list holds these)
    int n2 = 128;
    while (n2 < P) n2 <<= 1;
    const int64_t a0 = dclamp((int64_t)pA[r] - 1, 0, nnzA);
    const int64_t a1 = dclamp((int64_t)pA[r + 1] - 1, a0, nnzA);
    // 1. products in generation order: A entries in chunks of ROW_WG threads
    int64_t base = 0;
    for (int64_t ac = a0; ac < a1; ac += ROW_WG) {
        const int64_t a = ac + tid;
        int64_t bs = 0, bl = 0;
        if (a < a1) {
            const int64_t k = dclamp((int64_t)cA[a] - 1, 0, kdim - 1);
            bs = dclamp((int64_t)pB[k] - 1, 0, nnzB);
            bl = dclamp((int64_t)pB[k + 1] - 1, bs, nnzB) - bs;
        }
        int64_t chunk = 0;
        const int64_t ex = wg_exclusive_sum<int64_t>(bl, lds4, &chunk);
        if (a < a1) {
            const T va = NUM ? vA[a] : T(0);
            for (int64_t j = 0; j < bl; ++j) {
                const int64_t p = base + ex + j;
                if (p >= P) break;
                const int64_t col = dclamp((int64_t)cB[bs + j] - 1, 0, n - 1);
                sk[p] = sg_key(col, (int)p);
                if (NUM) sv[p] = va * vB[bs + j];
            }
        }
        base += chunk;
    }
    for (int p = P + tid; p < n2; p += ROW_WG) sk[p] = SG_NONE;
    __syncthreads();
    // 2. bitonic sort in LDS
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < n2 / 2; t += ROW_WG) {
                const int i = 2 * t - (t & (j - 1));
                const int l = i + j;
                const bool asc = (i & k) == 0;
                const uint64_t ki = sk[i], kl = sk[l];
                if ((ki > kl) == asc) {
                    sk[i] = kl; sk[l] = ki;
                    if (NUM) { const T t2 = sv[i]; sv[i] = sv[l]; sv[l] = t2; }
                }
            }
            __syncthreads();
        }
    }
    // 3. heads: thread tid owns the contiguous elements [tid*per, (tid+1)*per)
    const int per = SG_BLOCK_P / ROW_WG;
    int64_t mine = 0;
    for (int q = 0; q < per; ++q) {
        const int i = tid * per + q;
        if (i < P && (i == 0 || sg_col(sk[i]) != sg_col(sk[i - 1]))) ++mine;
    }
    int64_t total = 0;
    int64_t rank = wg_exclusive_sum<int64_t>(mine, lds4, &total);
    if (!NUM || STAGE) {
        if (tid == 0) rowcnt[r + 1] = (I)total;
        if (!NUM) return;
    }
    const int64_t off = (STAGE ? pp[r] : (int64_t)rowptrC[r]) - 1;
    for (int q = 0; q < per; ++q) {
        const int i = tid * per + q;
        if (i >= P) break;
        const int64_t col = sg_col(sk[i]);
        if (i != 0 && col == sg_col(sk[i - 1])) continue;
        T acc = T(0);
        for (int j = i; j < P && sg_col(sk[j]) == col; ++j) acc = acc + sv[j];
        cC[off + rank] = (I)(col + 1);
        vC[off + rank] = acc;
        ++rank;
    }
}

// Products of each row (cnt[r + 1]), the flags of rows over SG_WAVE_P products
// (f[r], for the row list of the workgroup kernel), out[0] the largest count
// and out[1] a bit per size class of the grouped row kernels that holds a row:
// 1 for (0,16], 2 for (16,32], 4 for (32,64].
template <typename I>
__global__ void k_sg_count_flags(int64_t m, const I* __restrict__ pA, const I* __restrict__ cA,
                                 const I* __restrict__ pB, int64_t kdim,
                                 int64_t* __restrict__ cnt, int64_t* __restrict__ f,
                                 unsigned long long* __restrict__ out) {
    __shared__ unsigned long long bmax;
    __shared__ unsigned int bcls;
    if (threadIdx.x == 0) {
        bmax = 0;
        bcls = 0;
    }
    __syncthreads();
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t P = 0;
    if (r < m) {
        for (int64_t a = (int64_t)pA[r] - 1; a < (int64_t)pA[r + 1] - 1; ++a) {
            const int64_t k = dclamp((int64_t)cA[a] - 1, 0, kdim - 1);
            P += (int64_t)pB[k + 1] - pB[k];
        }
        cnt[r + 1] = P;
        f[r] = P > SG_WAVE_P;
    }
    // the wavefront's largest count and size classes first (every lane takes
    // part), then one LDS atomic per wavefront: 64 lanes contending for one LDS
    // word serialise
    uint64_t pm = (uint64_t)P;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint64_t o = lane_xor(pm, off);
        pm = o > pm ? o : pm;
    }
    const unsigned cls = P <= 0 ? 0u : (P <= 16 ? 1u : (P <= 32 ? 2u : (P <= 64 ? 4u : 0u)));
    const unsigned wcls = (__ballot(cls & 1u) ? 1u : 0u) | (__ballot(cls & 2u) ? 2u : 0u) |
                          (__ballot(cls & 4u) ? 4u : 0u);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&bmax, (unsigned long long)pm);
        if (wcls) atomicOr(&bcls, wcls);
    }
    __syncthreads();
    // device-wide atomics on one address serialise across the whole grid: a block
    // adds its figures only when they raise what is already there (a stale read
    // is never above the true value, so it can only cost an unneeded atomic)
    if (threadIdx.x == 0) {
        if (bmax > *(volatile unsigned long long*)out) atomicMax(out, bmax);
        if (bcls & ~*(volatile unsigned long long*)(out + 1))
            atomicOr(out + 1, (unsigned long long)bcls);
    }
}

// The four figures the begin phase reads back, side by side for one copy:
// largest count, long rows, size classes, product-pointer end.
__global__ void k_sg_figures(const unsigned long long* __restrict__ mx,
                             const int64_t* __restrict__ long_incl, int64_t m,
                             const int64_t* __restrict__ pp, int64_t* __restrict__ out) {
    if (threadIdx.x != 0) return;
    out[0] = (int64_t)mx[0];
    out[1] = m > 0 ? long_incl[m - 1] : 0;
    out[2] = (int64_t)mx[1];
    out[3] = pp[m];
}

__global__ void k_sg_list(int64_t m, const int64_t* __restrict__ pp,
                          const int64_t* __restrict__ incl, int64_t* __restrict__ rows) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r < m && sg_products(pp, r) > SG_WAVE_P) rows[incl[r] - 1] = r;
}

template <typename T, typename I>
static int run_kron_coo(hipStream_t s, int64_t mB, int64_t nB, int64_t nnzA, const I* rA,
                        const I* cA, const T* vA, int64_t nnzB, const I* rB, const I* cB,
                        const T* vB, I* rC, I* cC, T* vC) {
    if (nnzA == 0 || nnzB == 0) return GSA_OK;
    unsigned g = (unsigned)(nnzA < 65536 ? nnzA : 65536);
    hipLaunchKernelGGL((k_kron_coo<T, I>), dim3(g), dim3(256), 0, s, nnzA, nnzB, rA, cA, vA, rB, cB,
                       vB, mB, nB, rC, cC, vC);
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

}  // namespace gsa

using namespace gsa;

extern "C" int gsa_coo_to_csr(void*, int, int, int64_t, int64_t, int64_t, const void*, const void*,
                              const void*, void*, void*, void*);

extern "C" {

int gsa_csr_dropzeros_count(void* stream, int vtype, int itype, int64_t m, const void* rowptr,
                            const void* nzval, void* rowptr_out, int64_t* nnz_out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(m >= 0 && nnz_out, GSA_ERR_INVALID, "gsa_csr_dropzeros_count: bad arguments");
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        I last = 1;
        if (m > 0) {
            GSA_HIP(hipMemcpyAsync(&last, (const I*)rowptr + m, sizeof(I), hipMemcpyDeviceToHost, s));
            GSA_HIP(hipStreamSynchronize(s));
        }
        const int64_t nnz = (int64_t)last - 1;
        if (m == 0 || nnz <= 0) {
            GSA_TRY(zero_fill(s, rowptr_out, sizeof(I) * (m + 1)));
            return counts_to_ptr<I>(s, (I*)rowptr_out, m, nnz_out);
        }
        const int64_t nblocks = nnz / SPMV_NB + 1;
        Scratch plan(s), incl(s);
        GSA_TRY(plan.alloc(sizeof(int64_t) * (nblocks + 1)));
        GSA_TRY(incl.alloc(sizeof(int64_t) * nblocks));
        GSA_TRY(build_csr_plan(s, itype, m, nnz, rowptr, plan.as<int64_t>(), nblocks));
        GSA_TRY(dz_block_prefix<T>(s, nnz, (const T*)nzval, nblocks, incl.as<int64_t>()));
        hipLaunchKernelGGL((k_dz_rowptr<T, I>), dim3((unsigned)nblocks), dim3(ROW_WG), 0, s, m, nnz,
                           (const I*)rowptr, (const T*)nzval, (const int64_t*)plan.as<int64_t>(),
                           (const int64_t*)incl.as<int64_t>(), (I*)rowptr_out);
        GSA_HIP(hipGetLastError());
        int64_t kept = 0;
        GSA_HIP(hipMemcpyAsync(&kept, incl.as<int64_t>() + nblocks - 1, sizeof kept,
                               hipMemcpyDeviceToHost, s));
        GSA_HIP(hipStreamSynchronize(s));
        *nnz_out = kept;
        return (int)GSA_OK;
    });
}

int gsa_csr_dropzeros_fill(void* stream, int vtype, int itype, int64_t m, const void* rowptr,
                           const void* colval, const void* nzval, const void* rowptr_out,
                           void* colval_out, void* nzval_out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(m >= 0, GSA_ERR_INVALID, "gsa_csr_dropzeros_fill: bad arguments");
    (void)rowptr_out;  // each kept entry's place follows from the entries before it
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        if (m == 0) return (int)GSA_OK;
        I last = 1;
        GSA_HIP(hipMemcpyAsync(&last, (const I*)rowptr + m, sizeof(I), hipMemcpyDeviceToHost, s));
        GSA_HIP(hipStreamSynchronize(s));
        const int64_t nnz = (int64_t)last - 1;
        if (nnz <= 0) return (int)GSA_OK;
        const int64_t nblocks = nnz / SPMV_NB + 1;
        Scratch incl(s);
        GSA_TRY(incl.alloc(sizeof(int64_t) * nblocks));
        GSA_TRY(dz_block_prefix<T>(s, nnz, (const T*)nzval, nblocks, incl.as<int64_t>()));
        hipLaunchKernelGGL((k_dz_compact<T, I>), dim3((unsigned)nblocks), dim3(ROW_WG), 0, s, nnz,
                           (const I*)colval, (const T*)nzval, (const int64_t*)incl.as<int64_t>(),
                           (I*)colval_out, (T*)nzval_out);
        GSA_HIP(hipGetLastError());
        return (int)GSA_OK;
    });
}

int gsa_coo_dropzeros_count(void* stream, int vtype, int64_t nnz, const void* nzval,
                            int64_t* nnz_out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(nnz >= 0 && nnz_out, GSA_ERR_INVALID, "gsa_coo_dropzeros_count: bad arguments");
    hipStream_t s = as_stream(stream);
    return dispatch_va(vtype, [&](auto tv) {
        using T = decltype(tv);
        *nnz_out = 0;
        if (nnz == 0) return (int)GSA_OK;
        Scratch f(s);
        GSA_TRY(f.alloc(sizeof(int64_t) * nnz));
        hipLaunchKernelGGL(k_nonzero_flags<T>, dim3(grid_for(nnz, 256)), dim3(256), 0, s, nnz,
                           (const T*)nzval, f.as<int64_t>());
        GSA_TRY(inclusive_scan<int64_t>(s, f.as<int64_t>(), f.as<int64_t>(), nnz, 0));
        int64_t tot = 0;
        GSA_HIP(hipMemcpyAsync(&tot, f.as<int64_t>() + nnz - 1, sizeof(int64_t),
                               hipMemcpyDeviceToHost, s));
        GSA_HIP(hipStreamSynchronize(s));
        *nnz_out = tot;
        return (int)GSA_OK;
    });
}

int gsa_coo_dropzeros_fill(void* stream, int vtype, int itype, int64_t nnz, const void* rowind,
                           const void* colind, const void* nzval, void* rowind_out,
                           void* colind_out, void* nzval_out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(nnz >= 0, GSA_ERR_INVALID, "gsa_coo_dropzeros_fill: bad arguments");
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        if (nnz == 0) return (int)GSA_OK;
        Scratch f(s);
        GSA_TRY(f.alloc(sizeof(int64_t) * nnz));
        hipLaunchKernelGGL(k_nonzero_flags<T>, dim3(grid_for(nnz, 256)), dim3(256), 0, s, nnz,
                           (const T*)nzval, f.as<int64_t>());
        GSA_TRY(inclusive_scan<int64_t>(s, f.as<int64_t>(), f.as<int64_t>(), nnz, 0));
        hipLaunchKernelGGL((k_compact<T, I>), dim3(grid_for(nnz, 256)), dim3(256), 0, s, nnz,
                           (const int64_t*)f.as<int64_t>(), (const I*)rowind, (const I*)colind,
                           (const T*)nzval, (I*)rowind_out, (I*)colind_out, (T*)nzval_out);
        GSA_HIP(hipGetLastError());
        return (int)GSA_OK;
    });
}

int gsa_csr_add_count(void* stream, int vtype, int itype, int64_t m, int64_t nnzA,
                      const void* rowptrA, const void* colA, const void* valA, int64_t nnzB,
                      const void* rowptrB, const void* colB, const void* valB, void* rowptrC,
                      int64_t* nnzC) {
    GSA_TRY(ensure_init());
    GSA_CHECK(m >= 0 && nnzA >= 0 && nnzB >= 0 && nnzC, GSA_ERR_INVALID,
              "gsa_csr_add_count: bad arguments");
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        GSA_TRY((merge_rows<T, I, false, false>(s, m, 0, nnzA, (const I*)rowptrA, (const I*)colA,
                                                 (const T*)valA, nnzB, (const I*)rowptrB,
                                                 (const I*)colB, (const T*)valB, (I*)rowptrC,
                                                 (I*)nullptr, (T*)nullptr)));
        return counts_to_ptr<I>(s, (I*)rowptrC, m, nnzC);
    });
}

int gsa_csr_add_fill(void* stream, int vtype, int itype, int64_t m, int64_t nnzA,
                     const void* rowptrA, const void* colA, const void* valA, int64_t nnzB,
                     const void* rowptrB, const void* colB, const void* valB,
                     const void* rowptrC, void* colC, void* valC) {
    GSA_TRY(ensure_init());
    GSA_CHECK(m >= 0 && nnzA >= 0 && nnzB >= 0, GSA_ERR_INVALID,
              "gsa_csr_add_fill: bad arguments");
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        return merge_rows<T, I, false, true>(s, m, 0, nnzA, (const I*)rowptrA, (const I*)colA,
                                             (const T*)valA, nnzB, (const I*)rowptrB,
                                             (const I*)colB, (const T*)valB, (I*)rowptrC,
                                             (I*)colC, (T*)valC);
    });
}

int gsa_coo_add_begin(void* stream, int vtype, int itype, int64_t m, int64_t n, int64_t nnzA,
                      const void* rowA, const void* colA, const void* valA, int64_t nnzB,
                      const void* rowB, const void* colB, const void* valB, gsa_ctx* ctx,
                      int64_t* nnzC) {
    GSA_TRY(ensure_init());
    GSA_CHECK(ctx && nnzC && m >= 0 && n >= 0 && nnzA >= 0 && nnzB >= 0, GSA_ERR_INVALID,
              "gsa_coo_add_begin: bad arguments");
    const int64_t tot = nnzA + nnzB;
    GSA_CHECK(tot == 0 || (m > 0 && n > 0), GSA_ERR_INVALID,
              "stored entries in a matrix with a zero dimension");
    GSA_CHECK(m == 0 || n == 0 || (uint64_t)m <= UINT64_MAX / (uint64_t)n, GSA_ERR_INVALID,
              "gsa_coo_add_begin: m*n overflows 64-bit keys");
    hipStream_t s = as_stream(stream);
    *ctx = nullptr;
    *nnzC = 0;
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        auto* c = new gsa_ctx_s();
        c->kind = 1; c->stream = s; c->vtype = vtype; c->itype = itype; c->m = m; c->n = n;
        c->aux[0] = tot;
        auto fail = [&](int r) { gsa_ctx_destroy(c); return r; };
        if (tot > 0) {
            Scratch keys(s), perm(s), flag(s);
            int r;
            // one check of both operands' order decides the path (synchronises)
            if ((r = flag.alloc(sizeof(int))) != GSA_OK) return fail(r);
            if ((r = zero_fill(s, flag.p, sizeof(int))) != GSA_OK) return fail(r);
            hipLaunchKernelGGL((k_parts_order<I>), dim3(grid_for(tot, 256)), dim3(256), 0, s, nnzA,
                               tot, (const I*)rowA, (const I*)colA, (const I*)rowB, (const I*)colB,
                               m, n, flag.as<int>());
            int fl = 15;
            hipError_t e0 = hipMemcpyAsync(&fl, flag.p, sizeof(int), hipMemcpyDeviceToHost, s);
            if (e0 == hipSuccess) e0 = hipStreamSynchronize(s);
            if (e0 != hipSuccess) return fail(hip_fail(e0, "gsa_coo_add_begin"));
            // each operand's storage order: by (col, row), by (row, col)
            // The merge paths build a pointer of the result in the index type; with
            // 32-bit indices and 2^31 - 1 entries or more it could not hold the
            // total, so such sums take the sort path (64-bit keys, no pointer).
            const bool ptr_fits = sizeof(I) == 8 || tot < (int64_t)INT32_MAX;
            const bool cmA = !(fl & 1) && ptr_fits, cmB = !(fl & 4) && ptr_fits;
            const bool rmA = !(fl & 2) && ptr_fits, rmB = !(fl & 8) && ptr_fits;
            if (cmA && cmB) {
                // both operands in (col, row) order (COO from a CSC or a SparseMatrixCSC):
                // the row merge with the roles of rows and columns exchanged gives the
                // sum in column-major order directly.  Here the column pointers and the
                // counts; the finish fills the caller's arrays (the operands are read
                // again there) and expands the column pointer into their column indices.
                for (int b = 0; b < 3; ++b)
                    if ((r = ws_alloc(&c->buf[b], sizeof(I) * (n + 1), s)) != GSA_OK) return fail(r);
                I* qA = (I*)c->buf[1];
                I* qB = (I*)c->buf[2];
                I* ptrC = (I*)c->buf[0];
                hipLaunchKernelGGL((k_rows_ptr<I>), dim3(grid_for(nnzA + 1, 256)), dim3(256), 0, s,
                                   nnzA, (const I*)colA, n, qA);
                hipLaunchKernelGGL((k_rows_ptr<I>), dim3(grid_for(nnzB + 1, 256)), dim3(256), 0, s,
                                   nnzB, (const I*)colB, n, qB);
                if ((r = merge_rows<T, I, true, false>(s, n, m, nnzA, (const I*)qA, (const I*)rowA,
                                                       (const T*)valA, nnzB, (const I*)qB,
                                                       (const I*)rowB, (const T*)valB, ptrC,
                                                       (I*)nullptr, (T*)nullptr)) != GSA_OK)
                    return fail(r);
                int64_t total = 0;
                if ((r = counts_to_ptr<I>(s, ptrC, n, &total)) != GSA_OK) return fail(r);
                c->in[0] = rowA; c->in[1] = valA; c->in[2] = rowB; c->in[3] = valB;
                c->aux[1] = 2;  // column merge: the finish fills
                c->aux[2] = nnzA;
                c->nnz_out = total;
                *nnzC = total;
                *ctx = c;
                return (int)GSA_OK;
            }
            if (rmA && rmB) {
                // both operands in (row, col) order (COO made from a CSR): a merge of
                // each row's entries, then (in the finish) one stable sort by column
                // that carries each entry's (row, value) into the caller's arrays
                Scratch pA(s), pB(s);
                if ((r = pA.alloc(sizeof(I) * (m + 1))) != GSA_OK) return fail(r);
                if ((r = pB.alloc(sizeof(I) * (m + 1))) != GSA_OK) return fail(r);
                hipLaunchKernelGGL((k_rows_ptr<I>), dim3(grid_for(nnzA + 1, 256)), dim3(256), 0, s,
                                   nnzA, (const I*)rowA, m, pA.as<I>());
                hipLaunchKernelGGL((k_rows_ptr<I>), dim3(grid_for(nnzB + 1, 256)), dim3(256), 0, s,
                                   nnzB, (const I*)rowB, m, pB.as<I>());
                if ((r = ws_alloc(&c->buf[0], sizeof(I) * (m + 1), s)) != GSA_OK) return fail(r);
                I* ptrC = (I*)c->buf[0];
                if ((r = merge_rows<T, I, true, false>(s, m, n, nnzA, (const I*)pA.as<I>(),
                                                       (const I*)colA, (const T*)valA, nnzB,
                                                       (const I*)pB.as<I>(), (const I*)colB,
                                                       (const T*)valB, ptrC, (I*)nullptr,
                                                       (T*)nullptr)) != GSA_OK)
                    return fail(r);
                int64_t total = 0;
                if ((r = counts_to_ptr<I>(s, ptrC, m, &total)) != GSA_OK) return fail(r);
                if (total > 0) {
                    if ((r = ws_alloc(&c->buf[1], sizeof(I) * total, s)) != GSA_OK) return fail(r);
                    if ((r = ws_alloc(&c->buf[2], sizeof(T) * total, s)) != GSA_OK) return fail(r);
                    if ((r = merge_rows<T, I, true, true>(s, m, n, nnzA, (const I*)pA.as<I>(),
                                                          (const I*)colA, (const T*)valA, nnzB,
                                                          (const I*)pB.as<I>(), (const I*)colB,
                                                          (const T*)valB, ptrC, (I*)c->buf[1],
                                                          (T*)c->buf[2])) != GSA_OK)
                        return fail(r);
                }
                hipError_t e1 = hipGetLastError();
                if (e1 != hipSuccess) return fail(hip_fail(e1, "gsa_coo_add_begin"));
                c->aux[1] = 1;  // row merge: the finish sorts by column
                c->nnz_out = total;
                *nnzC = total;
                *ctx = c;
                return (int)GSA_OK;
            }
            const uint64_t* skeys = nullptr;  // keys in column-major order
            // mixed or unordered storage: the concatenation sorted stably by (col, row),
            // a row-major operand by its column alone and then merged with the other,
            // else the radix sort; positions are 32-bit when they fit
            auto order = [&](auto pos_t) -> int {
                using V = decltype(pos_t);
                GSA_TRY(keys.alloc(sizeof(uint64_t) * 2 * tot));
                GSA_TRY(perm.alloc(sizeof(V) * tot));
                uint64_t* k0 = keys.as<uint64_t>();
                uint64_t* k1 = k0 + tot;
                hipLaunchKernelGGL((k_concat_keys<I, V>), dim3(grid_for(tot, 256)), dim3(256), 0,
                                   s, nnzA, (const I*)rowA, (const I*)colA, nnzB, (const I*)rowB,
                                   (const I*)colB, m, n, k0, perm.as<V>());
                if ((cmA || rmA) && (cmB || rmB) && n <= (int64_t)UINT32_MAX &&
                    tot < (int64_t)UINT32_MAX) {
                    // an operand in row-major order (a COO made from a CSR) becomes
                    // column-major by a stable sort on the column alone; then the merge
                    Scratch k2(s), ppos(s), cols(s), lpos(s);
                    GSA_TRY(k2.alloc(sizeof(uint64_t) * tot));
                    GSA_TRY(ppos.alloc(sizeof(V) * tot));
                    const int64_t mx = nnzA > nnzB ? nnzA : nnzB;
                    GSA_TRY(cols.alloc(sizeof(uint32_t) * (mx > 0 ? mx : 1)));
                    GSA_TRY(lpos.alloc(sizeof(uint32_t) * (mx > 0 ? mx : 1)));
                    for (int part = 0; part < 2; ++part) {
                        const int64_t off = part ? nnzA : 0, cnt = part ? nnzB : nnzA;
                        if (cnt == 0) continue;
                        const bool colsorted = part ? cmB : cmA;
                        const unsigned gp = grid_for(cnt, 256);
                        hipLaunchKernelGGL(k_part_cols, dim3(gp), dim3(256), 0, s, cnt,
                                           (const uint64_t*)(k0 + off), (uint64_t)m,
                                           cols.as<uint32_t>(), lpos.as<uint32_t>());
                        if (!colsorted)
                            GSA_TRY(radix_sort_pairs_u32_u32(s, cols.as<uint32_t>(),
                                                             lpos.as<uint32_t>(), cnt,
                                                             bits_for((uint64_t)(n - 1))));
                        hipLaunchKernelGGL((k_part_apply<V>), dim3(gp), dim3(256), 0, s, cnt, off,
                                           (const uint32_t*)lpos.as<uint32_t>(),
                                           (const uint64_t*)(k0 + off), k2.as<uint64_t>() + off,
                                           ppos.as<V>() + off);
                    }
                    GSA_TRY(merge_sorted_parts<V>(s, nnzA, tot, k2.as<uint64_t>(), k1,
                                                  perm.as<V>()));
                    hipLaunchKernelGGL((k_compose<V>), dim3(grid_for(tot, 256)), dim3(256), 0, s,
                                       tot, (const V*)ppos.as<V>(), perm.as<V>());
                    GSA_HIP(hipGetLastError());
                    skeys = k1;
                    return GSA_OK;
                }
                const int bits = bits_for((uint64_t)(n - 1) * (uint64_t)m + (uint64_t)(m - 1));
                if constexpr (std::is_same_v<V, uint32_t>)
                    GSA_TRY(radix_sort_pairs_u64_u32(s, k0, perm.as<V>(), tot, bits));
                else
                    GSA_TRY(radix_sort_pairs_u64(s, k0, perm.as<V>(), tot, bits));
                skeys = k0;
                return GSA_OK;
            };
            const bool pos32 = tot < (int64_t)UINT32_MAX;
            if ((r = pos32 ? order(uint32_t{}) : order(uint64_t{})) != GSA_OK) return fail(r);
            for (int b = 0; b < 4; ++b) {
                size_t sz = b < 2 ? sizeof(I) * tot : (b == 2 ? sizeof(T) * tot : sizeof(T) * tot);
                if ((r = ws_alloc(&c->buf[b], sz, s)) != GSA_OK) return fail(r);
            }
            if ((r = ws_alloc(&c->buf[4], sizeof(int64_t) * tot, s)) != GSA_OK) return fail(r);
            I* ro = (I*)c->buf[0]; I* co = (I*)c->buf[1]; T* vo = (T*)c->buf[2];
            T* sums = (T*)c->buf[3]; int64_t* keep = (int64_t*)c->buf[4];
            if (pos32)
                hipLaunchKernelGGL((k_sorted_entries<T, I, uint32_t>), dim3(grid_for(tot, 256)),
                                   dim3(256), 0, s, tot, (const uint32_t*)perm.p, skeys,
                                   (uint64_t)m, nnzA, (const T*)valA, (const T*)valB, ro, co, vo);
            else
                hipLaunchKernelGGL((k_sorted_entries<T, I, uint64_t>), dim3(grid_for(tot, 256)),
                                   dim3(256), 0, s, tot, (const uint64_t*)perm.p, skeys,
                                   (uint64_t)m, nnzA, (const T*)valA, (const T*)valB, ro, co, vo);
            hipLaunchKernelGGL(k_run_sums<T>, dim3(grid_for(tot, 256)), dim3(256), 0, s, tot,
                               skeys, (const T*)vo, sums, keep);
            if ((r = inclusive_scan<int64_t>(s, keep, keep, tot, 0)) != GSA_OK) return fail(r);
            hipError_t e = hipMemcpyAsync(&c->nnz_out, keep + tot - 1, sizeof(int64_t),
                                          hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return fail(hip_fail(e, "gsa_coo_add_begin"));
        }
        *nnzC = c->nnz_out;
        *ctx = c;
        return (int)GSA_OK;
    });
}

int gsa_coo_add_finish(gsa_ctx ctx, void* rowC, void* colC, void* valC) {
    GSA_CHECK(ctx && ctx->kind == 1, GSA_ERR_INVALID, "gsa_coo_add_finish: bad context");
    hipStream_t s = ctx->stream;
    const int64_t tot = ctx->aux[0];
    int r = dispatch_vai(ctx->vtype, ctx->itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        if (tot == 0) return (int)GSA_OK;
        if (ctx->aux[1] == 2) {  // column merge: fill the rows and values, expand the columns
            const int64_t nz = ctx->nnz_out, m = ctx->m, n = ctx->n;
            if (nz == 0) return (int)GSA_OK;
            const int64_t nnzA = ctx->aux[2], nnzB = tot - nnzA;
            GSA_TRY((merge_rows<T, I, true, true>(s, n, m, nnzA, (const I*)ctx->buf[1],
                                                  (const I*)ctx->in[0], (const T*)ctx->in[1], nnzB,
                                                  (const I*)ctx->buf[2], (const I*)ctx->in[2],
                                                  (const T*)ctx->in[3], (I*)ctx->buf[0], (I*)rowC,
                                                  (T*)valC)));
            return expand_rows(s, ctx->itype, n, nz, ctx->buf[0], colC);
        }
        if (ctx->aux[1] == 1) {  // row merge: stable sort by column, (row, value) carried
            const int64_t nz = ctx->nnz_out, m = ctx->m, n = ctx->n;
            if (nz == 0) return (int)GSA_OK;
            // the rows come from the result's row pointer (the first pass finds them
            // tile by tile), expanded first only for 2^31 rows or more
            Scratch rows(s), keys(s);
            const bool from_ptr = m < ((int64_t)1 << 31);
            if (!from_ptr) {
                GSA_TRY(rows.alloc(sizeof(I) * nz));
                GSA_TRY(expand_rows(s, ctx->itype, m, nz, ctx->buf[0], rows.p));
            }
            const int bits = bits_for((uint64_t)(n - 1));
            auto run = [&](auto key_t) -> int {
                using K = decltype(key_t);
                GSA_TRY(keys.alloc(sizeof(K) * nz));
                const KeyFromIdx<K, I> key0{(const I*)ctx->buf[1], nullptr, n, 1};
                GSA_TRY((radix_sort_payload<K, T, I>(
                    s, nz, bits, key0, (const I*)rows.p, (const T*)ctx->buf[2], keys.as<K>(),
                    (I*)rowC, (T*)valC, from_ptr ? (const I*)ctx->buf[0] : nullptr, m)));
                hipLaunchKernelGGL((k_keys_to_idx<I, K>), dim3(grid_for(nz, 256)), dim3(256), 0, s,
                                   nz, (const K*)keys.as<K>(), (I*)colC);
                GSA_HIP(hipGetLastError());
                return (int)GSA_OK;
            };
            return bits <= 32 ? run(uint32_t{}) : run(uint64_t{});
        }
        hipLaunchKernelGGL((k_compact<T, I>), dim3(grid_for(tot, 256)), dim3(256), 0, s, tot,
                           (const int64_t*)ctx->buf[4], (const I*)ctx->buf[0],
                           (const I*)ctx->buf[1], (const T*)ctx->buf[3], (I*)rowC, (I*)colC,
                           (T*)valC);
        GSA_HIP(hipGetLastError());
        return (int)GSA_OK;
    });
    gsa_ctx_destroy(ctx);
    return r;
}

int gsa_kron_coo(void* stream, int vtype, int itype, int64_t mB, int64_t nB, int64_t nnzA,
                 const void* rowA, const void* colA, const void* valA, int64_t nnzB,
                 const void* rowB, const void* colB, const void* valB, void* rowC, void* colC,
                 void* valC) {
    GSA_TRY(ensure_init());
    GSA_CHECK(mB >= 0 && nB >= 0 && nnzA >= 0 && nnzB >= 0, GSA_ERR_INVALID,
              "gsa_kron_coo: bad arguments");
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        return run_kron_coo<T, I>(s, mB, nB, nnzA, (const I*)rowA, (const I*)colA, (const T*)valA,
                                  nnzB, (const I*)rowB, (const I*)colB, (const T*)valB, (I*)rowC,
                                  (I*)colC, (T*)valC);
    });
}

int gsa_kron_csr(void* stream, int vtype, int itype, int64_t mA, int64_t nA, int64_t mB,
                 int64_t nB, const void* rowptrA, const void* colA, const void* valA, int64_t nnzA,
                 const void* rowptrB, const void* colB, const void* valB, int64_t nnzB,
                 void* rowptrC, void* colC, void* valC) {
    GSA_TRY(ensure_init());
    GSA_CHECK(mA >= 0 && nA >= 0 && mB >= 0 && nB >= 0 && nnzA >= 0 && nnzB >= 0, GSA_ERR_INVALID,
              "gsa_kron_csr: bad arguments");
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        const int64_t mC = mA * mB, nnzC = nnzA * nnzB;
        // canonical operands?
        Scratch flag(s);
        GSA_TRY(flag.alloc(sizeof(int)));
        GSA_TRY(zero_fill(s, flag.p, sizeof(int)));
        if (mA > 0)
            hipLaunchKernelGGL(k_check_canonical<I>, dim3(grid_for(mA, 256)), dim3(256), 0, s, mA,
                               (const I*)rowptrA, (const I*)colA, flag.as<int>());
        if (mB > 0)
            hipLaunchKernelGGL(k_check_canonical<I>, dim3(grid_for(mB, 256)), dim3(256), 0, s, mB,
                               (const I*)rowptrB, (const I*)colB, flag.as<int>());
        int bad = 0;
        GSA_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
        GSA_HIP(hipStreamSynchronize(s));
        if (!bad) {
            hipLaunchKernelGGL(k_kron_rowptr<I>, dim3(grid_for(mC + 1, 256)), dim3(256), 0, s, mA, mB,
                               nnzB, (const I*)rowptrA, (const I*)rowptrB, (I*)rowptrC);
            if (nnzC > 0 && mB > 0) {
                const int64_t waves = mC < 256 * 32 * 4 * 8 ? mC : 256 * 32 * 4 * 8;
                const unsigned g = (unsigned)cdiv(waves, 4);
                hipLaunchKernelGGL((k_kron_csr_fill<T, I>), dim3(g), dim3(256), 0, s, mA, mB, nB, nnzB,
                                   (const I*)rowptrA, (const I*)colA, (const T*)valA,
                                   (const I*)rowptrB, (const I*)colB, (const T*)valB, (I*)colC,
                                   (T*)valC);
            }
            GSA_HIP(hipGetLastError());
            return (int)GSA_OK;
        }
        // reference route: CSR -> COO, kron COO, COO -> CSR (keyed sort)
        Scratch rA(s), rB(s), rC(s), cCc(s), vCc(s);
        GSA_TRY(rA.alloc(sizeof(I) * (nnzA ? nnzA : 1)));
        GSA_TRY(rB.alloc(sizeof(I) * (nnzB ? nnzB : 1)));
        GSA_TRY(expand_rows(s, itype, mA, nnzA, rowptrA, rA.p));
        GSA_TRY(expand_rows(s, itype, mB, nnzB, rowptrB, rB.p));
        GSA_TRY(rC.alloc(sizeof(I) * (nnzC ? nnzC : 1)));
        GSA_TRY(cCc.alloc(sizeof(I) * (nnzC ? nnzC : 1)));
        GSA_TRY(vCc.alloc(sizeof(T) * (nnzC ? nnzC : 1)));
        GSA_TRY((run_kron_coo<T, I>(s, mB, nB, nnzA, rA.as<I>(), (const I*)colA, (const T*)valA, nnzB,
                                    rB.as<I>(), (const I*)colB, (const T*)valB, rC.as<I>(),
                                    cCc.as<I>(), vCc.as<T>())));
        return gsa_coo_to_csr(s, vtype, itype, mC, nA * nB, nnzC, rC.p, cCc.p, vCc.p, rowptrC, colC,
                              valC);
    });
}

int gsa_spgemm_csr_begin(void* stream, int vtype, int itype, int64_t m, int64_t kdim, int64_t n,
                         const void* rowptrA, const void* colA, const void* valA, int64_t nnzA,
                         const void* rowptrB, const void* colB, const void* valB, int64_t nnzB,
                         gsa_ctx* ctx, int64_t* nnzC) {
    GSA_TRY(ensure_init());
    GSA_CHECK(ctx && nnzC && m >= 0 && kdim >= 0 && n >= 0 && nnzA >= 0 && nnzB >= 0,
              GSA_ERR_INVALID, "gsa_spgemm_csr_begin: bad arguments");
    GSA_CHECK(m == 0 || n == 0 || (uint64_t)m <= UINT64_MAX / (uint64_t)n, GSA_ERR_INVALID,
              "gsa_spgemm_csr_begin: m*n overflows 64-bit keys");
    hipStream_t s = as_stream(stream);
    *ctx = nullptr;
    *nnzC = 0;
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        auto* c = new gsa_ctx_s();
        c->kind = 2; c->stream = s; c->vtype = vtype; c->itype = itype; c->m = m; c->n = n;
        auto fail = [&](int r) { gsa_ctx_destroy(c); return r; };
        int r;
        // per-row counts (ctx buf[0]: I[m+1]) always allocated
        if ((r = ws_alloc(&c->buf[0], sizeof(I) * (m + 1), s)) != GSA_OK) return fail(r);
        if ((r = zero_fill(s, c->buf[0], sizeof(I) * (m + 1))) != GSA_OK) return fail(r);
        if (m == 0 || n == 0 || nnzA == 0 || nnzB == 0 || kdim == 0) {
            *ctx = c;
            return (int)GSA_OK;
        }
        // the product pointer: kept in the context for the numeric phase or finish
        if ((r = ws_alloc(&c->buf[3], sizeof(int64_t) * (m + 1), s)) != GSA_OK) return fail(r);
        int64_t* pp = (int64_t*)c->buf[3];
        // per-row products with the long-row flags, largest count and size classes,
        // both scans, then one read-back of the four figures
        Scratch fl(s), mx(s), fig(s);
        if ((r = fl.alloc(sizeof(int64_t) * (m > 0 ? m : 1))) != GSA_OK) return fail(r);
        if ((r = mx.alloc(2 * sizeof(unsigned long long))) != GSA_OK) return fail(r);
        if ((r = fig.alloc(4 * sizeof(int64_t))) != GSA_OK) return fail(r);
        if ((r = zero_fill(s, mx.p, 2 * sizeof(unsigned long long))) != GSA_OK) return fail(r);
        if (m > 0)
            hipLaunchKernelGGL((k_sg_count_flags<I>), dim3((unsigned)cdiv(m, 256)), dim3(256), 0, s,
                               m, (const I*)rowptrA, (const I*)colA, (const I*)rowptrB, kdim,
                               pp, fl.as<int64_t>(), (unsigned long long*)mx.p);
        if ((r = counts_to_ptr<int64_t>(s, pp, m, nullptr)) != GSA_OK) return fail(r);
        if ((r = inclusive_scan<int64_t>(s, fl.as<int64_t>(), fl.as<int64_t>(), m, 0)) != GSA_OK)
            return fail(r);
        int64_t h[4] = {0, 0, 0, 1};  // largest count, long rows, classes, pointer end
        {
            hipLaunchKernelGGL(k_sg_figures, dim3(1), dim3(64), 0, s,
                               (const unsigned long long*)mx.p, (const int64_t*)fl.as<int64_t>(), m,
                               (const int64_t*)pp, fig.as<int64_t>());
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(h, fig.p, sizeof h, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return fail(hip_fail(e, "gsa_spgemm_csr_begin"));
        }
        const int64_t P = h[3] - 1;
        c->aux[0] = P;
        if (P == 0) { *ctx = c; return (int)GSA_OK; }
        // row-local path when every row's products fit one workgroup's sort
        {
            hipError_t e = hipSuccess;
            const int64_t maxP = h[0], nbig = h[1];
            const unsigned classes = (unsigned)h[2];
            if (maxP <= SG_BLOCK_P && n < (int64_t)1 << 40) {
                // keep the long-row list for the numeric phase
                if (nbig > 0) {
                    if ((r = ws_alloc(&c->buf[4], sizeof(int64_t) * nbig, s)) != GSA_OK) return fail(r);
                    hipLaunchKernelGGL(k_sg_list, dim3((unsigned)cdiv(m, 256)), dim3(256), 0, s, m,
                                       (const int64_t*)pp, (const int64_t*)fl.as<int64_t>(),
                                       (int64_t*)c->buf[4]);
                }
                c->aux[1] = 1;  // row-local path
                c->aux[2] = nbig;
                c->aux[3] = kdim;
                c->aux[4] = nnzA;
                c->aux[5] = nnzB;
                c->aux[6] = classes;
                c->in[0] = rowptrA; c->in[1] = colA; c->in[2] = valA;
                c->in[3] = rowptrB; c->in[4] = colB; c->in[5] = valB;
                I* cnt = (I*)c->buf[0];
                // a staging area of modest size: one pass sorts, sums and counts every
                // row; the finish only moves the rows
                const size_t stage = (size_t)P * (sizeof(I) + sizeof(T));
                if (stage <= ((size_t)4 << 30)) {
                    if ((r = ws_alloc(&c->buf[5], sizeof(I) * P, s)) != GSA_OK) return fail(r);
                    if ((r = ws_alloc(&c->buf[6], sizeof(T) * P, s)) != GSA_OK) return fail(r);
                    sg_rows_small<true, T, I, true>(s, m, n, kdim, nnzA, nnzB, (const I*)rowptrA,
                                                    (const I*)colA, (const T*)valA,
                                                    (const I*)rowptrB, (const I*)colB,
                                                    (const T*)valB, (const int64_t*)c->buf[3],
                                                    cnt, (const I*)nullptr, (I*)c->buf[5],
                                                    (T*)c->buf[6], classes);
                    if (nbig > 0)
                        hipLaunchKernelGGL((k_sg_row_block<true, T, I, true>), dim3((unsigned)nbig),
                                           dim3(ROW_WG), 0, s, n, kdim, nnzA, nnzB,
                                           (const I*)rowptrA, (const I*)colA, (const T*)valA,
                                           (const I*)rowptrB, (const I*)colB, (const T*)valB,
                                           (const int64_t*)c->buf[3], (const int64_t*)c->buf[4],
                                           cnt, (const I*)nullptr, (I*)c->buf[5], (T*)c->buf[6]);
                    if ((e = hipGetLastError()) != hipSuccess) return fail(hip_fail(e, "k_sg_row"));
                    c->aux[1] = 2;  // staged
                    int64_t total = 0;
                    if ((r = counts_to_ptr<I>(s, cnt, m, &total)) != GSA_OK) return fail(r);
                    c->nnz_out = total;
                    *nnzC = total;
                    *ctx = c;
                    return (int)GSA_OK;
                }
                sg_rows_small<false, T, I>(s, m, n, kdim, nnzA, nnzB, (const I*)rowptrA,
                                           (const I*)colA, (const T*)valA, (const I*)rowptrB,
                                           (const I*)colB, (const T*)valB,
                                           (const int64_t*)c->buf[3], cnt, (const I*)nullptr,
                                           (I*)nullptr, (T*)nullptr, classes);
                if (nbig > 0)
                    hipLaunchKernelGGL((k_sg_row_block<false, T, I>), dim3((unsigned)nbig), dim3(ROW_WG),
                                       0, s, n, kdim, nnzA, nnzB, (const I*)rowptrA, (const I*)colA,
                                       (const T*)valA, (const I*)rowptrB, (const I*)colB,
                                       (const T*)valB, (const int64_t*)c->buf[3],
                                       (const int64_t*)c->buf[4], cnt, (const I*)nullptr,
                                       (I*)nullptr, (T*)nullptr);
                if ((e = hipGetLastError()) != hipSuccess) return fail(hip_fail(e, "k_sg_row"));
                int64_t total = 0;
                if ((r = counts_to_ptr<I>(s, cnt, m, &total)) != GSA_OK) return fail(r);
                c->nnz_out = total;
                *nnzC = total;
                *ctx = c;
                return (int)GSA_OK;
            }
        }
        Scratch keys(s), prods(s), perm(s), sorted(s), flags(s);
        if ((r = keys.alloc(sizeof(uint64_t) * P)) != GSA_OK) return fail(r);
        if ((r = prods.alloc(sizeof(T) * P)) != GSA_OK) return fail(r);
        if ((r = perm.alloc(sizeof(uint64_t) * P)) != GSA_OK) return fail(r);
        hipLaunchKernelGGL((k_sg_expand<T, I, uint64_t>), dim3(grid_for(m, 256)), dim3(256), 0, s, m, n,
                           kdim, (const I*)rowptrA, (const I*)colA, (const T*)valA,
                           (const I*)rowptrB, (const I*)colB, (const T*)valB,
                           (const int64_t*)pp, keys.as<uint64_t>(), prods.as<T>(),
                           perm.as<uint64_t>());
        const int bits = bits_for((uint64_t)(m - 1) * (uint64_t)n + (uint64_t)(n - 1));
        if ((r = radix_sort_pairs_u64(s, keys.as<uint64_t>(), perm.as<uint64_t>(), P, bits)) != GSA_OK)
            return fail(r);
        if ((r = sorted.alloc(sizeof(T) * P)) != GSA_OK) return fail(r);
        hipLaunchKernelGGL((k_gather_vals<T, uint64_t>), dim3(grid_for(P, 256)), dim3(256), 0, s, P,
                           (const uint64_t*)perm.as<uint64_t>(), (const T*)prods.as<T>(),
                           sorted.as<T>());
        if ((r = flags.alloc(sizeof(int64_t) * P)) != GSA_OK) return fail(r);
        hipLaunchKernelGGL(k_sg_heads<T>, dim3(grid_for(P, 256)), dim3(256), 0, s, P,
                           (const uint64_t*)keys.as<uint64_t>(), flags.as<int64_t>());
        if ((r = inclusive_scan<int64_t>(s, flags.as<int64_t>(), flags.as<int64_t>(), P, 0)) != GSA_OK)
            return fail(r);
        int64_t total = 0;
        hipError_t e = hipMemcpyAsync(&total, flags.as<int64_t>() + P - 1, sizeof(int64_t),
                           hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return fail(hip_fail(e, "gsa_spgemm_csr_begin"));
        c->nnz_out = total;
        if ((r = ws_alloc(&c->buf[1], sizeof(uint64_t) * total, s)) != GSA_OK) return fail(r);
        if ((r = ws_alloc(&c->buf[2], sizeof(T) * total, s)) != GSA_OK) return fail(r);
        hipLaunchKernelGGL((k_sg_sum<T, I>), dim3(grid_for(P, 256)), dim3(256), 0, s, P, n,
                           (const uint64_t*)keys.as<uint64_t>(), (const T*)sorted.as<T>(),
                           (const int64_t*)flags.as<int64_t>(), (uint64_t*)c->buf[1],
                           (T*)c->buf[2], (I*)c->buf[0] + 1);
        e = hipGetLastError();
        if (e != hipSuccess) return fail(hip_fail(e, "k_sg_sum"));
        *nnzC = total;
        *ctx = c;
        return (int)GSA_OK;
    });
}

int gsa_spgemm_csr_finish(gsa_ctx ctx, void* rowptrC, void* colC, void* valC) {
    GSA_CHECK(ctx && ctx->kind == 2, GSA_ERR_INVALID, "gsa_spgemm_csr_finish: bad context");
    hipStream_t s = ctx->stream;
    int r = dispatch_vai(ctx->vtype, ctx->itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        const int64_t m = ctx->m;
        GSA_HIP(hipMemcpyAsync(rowptrC, ctx->buf[0], sizeof(I) * (m + 1), hipMemcpyDeviceToDevice, s));
        if (ctx->aux[1] == 2) {  // staged: move the rows to their place
            if (ctx->nnz_out > 0)
                hipLaunchKernelGGL((k_sg_compact<T, I>), dim3(grid_for(m * 16, 256)), dim3(256), 0,
                                   s, m, (const int64_t*)ctx->buf[3], (const I*)rowptrC,
                                   (const I*)ctx->buf[5], (const T*)ctx->buf[6], (I*)colC,
                                   (T*)valC);
            GSA_HIP(hipGetLastError());
            return (int)GSA_OK;
        }
        if (ctx->aux[1] == 1) {  // row-local: the pointer is final; write the rows
            if (ctx->nnz_out > 0) {
                const int64_t n = ctx->n, kdim = ctx->aux[3], nnzA = ctx->aux[4], nnzB = ctx->aux[5];
                const I* pA = (const I*)ctx->in[0];
                const I* cA = (const I*)ctx->in[1];
                const T* vA = (const T*)ctx->in[2];
                const I* pB = (const I*)ctx->in[3];
                const I* cB = (const I*)ctx->in[4];
                const T* vB = (const T*)ctx->in[5];
                sg_rows_small<true, T, I>(s, m, n, kdim, nnzA, nnzB, pA, cA, vA, pB, cB, vB,
                                          (const int64_t*)ctx->buf[3], (I*)nullptr,
                                          (const I*)rowptrC, (I*)colC, (T*)valC,
                                          (unsigned)ctx->aux[6]);
                if (ctx->aux[2] > 0)
                    hipLaunchKernelGGL((k_sg_row_block<true, T, I>), dim3((unsigned)ctx->aux[2]),
                                       dim3(ROW_WG), 0, s, n, kdim, nnzA, nnzB, pA, cA, vA, pB, cB,
                                       vB, (const int64_t*)ctx->buf[3],
                                       (const int64_t*)ctx->buf[4], (I*)nullptr, (const I*)rowptrC,
                                       (I*)colC, (T*)valC);
                GSA_HIP(hipGetLastError());
            }
            return (int)GSA_OK;
        }
        GSA_TRY(counts_to_ptr<I>(s, (I*)rowptrC, m, nullptr));
        if (ctx->nnz_out > 0) {
            hipLaunchKernelGGL((k_sg_write<T, I>), dim3(grid_for(ctx->nnz_out, 256)), dim3(256), 0, s,
                               ctx->nnz_out, ctx->n, (const uint64_t*)ctx->buf[1],
                               (const T*)ctx->buf[2], (I*)colC, (T*)valC);
            GSA_HIP(hipGetLastError());
        }
        return (int)GSA_OK;
    });
    gsa_ctx_destroy(ctx);
    return r;
}

}  // extern "C"
