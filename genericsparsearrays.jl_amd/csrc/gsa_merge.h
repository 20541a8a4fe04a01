// gsa_merge.h — the row merge of sparse + on gfx950: CSR + CSR, CSC + CSC (on
// the column pointers), and COO + COO of operands stored in row or column order.
#pragma once
#include "gsa_rows.h"

namespace gsa {

// Row merge of A + B.
//   RUNS = false: CSR + (kernel_add_csr!, matrix_csr_kernels.jl:188-248): a
//     two-pointer merge, equal columns summed as vA + vB;
//   RUNS = true: COO + of operands stored in (row, col) order: equal (row, col)
//     entries are summed left to right, A's in storage order then B's -- the
//     order of the reference's stable sort of the concatenation and its run
//     sums (matrix_coo.jl:300-384, matrix_coo_kernels.jl:260-310); columns are
//     clamped into range as the sort keys are.
// Exact zeros are dropped.  Without FILL each row's count goes to ptrC[r + 1];
// with FILL row r's entries go to positions ptrC[r] - 1 onwards.
//
// One workgroup per block of consecutive rows; the rows per block (at most
// MB_WG) are chosen from the mean row length so that a block holds about MB_CAP
// entries (A's and B's), which gives dense matrices enough blocks to fill the
// chip.
//  * A block whose entries fit MB_STAGE, in rows of at most MB_SHORT entries, is
//    read into LDS together (coalesced) and each row is merged there by one
//    thread, the two-pointer loop (merge_row_serial).
//  * Any other block is merged in runs of whole rows whose entries fit MB_CAP
//    staged slots.  Each thread gives its MB_CH consecutive staged entries their
//    rows (a search among the staged row starts, then a step) and one 64-bit
//    key, row << MB_KB | column key; then every lane merges MB_CH consecutive
//    positions of the output order, found by a merge-path search on those keys
//    in which A's entry goes first between equal keys.  A row of more than
//    MB_CAP entries is merged the same way in chunks of MB_CAP positions, split
//    by merge-path searches over the row in global memory; the entry just
//    before a chunk's part of A and the one just after its part of B are staged
//    too, so an equal pair that a chunk boundary separates is still summed once.
//  * A run whose rows, or a long row whose columns, do not strictly increase (or
//    reach 2^MB_KB), and a block whose row pointers decrease, are merged row by
//    row by one thread from the operands, which keeps the two-pointer semantics
//    for such input.  Every read is clamped into the operands' arrays.
constexpr int MB_WG = 256;
// one-thread-per-row blocks: 45 KB of Float64/Int64 entries, so that with the
// row starts and output starts three workgroups fit a CU's LDS; rows of at most
// MB_SHORT entries
constexpr int MB_STAGE = 2880;
constexpr int MB_SHORT = 64;
constexpr int MB_CAP = 2048;    // merge-path runs and chunks
constexpr int MB_CH = MB_CAP / MB_WG;  // output-order positions per lane
constexpr int MB_KB = 40;              // bits of the column key in a packed key
constexpr uint64_t MB_KMASK = (uint64_t(1) << MB_KB) - 1;
constexpr uint64_t MB_NONE = ~uint64_t(0);  // a key that no staged entry has

// The serial merge of one row; colA(k) etc. read entry k of A / B.  With FILL
// the row's entries go to positions w onwards; without, its count to ptrC[r + 1].
template <typename T, typename I, bool RUNS, bool FILL, typename CA, typename VA, typename CB,
          typename VB>
__device__ __forceinline__ void merge_row_serial(int64_t r, int64_t n, int64_t a, int64_t ea,
                                                 int64_t b, int64_t eb, CA colA, VA valA,
                                                 CB colB, VB valB, I* __restrict__ ptrC,
                                                 int64_t w, I* __restrict__ cC,
                                                 T* __restrict__ vC) {
    int64_t cnt = 0;
    auto emit = [&](I c, T v) {
        if (v == T(0)) return;
        if (FILL) {
            cC[w] = c;
            vC[w] = v;
            ++w;
        }
        ++cnt;
    };
    while (a < ea || b < eb) {
        if constexpr (RUNS) {
            const int64_t ca = a < ea ? dclamp((int64_t)colA(a) - 1, 0, n - 1) : n;
            const int64_t cb = b < eb ? dclamp((int64_t)colB(b) - 1, 0, n - 1) : n;
            const int64_t c = ca < cb ? ca : cb;
            T sum = T(0);
            bool first = true;
            while (a < ea && dclamp((int64_t)colA(a) - 1, 0, n - 1) == c) {
                sum = first ? valA(a) : sum + valA(a);
                first = false;
                ++a;
            }
            while (b < eb && dclamp((int64_t)colB(b) - 1, 0, n - 1) == c) {
                sum = first ? valB(b) : sum + valB(b);
                first = false;
                ++b;
            }
            emit((I)(c + 1), sum);
        } else {
            if (a < ea && (b >= eb || colA(a) < colB(b))) {
                emit(colA(a), valA(a));
                ++a;
            } else if (b < eb && (a >= ea || colA(a) > colB(b))) {
                emit(colB(b), valB(b));
                ++b;
            } else {
                emit(colA(a), valA(a) + valB(b));
                ++a;
                ++b;
            }
        }
    }
    if (!FILL) ptrC[r + 1] = (I)cnt;
}

// The staged part of a run or chunk: A's entries in slots [1, 1 + nA), B's in
// [1 + nA, 1 + nA + nB); slot 0 holds the entry before A's part and slot
// 1 + nA + nB the entry after B's part (MB_NONE when there is none).  Keys are
// row << MB_KB | column key, the row counted within the block and the column
// key being the column for CSR and the clamped zero-based column for RUNS.
//
// One lane merges the output-order positions [d0, d1) (d1 - d0 <= CH): a
// merge-path search finds how many of A's entries come before d0, then each
// position takes A's entry when its key is not greater than B's.  An A entry
// whose key equals the B entry it is compared with carries vA + vB; that B
// entry then comes next and is skipped.  Bit u of the result is set when
// position d0 + u yields an entry (nonzero value); its column and value are in
// oc[u], ov[u].
template <typename T, typename I, bool RUNS, int CH>
__device__ __forceinline__ unsigned mb_lane_merge(int nA, int nB, int d0, int d1,
                                                  const uint64_t* s_k, const T* s_v, I* oc,
                                                  T* ov) {
    auto out_col = [&](uint64_t k) -> I { return (I)((int64_t)(k & MB_KMASK) + (RUNS ? 1 : 0)); };
    const int sB = 1 + nA;
    int lo = d0 - nB > 0 ? d0 - nB : 0, hi = d0 < nA ? d0 : nA;
    while (lo < hi) {  // A(mid) comes before B(d0 - 1 - mid) unless B's key is smaller
        const int mid = (lo + hi) >> 1;
        if (s_k[1 + mid] <= s_k[sB + d0 - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    int i = lo, j = d0 - lo;
    unsigned keep = 0;
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        if (d0 + u < d1) {
            // slot 1 + nA is B's first entry when i == nA; sB + nB the slot after
            // B's part when j == nB
            const uint64_t ka = s_k[1 + i], kb = s_k[sB + j];
            if (j >= nB || (i < nA && ka <= kb)) {
                T v = s_v[1 + i];
                if (ka == kb) v = v + s_v[sB + j];
                oc[u] = out_col(ka);
                ov[u] = v;
                if (!(v == T(0))) keep |= 1u << u;
                ++i;
            } else {  // slot i holds A's entry i - 1 (the slot before A's part when i == 0)
                const T v = s_v[sB + j];
                oc[u] = out_col(kb);
                ov[u] = v;
                if (s_k[i] != kb && !(v == T(0))) keep |= 1u << u;
                ++j;
            }
        }
    }
    return keep;
}

// The row of staged entry k: the last of the rows [lo, hi) whose start rs[row]
// is at or before k (rs[lo] <= k).
template <typename R>
__device__ __forceinline__ int mb_row_of(const R* rs, int lo, int hi, int k) {
    hi -= 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rs[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T, typename I>
constexpr int mb_lds_bytes() {
    constexpr int one = MB_STAGE * (int)(sizeof(I) + sizeof(T));
    constexpr int runs = (MB_CAP + 2) * (int)(sizeof(uint64_t) + sizeof(T));
    return one > runs ? one : runs;
}

// Row block blockIdx: rows [r0, r0 + rpb), rpb <= MB_WG.
template <typename T, typename I, bool RUNS, bool FILL>
__global__ __launch_bounds__(MB_WG) void k_merge_blocks(
    int64_t m, int64_t n, int rpb, int64_t nnzA, const I* __restrict__ pA,
    const I* __restrict__ cA, const T* __restrict__ vA, int64_t nnzB, const I* __restrict__ pB,
    const I* __restrict__ cB, const T* __restrict__ vB, I* __restrict__ ptrC,
    I* __restrict__ cC, T* __restrict__ vC) {
    // one buffer: the whole block's entries (one thread per row), or a run's or
    // chunk's keys and values
    __shared__ __align__(16) unsigned char s_buf[mb_lds_bytes<T, I>()];
    __shared__ int rsA[MB_WG + 1], rsB[MB_WG + 1];  // row starts within the block
    __shared__ int64_t s_w[FILL ? MB_WG + 1 : 1];   // rows' output starts
    __shared__ int64_t s_split[MB_WG];              // a long row's chunk splits
    __shared__ int lds4[MB_WG / 64];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * rpb;
    if (r0 >= m) return;
    const int nrows = (int)(r0 + rpb < m ? rpb : m - r0);
    auto pa = [&](int64_t r) { return dclamp((int64_t)pA[r] - 1, 0, nnzA); };
    auto pb = [&](int64_t r) { return dclamp((int64_t)pB[r] - 1, 0, nnzB); };
    // column key, as an unsigned value: a column below 1 or of 2^MB_KB and more
    // exceeds MB_KMASK
    auto key = [&](I c) -> uint64_t {
        return RUNS ? (uint64_t)dclamp((int64_t)c - 1, 0, n - 1) : (uint64_t)(int64_t)c;
    };
    // row r from global memory by this thread, entry ranges clamped
    auto serial_global = [&](int64_t r) {
        const int64_t a = pa(r), b = pb(r);
        const int64_t ea = pa(r + 1) > a ? pa(r + 1) : a, eb = pb(r + 1) > b ? pb(r + 1) : b;
        merge_row_serial<T, I, RUNS, FILL>(
            r, n, a, ea, b, eb, [&](int64_t k) { return cA[k]; },
            [&](int64_t k) { return vA[k]; }, [&](int64_t k) { return cB[k]; },
            [&](int64_t k) { return vB[k]; }, ptrC, FILL ? (int64_t)ptrC[r] - 1 : 0, cC, vC);
    };

    // the block's row pointers (and output starts), relative to its first row
    const int64_t a_s = pa(r0), b_s = pb(r0);
    for (int k = tid; k <= nrows; k += MB_WG) {
        rsA[k] = (int)dclamp(pa(r0 + k) - a_s, 0, INT32_MAX);
        rsB[k] = (int)dclamp(pb(r0 + k) - b_s, 0, INT32_MAX);
        if (FILL && k < nrows) s_w[k] = (int64_t)ptrC[r0 + k] - 1;
    }
    __syncthreads();
    int dec = 0;
    for (int k = 1 + tid; k <= nrows; k += MB_WG)
        dec |= rsA[k] < rsA[k - 1] || rsB[k] < rsB[k - 1];
    // (a block of 2^29 entries or more in an operand: its offsets stay 64-bit there)
    if (tid == 0) dec |= rsA[nrows] >= (1 << 29) || rsB[nrows] >= (1 << 29);
    if (__syncthreads_or(dec)) {  // row pointers that decrease: one row per thread
        if (tid < nrows) serial_global(r0 + tid);
        return;
    }
    const int64_t total_entries = (int64_t)rsA[nrows] + rsB[nrows];
    int lng = 0;  // a row too long for one thread to merge while the others wait
    for (int k = tid; k < nrows; k += MB_WG)
        lng |= (rsA[k + 1] - rsA[k]) + (rsB[k + 1] - rsB[k]) > MB_SHORT;
    lng = __syncthreads_or(lng);

    if (total_entries <= MB_STAGE && !lng) {  // the whole block in LDS, one thread per row
        I* s_c = reinterpret_cast<I*>(s_buf);
        T* s_vv = reinterpret_cast<T*>(s_buf + MB_STAGE * sizeof(I));
        const int na = rsA[nrows], nb = rsB[nrows];
        for (int k = tid; k < na; k += MB_WG) {
            s_c[k] = cA[a_s + k];
            s_vv[k] = vA[a_s + k];
        }
        for (int k = tid; k < nb; k += MB_WG) {
            s_c[na + k] = cB[b_s + k];
            s_vv[na + k] = vB[b_s + k];
        }
        __syncthreads();
        if (tid < nrows)
            merge_row_serial<T, I, RUNS, FILL>(
                r0 + tid, n, rsA[tid], rsA[tid + 1], rsB[tid], rsB[tid + 1],
                [&](int64_t k) { return s_c[k]; }, [&](int64_t k) { return s_vv[k]; },
                [&](int64_t k) { return s_c[na + k]; }, [&](int64_t k) { return s_vv[na + k]; },
                ptrC, FILL ? s_w[tid] : 0, cC, vC);
        return;
    }

    uint64_t* s_k = reinterpret_cast<uint64_t*>(s_buf);
    T* s_v = reinterpret_cast<T*>(s_buf + (MB_CAP + 2) * sizeof(uint64_t));
    // stage nA entries of A from ga and nB of B from gb in slots 1 .. nA + nB
    auto stage = [&](int64_t ga, int nA, int64_t gb, int nB) {
        for (int k = tid; k < nA; k += MB_WG) {
            s_k[1 + k] = key(cA[ga + k]);
            s_v[1 + k] = vA[ga + k];
        }
        for (int k = tid; k < nB; k += MB_WG) {
            s_k[1 + nA + k] = key(cB[gb + k]);
            s_v[1 + nA + k] = vB[gb + k];
        }
    };
    for (int q0 = 0; q0 < nrows;) {
        const int f0 = rsA[q0] + rsB[q0];
        if (rsA[q0 + 1] + rsB[q0 + 1] - f0 > MB_CAP) {
            // --- row q0 alone, in chunks of MB_CAP output-order positions ---
            const int64_t q = r0 + q0;
            const int64_t a = a_s + rsA[q0], b = b_s + rsB[q0];
            const int64_t na = rsA[q0 + 1] - rsA[q0], nb = rsB[q0 + 1] - rsB[q0];
            int bad = 0;
            for (int64_t k = tid; k < na; k += MB_WG) {
                const uint64_t kc = key(cA[a + k]);
                bad |= kc > MB_KMASK || (k > 0 && key(cA[a + k - 1]) >= kc);
            }
            for (int64_t k = tid; k < nb; k += MB_WG) {
                const uint64_t kc = key(cB[b + k]);
                bad |= kc > MB_KMASK || (k > 0 && key(cB[b + k - 1]) >= kc);
            }
            if (__syncthreads_or(bad)) {
                if (tid == 0) serial_global(q);
            } else {
                const int64_t L = na + nb, nch = (L + MB_CAP - 1) / MB_CAP;
                const int64_t w0 = FILL ? s_w[q0] : 0;
                int64_t run = 0;
                // chunk boundaries c0 .. c0 + MB_WG - 1, then the chunks between them
                for (int64_t c0 = 0; c0 < nch; c0 += MB_WG - 1) {
                    const int64_t c = c0 + tid;
                    if (c <= nch) {
                        const int64_t d = c * MB_CAP < L ? c * MB_CAP : L;
                        int64_t lo2 = d - nb > 0 ? d - nb : 0, hi2 = d < na ? d : na;
                        while (lo2 < hi2) {
                            const int64_t mid = (lo2 + hi2) >> 1;
                            if (key(cA[a + mid]) <= key(cB[b + d - 1 - mid])) lo2 = mid + 1;
                            else hi2 = mid;
                        }
                        s_split[tid] = lo2;
                    }
                    __syncthreads();
                    const int64_t cend = c0 + MB_WG - 1 < nch ? c0 + MB_WG - 1 : nch;
                    for (int64_t cc = c0; cc < cend; ++cc) {
                        const int64_t d0 = cc * MB_CAP, d1 = d0 + MB_CAP < L ? d0 + MB_CAP : L;
                        const int64_t i0 = s_split[cc - c0], i1 = s_split[cc - c0 + 1];
                        const int64_t j0 = d0 - i0, j1 = d1 - i1;
                        const int nAc = (int)(i1 - i0), nBc = (int)(j1 - j0), sB = 1 + nAc;
                        stage(a + i0, nAc, b + j0, nBc);  // row 0: the key is the column key
                        if (tid == 0) {  // the row's entries just outside this chunk
                            const bool pre = i0 > 0, post = j1 < nb;
                            s_k[0] = pre ? key(cA[a + i0 - 1]) : MB_NONE;
                            s_v[0] = pre ? vA[a + i0 - 1] : T(0);
                            s_k[sB + nBc] = post ? key(cB[b + j1]) : MB_NONE;
                            s_v[sB + nBc] = post ? vB[b + j1] : T(0);
                        }
                        __syncthreads();
                        const int tot = nAc + nBc;
                        const int e0 = tid * MB_CH < tot ? tid * MB_CH : tot;
                        const int e1 = e0 + MB_CH < tot ? e0 + MB_CH : tot;
                        I oc[MB_CH];
                        T ov[MB_CH];
                        const unsigned keep =
                            mb_lane_merge<T, I, RUNS, MB_CH>(nAc, nBc, e0, e1, s_k, s_v, oc, ov);
                        int total = 0;
                        const int base = wg_exclusive_sum<int>(__popc(keep), lds4, &total);
                        if (FILL) {
                            I* s_oc = reinterpret_cast<I*>(s_k);
                            int w = base;
                            for (int u = 0; u < MB_CH; ++u)
                                if ((keep >> u) & 1) {
                                    s_oc[w] = oc[u];
                                    s_v[w] = ov[u];
                                    ++w;
                                }
                            __syncthreads();
                            for (int k = tid; k < total; k += MB_WG) {
                                cC[w0 + run + k] = s_oc[k];
                                vC[w0 + run + k] = s_v[k];
                            }
                        }
                        run += total;
                        __syncthreads();  // before the next chunk (or split batch) reuses the LDS
                    }
                }
                if (!FILL && tid == 0) ptrC[q + 1] = (I)run;
            }
            __syncthreads();
            ++q0;
            continue;
        }
        // --- rows [q0, q1): the most whole rows from q0 whose entries fit MB_CAP ---
        int q1 = q0 + 1, qh = nrows;
        while (q1 < qh) {
            const int mid = (q1 + qh + 1) >> 1;
            if (rsA[mid] + rsB[mid] - f0 <= MB_CAP) q1 = mid;
            else qh = mid - 1;
        }
        const int aoff = rsA[q0], boff = rsB[q0];
        const int nA = rsA[q1] - aoff, nB = rsB[q1] - boff, tot = nA + nB, sB = 1 + nA;
        stage(a_s + aoff, nA, b_s + boff, nB);
        if (tid == 0) {  // no entries before A's part or after B's
            s_k[0] = MB_NONE;
            s_v[0] = T(0);
            s_k[sB + nB] = MB_NONE;
            s_v[sB + nB] = T(0);
        }
        __syncthreads();
        // Each thread's MB_CH consecutive entries (A's part, then B's): the row by a
        // search for the first and a step or a search for the next ones, the packed
        // key, and the canonical check (columns strictly increase inside every row
        // of A and of B).  Slot 1 + x holds entry x of the concatenation; its left
        // neighbour may already hold a packed key, whose low bits are the column key
        // all the same.  Entry k of a part is entry off + k of the block's part (the
        // run starts a row).
        int bad = 0;
        {
            int row = q0, part = -1;
            uint64_t prev = 0;
            for (int u = 0; u < MB_CH; ++u) {
                const int x = tid * MB_CH + u;
                if (x >= tot) break;
                const int inA = x < nA;
                const int k = inA ? x : x - nA;
                const int kt = k + (inA ? aoff : boff);
                const int* rs = inA ? rsA : rsB;
                if (inA != part) {
                    row = mb_row_of(rs, q0, q1, kt);
                    part = inA;
                    prev = k > 0 ? (s_k[x] & MB_KMASK) : 0;  // slot x = entry x - 1
                } else if (row + 1 < q1 && rs[row + 1] <= kt) {
                    row = mb_row_of(rs, row + 1, q1, kt);
                }
                const uint64_t kc = s_k[1 + x];
                bad |= kc > MB_KMASK;
                bad |= k > 0 && rs[row] < kt && prev >= kc;  // entry k - 1 in the same row
                prev = kc;
                s_k[1 + x] = ((uint64_t)row << MB_KB) | (kc & MB_KMASK);
            }
        }
        if (__syncthreads_or(bad)) {  // every row of the run by one thread, from the operands
            for (int k = q0 + tid; k < q1; k += MB_WG) serial_global(r0 + k);
        } else {
            const int d0 = tid * MB_CH < tot ? tid * MB_CH : tot;
            const int d1 = d0 + MB_CH < tot ? d0 + MB_CH : tot;
            I oc[MB_CH];
            T ov[MB_CH];
            const unsigned keep =
                mb_lane_merge<T, I, RUNS, MB_CH>(nA, nB, d0, d1, s_k, s_v, oc, ov);
            int total = 0;
            const int base = wg_exclusive_sum<int>(__popc(keep), lds4, &total);
            // (every lane has finished its merge: the staged arrays may be reused)
            if (!FILL) {
                // kept entries before each output-order position, then each row's
                // count as the difference at its start and end
                int* s_pos = reinterpret_cast<int*>(s_k);
                for (int u = 0; u < MB_CH; ++u)
                    if (d0 + u < d1) s_pos[d0 + u] = base + __popc(keep & ((1u << u) - 1));
                if (tid == 0) s_pos[tot] = total;
                __syncthreads();
                for (int k = q0 + tid; k < q1; k += MB_WG) {
                    const int g0 = rsA[k] + rsB[k] - f0, g1 = rsA[k + 1] + rsB[k + 1] - f0;
                    ptrC[r0 + k + 1] = (I)(s_pos[g1] - s_pos[g0]);
                }
            } else {
                I* s_oc = reinterpret_cast<I*>(s_k);
                int w = base;
                for (int u = 0; u < MB_CH; ++u)
                    if ((keep >> u) & 1) {
                        s_oc[w] = oc[u];
                        s_v[w] = ov[u];
                        ++w;
                    }
                __syncthreads();
                const int64_t w0 = s_w[q0];
                for (int k = tid; k < total; k += MB_WG) {
                    cC[w0 + k] = s_oc[k];
                    vC[w0 + k] = s_v[k];
                }
            }
        }
        __syncthreads();  // before the next run (or long row) reuses the LDS
        q0 = q1;
    }
}

// ------------------------------------------------------------------ tiles --
// For longer mean rows the blocks above would hold too few rows to spread the
// work evenly, so the entries are cut into tiles of even size instead:
// g(r) = (entries of A and B before row r) + 2r, and tile t holds the rows r
// with t*MT_TILE <= g(r) < (t+1)*MT_TILE.  A tile thus has at most MT_ROWS =
// MT_TILE/2 rows, and a row of MT_TILE or more entries can only be its last
// row.  The other rows are merged in runs as above (at most MT_CAP staged
// entries, MT_WG lanes of MT_CH positions); the long row in chunks.  A bounds
// pass finds each tile's first row (a record per tile boundary).
constexpr int MT_TILE = 1792;
constexpr int MT_ROWS = MT_TILE / 2;
constexpr int MT_CAP = 2048;
constexpr int MT_WG = 512;
constexpr int MT_CH = MT_CAP / MT_WG;
static_assert(MT_TILE <= MT_CAP, "a row that is not long fits one run");
constexpr int MT_REC = 5;  // per tile boundary: row r, pa(r), pb(r), pa(r - 1), pb(r - 1)
constexpr int MT_AUX = 2 * (MT_ROWS + 1) * 2 > MT_WG * 8 ? 2 * (MT_ROWS + 1) * 2 : MT_WG * 8;

// Exclusive prefix sum over the MT_WG lanes of a workgroup; *total gets the sum.
__device__ __forceinline__ int mt_exclusive_sum(int v, int* lds8, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds8[wave] = x;
    __syncthreads();
    int pre = 0, all = 0;
#pragma unroll
    for (int w = 0; w < MT_WG / 64; ++w) {
        const int s = lds8[w];
        if (w < wave) pre += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return pre + x - v;
}

// Every tile boundary record (MT_REC values per boundary t = 0 .. ntiles) set to
// row m, and *flag to 0.
template <typename I>
__global__ void k_mt_init(int64_t m, int64_t nnzA, const I* __restrict__ pA, int64_t nnzB,
                          const I* __restrict__ pB, int64_t ntiles, int64_t* __restrict__ rec,
                          int* __restrict__ flag) {
    const int64_t am = dclamp((int64_t)pA[m] - 1, 0, nnzA), bm = dclamp((int64_t)pB[m] - 1, 0, nnzB);
    const int64_t am1 = dclamp((int64_t)pA[m - 1] - 1, 0, nnzA);
    const int64_t bm1 = dclamp((int64_t)pB[m - 1] - 1, 0, nnzB);
    const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (int64_t t = t0; t <= ntiles; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t* p = rec + MT_REC * t;
        p[0] = m;
        p[1] = am;
        p[2] = bm;
        p[3] = am1;
        p[4] = bm1;
    }
    if (t0 == 0) *flag = 0;
}

// Boundary t (t < ntiles) is the first row r with g(r) >= t * MT_TILE; its
// record holds r and the clamped entry offsets pa(r), pb(r), pa(r-1), pb(r-1)
// (the records were set to row m beforehand, which stays the record of every
// boundary past g(m), and of boundary ntiles).  Row r writes the boundaries
// whose t * MT_TILE lie in (g(r-1), g(r)].  A row pointer that decreases (after
// clamping into the arrays) sets *flag (a plain store of 1: every writer stores
// the same value): the tiles' kernel then merges every row by one thread.
template <typename I>
__global__ void k_mt_bounds(int64_t m, int64_t nnzA, const I* __restrict__ pA, int64_t nnzB,
                            const I* __restrict__ pB, int64_t ntiles, int64_t* __restrict__ rec,
                            int* __restrict__ flag) {
    auto pa = [&](int64_t r) { return dclamp((int64_t)pA[r] - 1, 0, nnzA); };
    auto pb = [&](int64_t r) { return dclamp((int64_t)pB[r] - 1, 0, nnzB); };
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r <= m;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ar = pa(r), br = pb(r);
        const int64_t g = ar + br + 2 * r;
        int64_t tlo = 0, ap = ar, bp = br;
        if (r > 0) {
            ap = pa(r - 1);
            bp = pb(r - 1);
            if (ar < ap || br < bp) *(volatile int*)flag = 1;
            tlo = (ap + bp + 2 * (r - 1)) / MT_TILE + 1;
        }
        int64_t thi = g / MT_TILE;
        thi = thi < ntiles - 1 ? thi : ntiles - 1;
        for (int64_t t = tlo; t <= thi; ++t) {
            int64_t* p = rec + MT_REC * t;
            p[0] = r;
            p[1] = ar;
            p[2] = br;
            p[3] = ap;
            p[4] = bp;
        }
    }
}

// Tile t (rows from boundary t to boundary t + 1), XCD-contiguous tile order.
template <typename T, typename I, bool RUNS, bool FILL>
__global__ __launch_bounds__(MT_WG) void k_merge_tiles(
    int64_t m, int64_t n, int64_t ntiles, const int64_t* __restrict__ rec,
    const int* __restrict__ flag, int64_t nnzA, const I* __restrict__ pA,
    const I* __restrict__ cA, const T* __restrict__ vA, int64_t nnzB, const I* __restrict__ pB,
    const I* __restrict__ cB, const T* __restrict__ vB, I* __restrict__ ptrC,
    I* __restrict__ cC, T* __restrict__ vC) {
    __shared__ uint64_t s_k[MT_CAP + 2];
    __shared__ T s_v[MT_CAP + 2];
    // the staged row starts (A's, then B's), or a long row's chunk splits
    __shared__ __align__(8) unsigned char s_aux[MT_AUX];
    __shared__ int lds8[MT_WG / 64];
    uint16_t* rsA = reinterpret_cast<uint16_t*>(s_aux);
    uint16_t* rsB = rsA + MT_ROWS + 1;
    int64_t* s_split = reinterpret_cast<int64_t*>(s_aux);
    const int tid = threadIdx.x;
    auto pa = [&](int64_t r) { return dclamp((int64_t)pA[r] - 1, 0, nnzA); };
    auto pb = [&](int64_t r) { return dclamp((int64_t)pB[r] - 1, 0, nnzB); };
    auto key = [&](I c) -> uint64_t {
        return RUNS ? (uint64_t)dclamp((int64_t)c - 1, 0, n - 1) : (uint64_t)(int64_t)c;
    };
    auto serial_global = [&](int64_t r) {
        const int64_t a = pa(r), b = pb(r);
        const int64_t ea = pa(r + 1) > a ? pa(r + 1) : a, eb = pb(r + 1) > b ? pb(r + 1) : b;
        merge_row_serial<T, I, RUNS, FILL>(
            r, n, a, ea, b, eb, [&](int64_t k) { return cA[k]; },
            [&](int64_t k) { return vA[k]; }, [&](int64_t k) { return cB[k]; },
            [&](int64_t k) { return vB[k]; }, ptrC, FILL ? (int64_t)ptrC[r] - 1 : 0, cC, vC);
    };
    if (*flag) {  // pointer arrays that decrease: one row per thread
        for (int64_t r = blockIdx.x * (int64_t)MT_WG + tid; r < m; r += (int64_t)gridDim.x * MT_WG)
            serial_global(r);
        return;
    }
    const int64_t t = xcd_tile(blockIdx.x, ntiles);
    if (t >= ntiles) return;
    // boundaries t and t + 1: rows [lo, hi); the entries of row hi - 1, the only
    // row that may be long, start at (qa, qb) and end at (ae, be)
    const int64_t* b0 = rec + MT_REC * t;
    const int64_t* b1 = b0 + MT_REC;
    const int64_t lo = dclamp(b0[0], 0, m), hi = dclamp(b1[0], lo, m);
    if (lo >= hi) return;
    const int64_t a_s = b0[1], b_s = b0[2], ae = b1[1], be = b1[2], qa = b1[3], qb = b1[4];
    const int64_t q = hi - 1;
    const bool has_long = (ae - qa) + (be - qb) >= MT_TILE;
    const int64_t qe = has_long ? q : hi;
    const int64_t wlo = FILL ? (int64_t)ptrC[lo] - 1 : 0;  // read now, used after the merge

    // A run's (or a chunk's) entries: nA of A from position ga and nB of B from
    // gb, staged in slots 1 .. nA + nB.  The count variant reads them all in one
    // batch (thread tid takes entries tid + u * MT_WG of A's part followed by
    // B's: load, then put); the fill variant, which needs more registers for its
    // outputs, stores each as it arrives (stage), which keeps its occupancy.
    I pc[MT_CH];
    T pv[MT_CH];
    auto stage = [&](int64_t ga, int nA, int64_t gb, int nB) {
        for (int k = tid; k < nA; k += MT_WG) {
            s_k[1 + k] = key(cA[ga + k]);
            s_v[1 + k] = vA[ga + k];
        }
        for (int k = tid; k < nB; k += MT_WG) {
            s_k[1 + nA + k] = key(cB[gb + k]);
            s_v[1 + nA + k] = vB[gb + k];
        }
    };
    auto load = [&](int64_t ga, int nA, int64_t gb, int nB) {
#pragma unroll
        for (int u = 0; u < MT_CH; ++u) {
            const int x = tid + u * MT_WG;
            if (x < nA) {
                pc[u] = cA[ga + x];
                pv[u] = vA[ga + x];
            } else if (x < nA + nB) {
                pc[u] = cB[gb + (x - nA)];
                pv[u] = vB[gb + (x - nA)];
            }
        }
    };
    auto put = [&](int nA, int nB) {
#pragma unroll
        for (int u = 0; u < MT_CH; ++u) {
            const int x = tid + u * MT_WG;
            if (x < nA + nB) {
                s_k[1 + x] = key(pc[u]);
                s_v[1 + x] = pv[u];
            }
        }
    };

    if (qe > lo) {  // the short rows [lo, qe), in runs
        const int nrows = (int)(qe - lo);
        // when one run holds all the short rows (most tiles), their entries are
        // read in the same batch as the row pointers
        const int64_t nA1 = (has_long ? qa : ae) - a_s, nB1 = (has_long ? qb : be) - b_s;
        const bool one = nA1 + nB1 <= MT_CAP;
        constexpr int RS = (MT_ROWS + 1 + MT_WG - 1) / MT_WG;
        I pra[RS], prb[RS];
#pragma unroll
        for (int u = 0; u < RS; ++u) {
            const int k = tid + u * MT_WG;
            if (k <= nrows) {
                pra[u] = pA[lo + k];
                prb[u] = pB[lo + k];
            }
        }
        if (!FILL && one) load(a_s, (int)nA1, b_s, (int)nB1);
#pragma unroll
        for (int u = 0; u < RS; ++u) {
            const int k = tid + u * MT_WG;
            if (k <= nrows) {
                rsA[k] = (uint16_t)(dclamp((int64_t)pra[u] - 1, 0, nnzA) - a_s);
                rsB[k] = (uint16_t)(dclamp((int64_t)prb[u] - 1, 0, nnzB) - b_s);
            }
        }
        __syncthreads();
        for (int r0 = 0; r0 < nrows;) {
            // rows [r0, r1): the most whole rows from r0 whose entries fit MT_CAP
            const int f0 = rsA[r0] + rsB[r0];
            int r1 = r0 + 1, rh = nrows;
            while (r1 < rh) {
                const int mid = (r1 + rh + 1) >> 1;
                if (rsA[mid] + rsB[mid] - f0 <= MT_CAP) r1 = mid;
                else rh = mid - 1;
            }
            const int aoff = rsA[r0], boff = rsB[r0];
            const int nA = rsA[r1] - aoff, nB = rsB[r1] - boff, tot = nA + nB, sB = 1 + nA;
            const int64_t w0 = r0 == 0 ? wlo : (FILL ? (int64_t)ptrC[lo + r0] - 1 : 0);
            if (FILL) {
                stage(a_s + aoff, nA, b_s + boff, nB);
            } else {
                if (!(one && r0 == 0)) load(a_s + aoff, nA, b_s + boff, nB);
                put(nA, nB);
            }
            if (tid == 0) {  // no entries before A's part or after B's
                s_k[0] = MB_NONE;
                s_v[0] = T(0);
                s_k[sB + nB] = MB_NONE;
                s_v[sB + nB] = T(0);
            }
            __syncthreads();
            // rows, packed keys and the canonical check, as in k_merge_blocks
            int bad = 0;
            {
                int row = r0, part = -1;
                uint64_t prev = 0;
                for (int u = 0; u < MT_CH; ++u) {
                    const int x = tid * MT_CH + u;
                    if (x >= tot) break;
                    const int inA = x < nA;
                    const int k = inA ? x : x - nA;
                    const int kt = k + (inA ? aoff : boff);
                    const uint16_t* rs = inA ? rsA : rsB;
                    if (inA != part) {
                        row = mb_row_of(rs, r0, r1, kt);
                        part = inA;
                        prev = k > 0 ? (s_k[x] & MB_KMASK) : 0;  // slot x = entry x - 1
                    } else if (row + 1 < r1 && rs[row + 1] <= kt) {
                        row = mb_row_of(rs, row + 1, r1, kt);
                    }
                    const uint64_t kc = s_k[1 + x];
                    bad |= kc > MB_KMASK;
                    bad |= k > 0 && rs[row] < kt && prev >= kc;  // entry k - 1 in the same row
                    prev = kc;
                    s_k[1 + x] = ((uint64_t)row << MB_KB) | (kc & MB_KMASK);
                }
            }
            if (__syncthreads_or(bad)) {  // every row of the run by one thread
                for (int64_t r = lo + r0 + tid; r < lo + r1; r += MT_WG) serial_global(r);
            } else {
                const int d0 = tid * MT_CH < tot ? tid * MT_CH : tot;
                const int d1 = d0 + MT_CH < tot ? d0 + MT_CH : tot;
                I oc[MT_CH];
                T ov[MT_CH];
                const unsigned keep =
                    mb_lane_merge<T, I, RUNS, MT_CH>(nA, nB, d0, d1, s_k, s_v, oc, ov);
                int total = 0;
                const int base = mt_exclusive_sum(__popc(keep), lds8, &total);
                if (!FILL) {
                    int* s_pos = reinterpret_cast<int*>(s_k);
                    for (int u = 0; u < MT_CH; ++u)
                        if (d0 + u < d1) s_pos[d0 + u] = base + __popc(keep & ((1u << u) - 1));
                    if (tid == 0) s_pos[tot] = total;
                    __syncthreads();
                    for (int k = r0 + tid; k < r1; k += MT_WG) {
                        const int g0 = rsA[k] + rsB[k] - f0, g1 = rsA[k + 1] + rsB[k + 1] - f0;
                        ptrC[lo + k + 1] = (I)(s_pos[g1] - s_pos[g0]);
                    }
                } else {
                    I* s_oc = reinterpret_cast<I*>(s_k);
                    int w = base;
                    for (int u = 0; u < MT_CH; ++u)
                        if ((keep >> u) & 1) {
                            s_oc[w] = oc[u];
                            s_v[w] = ov[u];
                            ++w;
                        }
                    __syncthreads();
                    for (int k = tid; k < total; k += MT_WG) {
                        cC[w0 + k] = s_oc[k];
                        vC[w0 + k] = s_v[k];
                    }
                }
            }
            __syncthreads();  // before the next run (or the long row) reuses the LDS
            r0 = r1;
        }
    }

    if (has_long) {  // row q, in chunks of MT_CAP output-order positions
        const int64_t a = qa, b = qb, na = ae - qa, nb = be - qb;
        int bad = 0;
        for (int64_t k = tid; k < na; k += MT_WG) {
            const uint64_t kc = key(cA[a + k]);
            bad |= kc > MB_KMASK || (k > 0 && key(cA[a + k - 1]) >= kc);
        }
        for (int64_t k = tid; k < nb; k += MT_WG) {
            const uint64_t kc = key(cB[b + k]);
            bad |= kc > MB_KMASK || (k > 0 && key(cB[b + k - 1]) >= kc);
        }
        if (__syncthreads_or(bad)) {
            if (tid == 0) serial_global(q);
            return;
        }
        const int64_t L = na + nb, nch = (L + MT_CAP - 1) / MT_CAP;
        const int64_t w0 = FILL ? (int64_t)ptrC[q] - 1 : 0;
        int64_t run = 0;
        // chunk boundaries c0 .. c0 + MT_WG - 1, then the chunks between them
        for (int64_t c0 = 0; c0 < nch; c0 += MT_WG - 1) {
            const int64_t c = c0 + tid;
            if (c <= nch) {
                const int64_t d = c * MT_CAP < L ? c * MT_CAP : L;
                int64_t lo2 = d - nb > 0 ? d - nb : 0, hi2 = d < na ? d : na;
                while (lo2 < hi2) {
                    const int64_t mid = (lo2 + hi2) >> 1;
                    if (key(cA[a + mid]) <= key(cB[b + d - 1 - mid])) lo2 = mid + 1;
                    else hi2 = mid;
                }
                s_split[tid] = lo2;
            }
            __syncthreads();
            const int64_t cend = c0 + MT_WG - 1 < nch ? c0 + MT_WG - 1 : nch;
            for (int64_t cc = c0; cc < cend; ++cc) {
                const int64_t d0 = cc * MT_CAP, d1 = d0 + MT_CAP < L ? d0 + MT_CAP : L;
                const int64_t i0 = s_split[cc - c0], i1 = s_split[cc - c0 + 1];
                const int64_t j0 = d0 - i0, j1 = d1 - i1;
                const int nAc = (int)(i1 - i0), nBc = (int)(j1 - j0), sB = 1 + nAc;
                if (FILL) {  // row 0: the key is the column key
                    stage(a + i0, nAc, b + j0, nBc);
                } else {
                    load(a + i0, nAc, b + j0, nBc);
                    put(nAc, nBc);
                }
                if (tid == 0) {  // the row's entries just outside this chunk
                    const bool pre = i0 > 0, post = j1 < nb;
                    s_k[0] = pre ? key(cA[a + i0 - 1]) : MB_NONE;
                    s_v[0] = pre ? vA[a + i0 - 1] : T(0);
                    s_k[sB + nBc] = post ? key(cB[b + j1]) : MB_NONE;
                    s_v[sB + nBc] = post ? vB[b + j1] : T(0);
                }
                __syncthreads();
                const int tot = nAc + nBc;
                const int e0 = tid * MT_CH < tot ? tid * MT_CH : tot;
                const int e1 = e0 + MT_CH < tot ? e0 + MT_CH : tot;
                I oc[MT_CH];
                T ov[MT_CH];
                const unsigned keep =
                    mb_lane_merge<T, I, RUNS, MT_CH>(nAc, nBc, e0, e1, s_k, s_v, oc, ov);
                int total = 0;
                const int base = mt_exclusive_sum(__popc(keep), lds8, &total);
                if (FILL) {
                    I* s_oc = reinterpret_cast<I*>(s_k);
                    int w = base;
                    for (int u = 0; u < MT_CH; ++u)
                        if ((keep >> u) & 1) {
                            s_oc[w] = oc[u];
                            s_v[w] = ov[u];
                            ++w;
                        }
                    __syncthreads();
                    for (int k = tid; k < total; k += MT_WG) {
                        cC[w0 + run + k] = s_oc[k];
                        vC[w0 + run + k] = s_v[k];
                    }
                }
                run += total;
                __syncthreads();  // before the next chunk (or split batch) reuses the LDS
            }
        }
        if (!FILL && tid == 0) ptrC[q + 1] = (I)run;
    }
}

// mean entries per row (A's and B's) up to which the row blocks merge; above it
// the tiles
constexpr int64_t MB_MEAN_MAX = 12;

// The row merge of A + B over m rows (n: the column count, for RUNS' clamp).
template <typename T, typename I, bool RUNS, bool FILL>
int merge_rows(hipStream_t s, int64_t m, int64_t n, int64_t nnzA, const I* pA, const I* cA,
               const T* vA, int64_t nnzB, const I* pB, const I* cB, const T* vB, I* ptrC,
               I* cC, T* vC) {
    if (m <= 0) return GSA_OK;
    const int64_t per_row = (nnzA + nnzB + m - 1) / m;
    if (per_row <= MB_MEAN_MAX) {
        // rows per block: about MB_CAP entries at the mean row length
        int64_t rpb = per_row > 0 ? MB_CAP / per_row : MB_WG;
        rpb = rpb < 1 ? 1 : (rpb > MB_WG ? MB_WG : rpb);
        hipLaunchKernelGGL((k_merge_blocks<T, I, RUNS, FILL>), dim3((unsigned)cdiv(m, rpb)),
                           dim3(MB_WG), 0, s, m, n, (int)rpb, nnzA, pA, cA, vA, nnzB, pB, cB, vB,
                           ptrC, cC, vC);
    } else {
        const int64_t ntiles = (nnzA + nnzB + 2 * m) / MT_TILE + 1;
        Scratch R(s);
        GSA_TRY(R.alloc(sizeof(int64_t) * (MT_REC * (ntiles + 1) + 1)));
        int64_t* rec = R.as<int64_t>();
        int* flag = reinterpret_cast<int*>(rec + MT_REC * (ntiles + 1));
        hipLaunchKernelGGL((k_mt_init<I>), dim3(grid_for(ntiles + 1, 256)), dim3(256), 0, s, m,
                           nnzA, pA, nnzB, pB, ntiles, rec, flag);
        hipLaunchKernelGGL((k_mt_bounds<I>), dim3(grid_for(m + 1, 256)), dim3(256), 0, s, m,
                           nnzA, pA, nnzB, pB, ntiles, rec, flag);
        hipLaunchKernelGGL((k_merge_tiles<T, I, RUNS, FILL>),
                           dim3((unsigned)xcd_tile_grid(ntiles)), dim3(MT_WG), 0, s, m, n, ntiles,
                           (const int64_t*)rec, (const int*)flag, nnzA, pA, cA, vA, nnzB, pB, cB,
                           vB, ptrC, cC, vC);
    }
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

}  // namespace gsa
