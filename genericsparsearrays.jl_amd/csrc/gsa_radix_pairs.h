// gsa_radix_pairs.h — stable LSD radix sort that moves each entry's payload
// (its minor index and value) with its key: the sort behind the
// compressed-format conversions (conversions.jl:85-86, :171-212, :282-323,
// whose sortperm is stable, ext/GenericSparseArraysJLArraysExt.jl:6).
//
// A key/position sort followed by a gather by position reads every payload from
// a random line at the end.  Here the payload travels through every pass: each
// pass reads keys and payloads in storage order and writes each tile's digit
// runs contiguously, so every byte moves in streams.  The first pass computes
// its keys from the one-based index arrays and reads the payload from the
// separate minor-index and value arrays; the last writes the separate output
// arrays.  Up to 8 bits per pass (keys below 2^20, 1M columns, take three), in
// 2048-key tiles: three workgroups fit in one CU's LDS.  Measured on the 1M-row
// power-law matrix (Float64, Int64), CSR -> CSC: 472 us, against 621 us for the
// key/position sort and gather, 782 us for 10-bit digits in two passes (runs of
// two entries per digit and tile), 709 us for 4096-key tiles (one per CU).
//
// Per pass: k_rp_hist counts each tile's digits; a scan turns the digit-major
// [digits x tiles] counts into offsets; k_rp_scatter ranks each wavefront's keys
// stably with ballots (a per-wavefront counter per digit carries the rank from
// one 64-key round to the next), reorders the tile by digit in LDS and writes
// each digit run contiguously.  Equal keys keep their input order.
#pragma once
#include "gsa_scan.h"
#include "gsa_rows.h"

namespace gsa {

constexpr int RP_WG = 256;
constexpr int RP_WAVES = RP_WG / 64;
constexpr int RP_IPT = 8;                      // keys per thread
constexpr int RP_TILE = RP_WG * RP_IPT;        // 2048 keys per tile
constexpr int RP_WTILE = 64 * RP_IPT;          // 512 keys per wavefront
constexpr int RP_MAXBITS = 8;
constexpr int RP_BINS = 1 << RP_MAXBITS;
constexpr int RP_BPT = RP_BINS / RP_WG;        // digits per thread in the offset step
static_assert(RP_BINS % RP_WG == 0, "digit bins must be a multiple of the workgroup");

// (minor index, value) of one stored entry, side by side.
template <typename T, typename I>
struct alignas(sizeof(T) > sizeof(I) ? 2 * sizeof(T) : 2 * sizeof(I)) MinorVal {
    I minor;
    T val;
};

// Key of entry i from one-based indices: (major-1)*kdim + (minor-1), or major-1
// when minor is null (kdim = 1); indices are clamped into range.
template <typename K, typename I>
struct KeyFromIdx {
    const I* major;
    const I* minor;
    int64_t major_dim, kdim;
    __device__ __forceinline__ K operator()(int64_t i) const {
        const int64_t a = dclamp((int64_t)major[i] - 1, 0, major_dim - 1);
        const int64_t b = minor ? dclamp((int64_t)minor[i] - 1, 0, kdim - 1) : 0;
        return (K)((uint64_t)a * (uint64_t)kdim + (uint64_t)b);
    }
};
template <typename K>
struct KeyArr {
    const K* p;
    __device__ __forceinline__ K operator()(int64_t i) const { return p[i]; }
};
template <typename T, typename I>
struct PaySplit {
    const I* minor;
    const T* val;
    __device__ __forceinline__ MinorVal<T, I> operator()(int64_t i) const {
        MinorVal<T, I> m;
        m.minor = minor ? minor[i] : I(0);
        m.val = val[i];
        return m;
    }
};
template <typename P>
struct PayArr {
    const P* p;
    __device__ __forceinline__ P operator()(int64_t i) const { return p[i]; }
};
template <typename T, typename I>
struct SinkSplit {
    I* minor;
    T* val;
    __device__ __forceinline__ void operator()(int64_t i, const MinorVal<T, I>& m) const {
        minor[i] = m.minor;
        val[i] = m.val;
    }
};
template <typename P>
struct SinkArr {
    P* p;
    __device__ __forceinline__ void operator()(int64_t i, const P& m) const { p[i] = m; }
};

template <typename K>
__device__ __forceinline__ int rp_digit(K k, int shift, int nbits) {
    return (int)((k >> shift) & (K)((1u << nbits) - 1));
}

// Per-tile digit counts, digit-major: hist[d * ntiles + tile], d < 2^nbits.
template <typename K, typename KeySrc>
__global__ __launch_bounds__(RP_WG) void k_rp_hist(KeySrc key, int64_t n, int shift, int nbits,
                                                   int64_t* __restrict__ hist, int64_t ntiles) {
    __shared__ int cnt[RP_WAVES][RP_BINS];
    const int tid = threadIdx.x, wave = tid >> 6;
    const int nbins = 1 << nbits;
    for (int d = tid; d < nbins; d += RP_WG)
        for (int w = 0; w < RP_WAVES; ++w) cnt[w][d] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RP_TILE;
    K k[RP_IPT];
#pragma unroll
    for (int i = 0; i < RP_IPT; ++i) {
        const int64_t idx = base + (int64_t)i * RP_WG + tid;
        k[i] = idx < n ? key(idx) : (K)0;
    }
#pragma unroll
    for (int i = 0; i < RP_IPT; ++i) {
        const int64_t idx = base + (int64_t)i * RP_WG + tid;
        if (idx < n) atomicAdd(&cnt[wave][rp_digit(k[i], shift, nbits)], 1);
    }
    __syncthreads();
    for (int d = tid; d < nbins; d += RP_WG) {
        int c = 0;
        for (int w = 0; w < RP_WAVES; ++w) c += cnt[w][d];
        hist[(int64_t)d * ntiles + blockIdx.x] = c;
    }
}

// Orders one wavefront's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void rp_wave_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive sum over the 256 threads of one workgroup.
__device__ __forceinline__ int rp_block_excl_sum(int v, int* wsum) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int pre = 0;
    for (int w = 0; w < wave; ++w) pre += wsum[w];
    return pre + x - v;
}

// Where the first pass takes each entry's minor index from: the payload source
// (NoRows), or, for a CSR-to-CSC transposition, the row each entry lies in,
// found per tile from the one-based row pointer (TileRows; tile_rows holds the
// rows of each tile's first and last entries, from k_tile_rows).
struct NoRows {
    static constexpr bool active = false;
};
template <typename I>
struct TileRows {
    static constexpr bool active = true;
    const I* ptr;
    const int64_t* tile_rows;
};

// Rows (zero-based) of the first and last entry of every RP_TILE-entry tile:
// the last row r with ptr[r] - 1 <= position, by binary search.
template <typename I>
__global__ void k_tile_rows(int64_t m, int64_t nnz, const I* __restrict__ ptr, int64_t ntiles,
                            int64_t* __restrict__ out) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int64_t first = t * RP_TILE;
    const int64_t last = (first + RP_TILE < nnz ? first + RP_TILE : nnz) - 1;
    auto row_of = [&](int64_t pos) {
        int64_t lo = 0, hi = m - 1;  // the answer lies in [lo, hi]
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if ((int64_t)ptr[mid] - 1 <= pos) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    out[2 * t] = row_of(first);
    out[2 * t + 1] = row_of(last);
}

// Exclusive max over the 256 threads of one workgroup (values >= 0).
__device__ __forceinline__ uint32_t rp_block_excl_max(uint32_t v, uint32_t* wmax) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x = x > y ? x : y;
    }
    if (lane == 63) wmax[wave] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (int w = 0; w < wave; ++w) pre = pre > wmax[w] ? pre : wmax[w];
    const uint32_t excl = __shfl_up(x, 1, 64);
    const uint32_t inwave = lane == 0 ? 0u : excl;
    return pre > inwave ? pre : inwave;
}

template <typename K, typename P, typename KeySrc, typename PaySrc, typename Sink,
          typename RowSrc = NoRows>
__global__ __launch_bounds__(RP_WG) void k_rp_scatter(
    KeySrc key, PaySrc pay, int64_t n, int shift, int nbits, const int64_t* __restrict__ incl,
    const int64_t* __restrict__ hist, int64_t ntiles, K* __restrict__ keys_out, Sink sink,
    RowSrc rows = RowSrc{}) {
    __shared__ int wcnt[RP_WAVES][RP_BINS];   // per-wavefront digit counts, then offsets
    __shared__ int tile_start[RP_BINS];
    __shared__ int wsum[RP_WAVES];
    __shared__ int64_t gbase[RP_BINS];
    __shared__ K s_keys[RP_TILE];
    __shared__ P s_pay[RP_TILE];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nbins = 1 << nbits;
    const int64_t tile = xcd_tile(blockIdx.x, ntiles);
    if (tile >= ntiles) return;
    const int64_t base = tile * RP_TILE;
    const int tile_n = (int)(n - base < RP_TILE ? n - base : RP_TILE);
    for (int d = tid; d < nbins; d += RP_WG) {
        const int64_t at = (int64_t)d * ntiles + tile;
        gbase[d] = incl[at] - hist[at];  // exclusive global offset of this tile's run
        for (int w = 0; w < RP_WAVES; ++w) wcnt[w][d] = 0;
    }
    __syncthreads();

    // TileRows: each entry's row, relative to the tile's first row, in LDS (the
    // payload staging area, unused until the ranking is done): every row that
    // starts inside the tile marks its first slot (the largest row wins where
    // empty rows share a slot), then an inclusive max-scan fills the rest
    int64_t row0 = 0;
    uint32_t* rid = reinterpret_cast<uint32_t*>(s_pay);
    if constexpr (RowSrc::active) {
        row0 = rows.tile_rows[2 * tile];
        const int64_t row1 = rows.tile_rows[2 * tile + 1];
        for (int p = tid; p < tile_n; p += RP_WG) rid[p] = 0u;
        __syncthreads();
        for (int64_t r = row0 + 1 + tid; r <= row1; r += RP_WG) {
            const int64_t st = (int64_t)rows.ptr[r] - 1 - base;
            if (st > 0 && st < tile_n) atomicMax(&rid[st], (uint32_t)(r - row0));
        }
        __syncthreads();
        constexpr int PER = RP_TILE / RP_WG;
        uint32_t mx = 0;
        for (int q = 0; q < PER; ++q) {
            const int p = tid * PER + q;
            if (p < tile_n) mx = mx > rid[p] ? mx : rid[p];
        }
        uint32_t run = rp_block_excl_max(mx, reinterpret_cast<uint32_t*>(wsum));
        for (int q = 0; q < PER; ++q) {
            const int p = tid * PER + q;
            if (p < tile_n) {
                run = run > rid[p] ? run : rid[p];
                rid[p] = run;
            }
        }
        __syncthreads();
    }

    // this wavefront's entries: tile positions wave*RP_WTILE + round*64 + lane
    K k[RP_IPT];
    P v[RP_IPT];
#pragma unroll
    for (int i = 0; i < RP_IPT; ++i) {
        const int pos = wave * RP_WTILE + i * 64 + lane;
        if (pos < tile_n) {
            k[i] = key(base + pos);
            v[i] = pay(base + pos);
            if constexpr (RowSrc::active) v[i].minor = (decltype(v[i].minor))(row0 + rid[pos] + 1);
        }
    }
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int* my = wcnt[wave];
    int rank[RP_IPT];
#pragma unroll
    for (int i = 0; i < RP_IPT; ++i) {
        const int pos = wave * RP_WTILE + i * 64 + lane;
        const bool valid = pos < tile_n;
        const int d = valid ? rp_digit(k[i], shift, nbits) : 0;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < nbits; ++b) {
            const uint64_t bal = __ballot(valid && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? bal : ~bal;
        }
        // every lane reads its digit's count before the round's leader adds to
        // it: one wavefront, LDS operations in program order
        const int before = valid ? my[d] : 0;
        rank[i] = before + __popcll(peers & lt_mask);
        const int leader = valid ? (__ffsll((long long)peers) - 1) : -1;
        rp_wave_order();
        if (valid && lane == leader) my[d] = before + __popcll(peers);
        rp_wave_order();
    }
    __syncthreads();
    // thread tid owns digits [tid*RP_BPT, (tid+1)*RP_BPT): offsets of each
    // wavefront's run inside the digit, then the digit's start in the tile
    int own[RP_BPT];
    int mine = 0;
#pragma unroll
    for (int q = 0; q < RP_BPT; ++q) {
        const int d = tid * RP_BPT + q;
        int tot = 0;
        if (d < nbins) {
            for (int w = 0; w < RP_WAVES; ++w) {
                const int c = wcnt[w][d];
                wcnt[w][d] = tot;
                tot += c;
            }
        }
        own[q] = tot;
        mine += tot;
    }
    int start = rp_block_excl_sum(mine, wsum);
#pragma unroll
    for (int q = 0; q < RP_BPT; ++q) {
        const int d = tid * RP_BPT + q;
        if (d < nbins) tile_start[d] = start;
        start += own[q];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RP_IPT; ++i) {
        const int pos = wave * RP_WTILE + i * 64 + lane;
        if (pos < tile_n) {
            const int d = rp_digit(k[i], shift, nbits);
            const int p = tile_start[d] + my[d] + rank[i];
            s_keys[p] = k[i];
            s_pay[p] = v[i];
        }
    }
    __syncthreads();
    for (int p = tid; p < tile_n; p += RP_WG) {
        const K kk = s_keys[p];
        const int d = rp_digit(kk, shift, nbits);
        const int64_t dst = gbase[d] + (p - tile_start[d]);
        keys_out[dst] = kk;
        sink(dst, s_pay[p]);
    }
}

// Stable sort of n entries on the low `bits` bits of their keys (key0 over the
// index arrays).  keys_final receives the sorted keys, minor_out and val_out the
// entries' minor indices and values in sorted order.  The minor indices come
// from minor_in, or, when rows_ptr is given (m_rows < 2^31 rows, one-based
// pointer over the n entries in order), are the rows the entries lie in.
// Inputs and outputs must not overlap.
template <typename K, typename T, typename I>
int radix_sort_payload(hipStream_t s, int64_t n, int bits, KeyFromIdx<K, I> key0,
                       const I* minor_in, const T* val_in, K* keys_final, I* minor_out,
                       T* val_out, const I* rows_ptr = nullptr, int64_t m_rows = 0) {
    using P = MinorVal<T, I>;
    if (n <= 0) return GSA_OK;
    GSA_CHECK(bits >= 0 && bits <= (int)(8 * sizeof(K)), GSA_ERR_INVALID,
              "radix sort: %d bits in a %d-bit key", bits, (int)(8 * sizeof(K)));
    const int64_t ntiles = cdiv(n, RP_TILE);
    const int npass = bits == 0 ? 1 : (bits + RP_MAXBITS - 1) / RP_MAXBITS;
    Scratch hist(s), incl(s), kA(s), kB(s), pA(s), pB(s);
    GSA_TRY(hist.alloc(sizeof(int64_t) * RP_BINS * ntiles));
    GSA_TRY(incl.alloc(sizeof(int64_t) * RP_BINS * ntiles));
    if (npass >= 2) {
        GSA_TRY(kA.alloc(sizeof(K) * n));
        GSA_TRY(pA.alloc(sizeof(P) * n));
    }
    if (npass >= 3) {
        GSA_TRY(kB.alloc(sizeof(K) * n));
        GSA_TRY(pB.alloc(sizeof(P) * n));
    }
    GSA_CHECK(rows_ptr == nullptr || (m_rows > 0 && m_rows < ((int64_t)1 << 31)), GSA_ERR_INVALID,
              "radix sort: rows from a pointer need 0 < m < 2^31");
    Scratch trows(s);
    if (rows_ptr) {
        GSA_TRY(trows.alloc(sizeof(int64_t) * 2 * ntiles));
        hipLaunchKernelGGL((k_tile_rows<I>), dim3((unsigned)cdiv(ntiles, 256)), dim3(256), 0, s,
                           m_rows, n, rows_ptr, ntiles, trows.as<int64_t>());
    }
    const TileRows<I> rsrc{rows_ptr, trows.as<int64_t>()};
    const PaySplit<T, I> pay0{rows_ptr ? nullptr : minor_in, val_in};
    const SinkSplit<T, I> sink_out{minor_out, val_out};
    const unsigned grid = (unsigned)xcd_tile_grid(ntiles);
    const K* kin = nullptr;
    const P* pin = nullptr;
    int shift = 0;
    for (int p = 0; p < npass; ++p) {
        const int nb = bits / npass + (p < bits % npass ? 1 : 0);
        const bool last = p == npass - 1;
        K* kout = last ? keys_final : (p % 2 == 0 ? kA.as<K>() : kB.as<K>());
        P* pout = last ? nullptr : (p % 2 == 0 ? pA.as<P>() : pB.as<P>());
        const KeyArr<K> kprev{kin};
        const PayArr<P> pprev{pin};
        const SinkArr<P> sink_mid{pout};
        if (p == 0)
            hipLaunchKernelGGL((k_rp_hist<K, KeyFromIdx<K, I>>), dim3((unsigned)ntiles), dim3(RP_WG),
                               0, s, key0, n, shift, nb, hist.as<int64_t>(), ntiles);
        else
            hipLaunchKernelGGL((k_rp_hist<K, KeyArr<K>>), dim3((unsigned)ntiles), dim3(RP_WG), 0, s,
                               kprev, n, shift, nb, hist.as<int64_t>(), ntiles);
        GSA_TRY(inclusive_scan<int64_t>(s, hist.as<int64_t>(), incl.as<int64_t>(),
                                        ((int64_t)1 << nb) * ntiles, 0));
        const int64_t* ic = incl.as<int64_t>();
        const int64_t* hc = hist.as<int64_t>();
        if (p == 0 && last && rows_ptr)
            hipLaunchKernelGGL((k_rp_scatter<K, P, KeyFromIdx<K, I>, PaySplit<T, I>, SinkSplit<T, I>,
                                             TileRows<I>>),
                               dim3(grid), dim3(RP_WG), 0, s, key0, pay0, n, shift, nb, ic, hc,
                               ntiles, kout, sink_out, rsrc);
        else if (p == 0 && rows_ptr)
            hipLaunchKernelGGL((k_rp_scatter<K, P, KeyFromIdx<K, I>, PaySplit<T, I>, SinkArr<P>,
                                             TileRows<I>>),
                               dim3(grid), dim3(RP_WG), 0, s, key0, pay0, n, shift, nb, ic, hc,
                               ntiles, kout, sink_mid, rsrc);
        else if (p == 0 && last)
            hipLaunchKernelGGL((k_rp_scatter<K, P, KeyFromIdx<K, I>, PaySplit<T, I>, SinkSplit<T, I>>),
                               dim3(grid), dim3(RP_WG), 0, s, key0, pay0, n, shift, nb, ic, hc,
                               ntiles, kout, sink_out, NoRows{});
        else if (p == 0)
            hipLaunchKernelGGL((k_rp_scatter<K, P, KeyFromIdx<K, I>, PaySplit<T, I>, SinkArr<P>>),
                               dim3(grid), dim3(RP_WG), 0, s, key0, pay0, n, shift, nb, ic, hc,
                               ntiles, kout, sink_mid, NoRows{});
        else if (last)
            hipLaunchKernelGGL((k_rp_scatter<K, P, KeyArr<K>, PayArr<P>, SinkSplit<T, I>>),
                               dim3(grid), dim3(RP_WG), 0, s, kprev, pprev, n, shift, nb, ic, hc,
                               ntiles, kout, sink_out, NoRows{});
        else
            hipLaunchKernelGGL((k_rp_scatter<K, P, KeyArr<K>, PayArr<P>, SinkArr<P>>), dim3(grid),
                               dim3(RP_WG), 0, s, kprev, pprev, n, shift, nb, ic, hc, ntiles, kout,
                               sink_mid, NoRows{});
        GSA_HIP(hipGetLastError());
        kin = kout;
        pin = pout;
        shift += nb;
    }
    return GSA_OK;
}

}  // namespace gsa
