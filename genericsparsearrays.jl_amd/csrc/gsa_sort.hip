// gsa_sort.hip — stable LSD radix sort of (uint64 key, payload) pairs.
//
// This is the sortperm primitive of the conversions, COO +, and the
// expand-sort-compress SpGEMM (_sortperm_AK, src/helpers.jl:2; Base's stable
// semantics, ext/GenericSparseArraysJLArraysExt.jl:6): equal keys keep their
// input order, so the permutation equals Base.sortperm's.
//
// One 8-bit digit per pass, only over the bits the keys use.  Per pass:
//   hist    — each 4096-key tile counts its digits (LDS atomics),
//   scan    — exclusive prefix over the digit-major [256 x tiles] counts,
//   scatter — each tile ranks its keys stably (wavefront ballots give the rank
//             among equal digits inside a wavefront; per-wave counters order the
//             wavefronts and the 16 rounds), reorders them by digit in LDS, and
//             writes each digit run contiguously.
#include "gsa_scan.h"

namespace gsa {

constexpr int RS_WG = 256;
constexpr int RS_IPT = 16;
constexpr int RS_TILE = RS_WG * RS_IPT;
constexpr int RS_BINS = 256;

__device__ __forceinline__ int digit_of(uint64_t k, int shift, int nbits) {
    return (int)((k >> shift) & ((1ull << nbits) - 1));
}

__global__ __launch_bounds__(RS_WG) void k_rs_hist(const uint64_t* __restrict__ keys, int64_t n,
                                                   int shift, int nbits,
                                                   int64_t* __restrict__ hist, int64_t ntiles) {
    __shared__ int cnt[RS_BINS];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int i = 0; i < RS_IPT; ++i) {
        int64_t idx = base + (int64_t)i * RS_WG + threadIdx.x;
        if (idx < n) atomicAdd(&cnt[digit_of(keys[idx], shift, nbits)], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

template <typename V>
__global__ __launch_bounds__(RS_WG) void k_rs_scatter(
    const uint64_t* __restrict__ keys_in, const V* __restrict__ vals_in, int64_t n, int shift,
    int nbits, const int64_t* __restrict__ incl, const int64_t* __restrict__ hist, int64_t ntiles,
    uint64_t* __restrict__ keys_out, V* __restrict__ vals_out) {
    __shared__ int wcnt[RS_WG / 64][RS_BINS];
    __shared__ int running[RS_BINS];
    __shared__ int tile_start[RS_BINS];
    __shared__ int64_t gbase[RS_BINS];
    __shared__ uint64_t s_keys[RS_TILE];
    __shared__ V s_vals[RS_TILE];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const int64_t tile_n64 = n - base;
    const int tile_n = (int)(tile_n64 < RS_TILE ? tile_n64 : RS_TILE);

    running[tid] = 0;
    {
        int64_t h = hist[(int64_t)tid * ntiles + blockIdx.x];
        gbase[tid] = incl[(int64_t)tid * ntiles + blockIdx.x] - h;  // exclusive offset
    }
    for (int w = 0; w < RS_WG / 64; ++w) wcnt[w][tid] = 0;
    __syncthreads();

    uint64_t k[RS_IPT];
    V v[RS_IPT];
    int rank[RS_IPT];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    for (int i = 0; i < RS_IPT; ++i) {
        const int pos = i * RS_WG + tid;
        const bool valid = pos < tile_n;
        int d = 0;
        if (valid) {
            k[i] = keys_in[base + pos];
            v[i] = vals_in[base + pos];
            d = digit_of(k[i], shift, nbits);
        }
        // lanes of this wavefront holding the same digit
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < nbits; ++b) {
            uint64_t bal = __ballot(valid && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? bal : ~bal;
        }
        const int r_in_wave = __popcll(peers & lt_mask);
        const int leader = valid ? (__ffsll((long long)peers) - 1) : -1;
        if (valid && lane == leader) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int r = running[d] + r_in_wave;
            for (int w = 0; w < wave; ++w) r += wcnt[w][d];
            rank[i] = r;
        }
        __syncthreads();
        {   // advance the per-digit counters, clear the wave counts
            int add = 0;
            for (int w = 0; w < RS_WG / 64; ++w) { add += wcnt[w][tid]; wcnt[w][tid] = 0; }
            running[tid] += add;
        }
        __syncthreads();
    }
    // running[] now holds the tile histogram: exclusive scan -> tile_start
    if (tid == 0) {
        int acc = 0;
        for (int b = 0; b < RS_BINS; ++b) { tile_start[b] = acc; acc += running[b]; }
    }
    __syncthreads();
    for (int i = 0; i < RS_IPT; ++i) {
        const int pos = i * RS_WG + tid;
        if (pos < tile_n) {
            int d = digit_of(k[i], shift, nbits);
            int p = tile_start[d] + rank[i];
            s_keys[p] = k[i];
            s_vals[p] = v[i];
        }
    }
    __syncthreads();
    for (int p = tid; p < tile_n; p += RS_WG) {
        uint64_t kk = s_keys[p];
        int d = digit_of(kk, shift, nbits);
        int64_t dst = gbase[d] + (p - tile_start[d]);
        keys_out[dst] = kk;
        vals_out[dst] = s_vals[p];
    }
}

template <typename V>
static int radix_sort_impl(hipStream_t s, uint64_t* keys, V* vals, int64_t n, int bits) {
    if (n <= 1 || bits <= 0) return GSA_OK;
    const int64_t ntiles = cdiv(n, RS_TILE);
    Scratch kalt(s), valt(s), hist(s), incl(s);
    GSA_TRY(kalt.alloc(sizeof(uint64_t) * n));
    GSA_TRY(valt.alloc(sizeof(V) * n));
    GSA_TRY(hist.alloc(sizeof(int64_t) * RS_BINS * ntiles));
    GSA_TRY(incl.alloc(sizeof(int64_t) * RS_BINS * ntiles));
    uint64_t* kin = keys;
    V* vin = vals;
    uint64_t* kout = kalt.as<uint64_t>();
    V* vout = valt.as<V>();
    int passes = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        int nb = bits - shift < 8 ? bits - shift : 8;
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)ntiles), dim3(RS_WG), 0, s, kin, n, shift, nb,
                           hist.as<int64_t>(), ntiles);
        GSA_TRY(inclusive_scan<int64_t>(s, hist.as<int64_t>(), incl.as<int64_t>(),
                                        RS_BINS * ntiles, 0));
        hipLaunchKernelGGL(k_rs_scatter<V>, dim3((unsigned)ntiles), dim3(RS_WG), 0, s, kin, vin, n,
                           shift, nb, (const int64_t*)incl.as<int64_t>(),
                           (const int64_t*)hist.as<int64_t>(), ntiles, kout, vout);
        GSA_HIP(hipGetLastError());
        uint64_t* tk = kin; kin = kout; kout = tk;
        V* tv = vin; vin = vout; vout = tv;
        ++passes;
    }
    if (passes & 1) {  // result is in the scratch buffers
        GSA_HIP(hipMemcpyAsync(keys, kin, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, s));
        GSA_HIP(hipMemcpyAsync(vals, vin, sizeof(V) * n, hipMemcpyDeviceToDevice, s));
    }
    return GSA_OK;
}

int radix_sort_pairs_u64(hipStream_t s, uint64_t* keys, uint64_t* vals, int64_t n, int bits) {
    return radix_sort_impl<uint64_t>(s, keys, vals, n, bits);
}
int radix_sort_pairs_u64_u32(hipStream_t s, uint64_t* keys, uint32_t* vals, int64_t n, int bits) {
    return radix_sort_impl<uint32_t>(s, keys, vals, n, bits);
}

}  // namespace gsa
