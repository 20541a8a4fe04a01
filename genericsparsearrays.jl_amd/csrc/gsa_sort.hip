// gsa_sort.hip — stable LSD radix sort of (key, payload) pairs.
//
// This is the sortperm primitive of the conversions, COO +, and the
// expand-sort-compress SpGEMM (_sortperm_AK, src/helpers.jl:2; Base's stable
// semantics, ext/GenericSparseArraysJLArraysExt.jl:6): equal keys keep their
// input order, so the permutation equals Base.sortperm's.
//
// Keys are 32- or 64-bit unsigned, payloads 32- or 64-bit.  One 8-bit digit per
// pass, only over the bits the keys use.  Per pass:
//   hist    — each 4096-key tile counts its digits (per-wavefront LDS bins,
//             so partly sorted keys do not all hit one counter);
//   scan    — exclusive offsets over the digit-major [256 x tiles] counts;
//   scatter — each wavefront ranks its own 1024 keys stably (16 rounds of 64:
//             ballots give the rank among equal digits inside a round, a
//             per-wavefront counter per digit carries it across rounds, with
//             no barrier); one barrier later the tile knows every digit's
//             offset per wavefront, reorders its keys by digit in LDS and
//             writes each digit run contiguously.
#include "gsa_scan.h"

namespace gsa {

constexpr int RS_WG = 256;
constexpr int RS_WAVES = RS_WG / 64;
constexpr int RS_IPT = 16;                 // keys per thread (rounds per wavefront)
constexpr int RS_TILE = RS_WG * RS_IPT;    // 4096 keys per tile
constexpr int RS_WTILE = 64 * RS_IPT;      // 1024 keys per wavefront
constexpr int RS_BINS = 256;

template <typename K>
__device__ __forceinline__ int digit_of(K k, int shift, int nbits) {
    return (int)((k >> shift) & (K)((1u << nbits) - 1));
}

// Per-tile digit counts, digit-major: hist[d * ntiles + tile].
template <typename K>
__global__ __launch_bounds__(RS_WG) void k_rs_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                   int nbits, int64_t* __restrict__ hist,
                                                   int64_t ntiles) {
    __shared__ int cnt[RS_WAVES][RS_BINS];
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int w = 0; w < RS_WAVES; ++w) cnt[w][tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    K k[RS_IPT];
#pragma unroll
    for (int i = 0; i < RS_IPT; ++i) {
        const int64_t idx = base + (int64_t)i * RS_WG + tid;
        k[i] = idx < n ? keys[idx] : (K)0;
    }
#pragma unroll
    for (int i = 0; i < RS_IPT; ++i) {
        const int64_t idx = base + (int64_t)i * RS_WG + tid;
        if (idx < n) atomicAdd(&cnt[wave][digit_of(k[i], shift, nbits)], 1);
    }
    __syncthreads();
    int c = 0;
    for (int w = 0; w < RS_WAVES; ++w) c += cnt[w][tid];
    hist[(int64_t)tid * ntiles + blockIdx.x] = c;
}

// Orders one wavefront's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive sum over the 256 threads of one workgroup.
__device__ __forceinline__ int block_excl_sum(int v, int* wsum) {
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

template <typename K, typename V>
__global__ __launch_bounds__(RS_WG) void k_rs_scatter(
    const K* __restrict__ keys_in, const V* __restrict__ vals_in, int64_t n, int shift, int nbits,
    const int64_t* __restrict__ incl, const int64_t* __restrict__ hist, int64_t ntiles,
    K* __restrict__ keys_out, V* __restrict__ vals_out) {
    __shared__ int wcnt[RS_WAVES][RS_BINS];   // per-wavefront digit counts, then offsets
    __shared__ int tile_start[RS_BINS];
    __shared__ int wsum[RS_WAVES];
    __shared__ int64_t gbase[RS_BINS];
    __shared__ K s_keys[RS_TILE];
    __shared__ V s_vals[RS_TILE];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t tile = xcd_tile(blockIdx.x, ntiles);
    if (tile >= ntiles) return;
    const int64_t base = tile * RS_TILE;
    const int tile_n = (int)(n - base < RS_TILE ? n - base : RS_TILE);
    {
        const int64_t h = hist[(int64_t)tid * ntiles + tile];
        gbase[tid] = incl[(int64_t)tid * ntiles + tile] - h;  // exclusive global offset
    }
    for (int w = 0; w < RS_WAVES; ++w) wcnt[w][tid] = 0;
    __syncthreads();

    // this wavefront's keys: tile positions wave*1024 + round*64 + lane
    K k[RS_IPT];
    V v[RS_IPT];
#pragma unroll
    for (int i = 0; i < RS_IPT; ++i) {
        const int pos = wave * RS_WTILE + i * 64 + lane;
        if (pos < tile_n) {
            k[i] = keys_in[base + pos];
            v[i] = vals_in[base + pos];
        }
    }
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int* my = wcnt[wave];
    int rank[RS_IPT];
#pragma unroll
    for (int i = 0; i < RS_IPT; ++i) {
        const int pos = wave * RS_WTILE + i * 64 + lane;
        const bool valid = pos < tile_n;
        const int d = valid ? digit_of(k[i], shift, nbits) : 0;
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
        wave_lds_order();
        if (valid && lane == leader) my[d] = before + __popcll(peers);
        wave_lds_order();
    }
    __syncthreads();
    // digit `tid`: offsets of each wavefront's run inside the digit, tile start
    int tot = 0;
    for (int w = 0; w < RS_WAVES; ++w) {
        const int c = wcnt[w][tid];
        wcnt[w][tid] = tot;
        tot += c;
    }
    tile_start[tid] = block_excl_sum(tot, wsum);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_IPT; ++i) {
        const int pos = wave * RS_WTILE + i * 64 + lane;
        if (pos < tile_n) {
            const int d = digit_of(k[i], shift, nbits);
            const int p = tile_start[d] + my[d] + rank[i];
            s_keys[p] = k[i];
            s_vals[p] = v[i];
        }
    }
    __syncthreads();
    for (int p = tid; p < tile_n; p += RS_WG) {
        const K kk = s_keys[p];
        const int d = digit_of(kk, shift, nbits);
        const int64_t dst = gbase[d] + (p - tile_start[d]);
        keys_out[dst] = kk;
        vals_out[dst] = s_vals[p];
    }
}

template <typename K, typename V>
static int radix_sort_impl(hipStream_t s, K* keys, V* vals, int64_t n, int bits) {
    if (n <= 1 || bits <= 0) return GSA_OK;
    GSA_CHECK(bits <= (int)(8 * sizeof(K)), GSA_ERR_INVALID, "radix sort: %d bits in a %d-bit key",
              bits, (int)(8 * sizeof(K)));
    const int64_t ntiles = cdiv(n, RS_TILE);
    Scratch kalt(s), valt(s), hist(s), incl(s);
    GSA_TRY(kalt.alloc(sizeof(K) * n));
    GSA_TRY(valt.alloc(sizeof(V) * n));
    GSA_TRY(hist.alloc(sizeof(int64_t) * RS_BINS * ntiles));
    GSA_TRY(incl.alloc(sizeof(int64_t) * RS_BINS * ntiles));
    K* kin = keys;
    V* vin = vals;
    K* kout = kalt.as<K>();
    V* vout = valt.as<V>();
    int passes = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        const int nb = bits - shift < 8 ? bits - shift : 8;
        hipLaunchKernelGGL(k_rs_hist<K>, dim3((unsigned)ntiles), dim3(RS_WG), 0, s, (const K*)kin, n,
                           shift, nb, hist.as<int64_t>(), ntiles);
        GSA_TRY(inclusive_scan<int64_t>(s, hist.as<int64_t>(), incl.as<int64_t>(),
                                        RS_BINS * ntiles, 0));
        hipLaunchKernelGGL((k_rs_scatter<K, V>), dim3((unsigned)xcd_tile_grid(ntiles)), dim3(RS_WG), 0, s,
                           (const K*)kin, (const V*)vin, n, shift, nb,
                           (const int64_t*)incl.as<int64_t>(), (const int64_t*)hist.as<int64_t>(),
                           ntiles, kout, vout);
        GSA_HIP(hipGetLastError());
        K* tk = kin; kin = kout; kout = tk;
        V* tv = vin; vin = vout; vout = tv;
        ++passes;
    }
    if (passes & 1) {  // result is in the scratch buffers
        GSA_HIP(hipMemcpyAsync(keys, kin, sizeof(K) * n, hipMemcpyDeviceToDevice, s));
        GSA_HIP(hipMemcpyAsync(vals, vin, sizeof(V) * n, hipMemcpyDeviceToDevice, s));
    }
    return GSA_OK;
}

int radix_sort_pairs_u64(hipStream_t s, uint64_t* keys, uint64_t* vals, int64_t n, int bits) {
    return radix_sort_impl<uint64_t, uint64_t>(s, keys, vals, n, bits);
}
int radix_sort_pairs_u64_u32(hipStream_t s, uint64_t* keys, uint32_t* vals, int64_t n, int bits) {
    return radix_sort_impl<uint64_t, uint32_t>(s, keys, vals, n, bits);
}
int radix_sort_pairs_u32_u32(hipStream_t s, uint32_t* keys, uint32_t* vals, int64_t n, int bits) {
    return radix_sort_impl<uint32_t, uint32_t>(s, keys, vals, n, bits);
}

}  // namespace gsa
