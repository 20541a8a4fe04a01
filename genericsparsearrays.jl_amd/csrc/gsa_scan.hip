// gsa_scan.hip — device-wide inclusive prefix sums (the _cumsum_AK primitive,
// src/helpers.jl:3) used to turn per-row counts into one-based pointers and
// masks into write positions.  Reduce-then-scan over 4096-element tiles:
// tile sums, then a per-tile scan with its offset (up to 1024 tiles each tile
// adds up the sums before it; beyond, one block scans the tile sums first).
// Deterministic (integer).
#include "gsa_scan.h"

namespace gsa {

constexpr int SCAN_WG = 256;
constexpr int SCAN_IPT = 16;
constexpr int SCAN_TILE = SCAN_WG * SCAN_IPT;

template <typename T>
__device__ __forceinline__ T block_exclusive_sum(T v, T* lds, T* total) {
    // wave inclusive scan with shuffles, then across the 4 waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    T wpre = 0, all = 0;
    for (int w = 0; w < SCAN_WG / 64; ++w) {
        T s = lds[w];
        if (w < wave) wpre += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return wpre + x - v;
}

template <typename T>
__global__ __launch_bounds__(SCAN_WG) void k_tile_sums(const T* __restrict__ in, int64_t n,
                                                       T* __restrict__ sums) {
    __shared__ T lds[SCAN_WG / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    T s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
        int64_t idx = base + (int64_t)i * SCAN_WG + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    T total;
    (void)block_exclusive_sum(s, lds, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// exclusive scan of `count` tile sums in place, one block, looping
template <typename T>
__global__ __launch_bounds__(SCAN_WG) void k_scan_sums(T* __restrict__ sums, int64_t count) {
    __shared__ T lds[SCAN_WG / 64];
    T carry = 0;
    for (int64_t base = 0; base < count; base += SCAN_WG) {
        int64_t idx = base + threadIdx.x;
        T v = idx < count ? sums[idx] : T(0);
        T total;
        T ex = block_exclusive_sum(v, lds, &total);
        if (idx < count) sums[idx] = carry + ex;
        carry += total;
    }
}

// Sum of sums[0 .. count) over the whole workgroup (every thread gets it).
template <typename T>
__device__ __forceinline__ T block_prefix_of(const T* __restrict__ sums, int64_t count, T* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T part = 0;
    for (int64_t t = threadIdx.x; t < count; t += SCAN_WG) part += sums[t];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if (lane == 0) lds[wave] = part;
    __syncthreads();
    T all = 0;
    for (int w = 0; w < SCAN_WG / 64; ++w) all += lds[w];
    return all;
}

// Inclusive scan of each tile plus its offset; out may alias in.  The offset is
// offs[tile] (the scanned tile sums), or the sum of sums[0 .. tile) when offs is
// null and sums is not (few tiles: no separate scan of the sums), or zero.
// Each wavefront scans its 1024 consecutive elements as 16 coalesced rows of
// 64, each row's lane scan carried into the next.
template <typename T>
__global__ __launch_bounds__(SCAN_WG) void k_scan_tiles(const T* in, int64_t n,
                                                        const T* __restrict__ offs,
                                                        const T* __restrict__ sums, T* out,
                                                        T add) {
    __shared__ T lds[SCAN_WG / 64];
    __shared__ T lds2[SCAN_WG / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)wave * (64 * SCAN_IPT);
    T v[SCAN_IPT];
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
        const int64_t idx = base + i * 64 + lane;
        v[i] = idx < n ? in[idx] : T(0);
    }
    T carry = 0;
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
        T x = v[i];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const T y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        const T row = __shfl(x, 63, 64);
        v[i] = x + carry;
        carry += row;
    }
    if (lane == 0) lds[wave] = carry;
    T tile_off = add;
    if (offs) tile_off += offs[blockIdx.x];
    else if (sums) tile_off += block_prefix_of(sums, (int64_t)blockIdx.x, lds2);
    __syncthreads();
    T wpre = 0;
    for (int w = 0; w < wave; ++w) wpre += lds[w];
    const T o = tile_off + wpre;
#pragma unroll
    for (int i = 0; i < SCAN_IPT; ++i) {
        const int64_t idx = base + i * 64 + lane;
        if (idx < n) out[idx] = v[i] + o;
    }
}

// Up to this many tiles, each tile sums the tile sums before it itself.
constexpr int64_t SCAN_FUSE_TILES = 1024;

template <typename T>
int inclusive_scan(hipStream_t s, const T* in, T* out, int64_t n, T add) {
    if (n <= 0) return GSA_OK;
    const int64_t tiles = cdiv(n, SCAN_TILE);
    if (tiles == 1) {
        hipLaunchKernelGGL(k_scan_tiles<T>, dim3(1), dim3(SCAN_WG), 0, s, in, n, (const T*)nullptr,
                           (const T*)nullptr, out, add);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    }
    Scratch sums(s);
    GSA_TRY(sums.alloc(sizeof(T) * tiles));
    hipLaunchKernelGGL(k_tile_sums<T>, dim3((unsigned)tiles), dim3(SCAN_WG), 0, s, in, n,
                       sums.as<T>());
    if (tiles <= SCAN_FUSE_TILES) {
        hipLaunchKernelGGL(k_scan_tiles<T>, dim3((unsigned)tiles), dim3(SCAN_WG), 0, s, in, n,
                           (const T*)nullptr, (const T*)sums.as<T>(), out, add);
    } else {
        hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(SCAN_WG), 0, s, sums.as<T>(), tiles);
        hipLaunchKernelGGL(k_scan_tiles<T>, dim3((unsigned)tiles), dim3(SCAN_WG), 0, s, in, n,
                           (const T*)sums.as<T>(), (const T*)nullptr, out, add);
    }
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

template int inclusive_scan<int64_t>(hipStream_t, const int64_t*, int64_t*, int64_t, int64_t);
template int inclusive_scan<int32_t>(hipStream_t, const int32_t*, int32_t*, int64_t, int32_t);

// Zero `bytes` bytes with a kernel: the unaligned head bytes, the 8-byte
// words, then the tail bytes.  An ordinary kernel node under graph capture,
// unlike a memset node.
__global__ void k_zero_fill(uint8_t* __restrict__ p, int64_t head, int64_t words, int64_t tail) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    uint64_t* w = reinterpret_cast<uint64_t*>(p + head);
    for (int64_t i = t; i < words; i += stride) w[i] = 0;
    if (t < head) p[t] = 0;
    if (t < tail) p[head + words * 8 + t] = 0;
}

int zero_fill(hipStream_t s, void* p, size_t bytes) {
    if (bytes == 0) return GSA_OK;
    const int64_t n = (int64_t)bytes;
    int64_t head = (int64_t)((8 - ((uintptr_t)p & 7)) & 7);
    if (head > n) head = n;
    const int64_t words = (n - head) / 8;
    const int64_t tail = (n - head) - words * 8;
    hipLaunchKernelGGL(k_zero_fill, dim3(grid_for(words > 0 ? words : 1, 256)), dim3(256), 0, s,
                       reinterpret_cast<uint8_t*>(p), head, words, tail);
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

template <typename I>
__global__ void k_set_one(I* p) { *p = 1; }

// The 64-bit sum of m non-negative 32-bit counts (one atomic per workgroup).
__global__ __launch_bounds__(256) void k_sum_counts32(const int32_t* __restrict__ c, int64_t m,
                                                      unsigned long long* __restrict__ out) {
    __shared__ unsigned long long part[4];
    unsigned long long t = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x)
        t += (unsigned long long)(uint32_t)c[i];
    for (int off = 32; off >= 1; off >>= 1) t += __shfl_xor(t, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, part[0] + part[1] + part[2] + part[3]);
}

template <typename I>
int counts_to_ptr(hipStream_t s, I* ptr, int64_t m, int64_t* total) {
    // ptr[1..m] hold counts; make ptr one-based pointers (ptr[0] = 1).
    // With 32-bit indices and a total asked for, the counts are first summed in
    // 64 bits: a result of 2^31 - 1 entries or more has no 32-bit pointer, and
    // the call fails instead of handing back a wrapped one.
    Scratch wide(s);
    if constexpr (sizeof(I) == 4) {
        if (total && m > 0) {
            GSA_TRY(wide.alloc(sizeof(unsigned long long)));
            GSA_TRY(zero_fill(s, wide.p, sizeof(unsigned long long)));
            // at most 256 workgroups: each adds its part with one device-wide
            // atomic on the same address, and those serialise
            hipLaunchKernelGGL(k_sum_counts32, dim3((unsigned)(cdiv(m, 256) < 256 ? cdiv(m, 256) : 256)),
                               dim3(256), 0, s,
                               (const int32_t*)(ptr + 1), m, wide.as<unsigned long long>());
            GSA_HIP(hipGetLastError());
        }
    }
    hipLaunchKernelGGL(k_set_one<I>, dim3(1), dim3(1), 0, s, ptr);
    if (m > 0) GSA_TRY(inclusive_scan<I>(s, ptr + 1, ptr + 1, m, (I)1));
    if (total) {
        I last = 1;
        unsigned long long sum64 = 0;
        GSA_HIP(hipMemcpyAsync(&last, ptr + m, sizeof(I), hipMemcpyDeviceToHost, s));
        if (wide.p)
            GSA_HIP(hipMemcpyAsync(&sum64, wide.p, sizeof sum64, hipMemcpyDeviceToHost, s));
        GSA_HIP(hipStreamSynchronize(s));
        if (wide.p && sum64 >= (unsigned long long)INT32_MAX) {
            *total = (int64_t)sum64;
            GSA_CHECK(false, GSA_ERR_INVALID,
                      "the result has %llu stored entries: more than 32-bit indices can point to",
                      sum64);
        }
        *total = (int64_t)last - 1;
    }
    return GSA_OK;
}

template int counts_to_ptr<int64_t>(hipStream_t, int64_t*, int64_t, int64_t*);
template int counts_to_ptr<int32_t>(hipStream_t, int32_t*, int64_t, int64_t*);

}  // namespace gsa
