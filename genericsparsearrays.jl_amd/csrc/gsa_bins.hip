// gsa_bins.hip — scatter-shaped SpMV (CSC-N, COO-N in any entry order) by row bins.
//
// The reference adds every stored entry into y with an atomic
// (kernel_spmatmul_csc_N!, src/matrix_csc/matrix_csc_kernels.jl:1-22;
// kernel_spmatmul_coo_N!, src/matrix_coo/matrix_coo_kernels.jl:1-20), so
// its summation order is undefined.  Here y is cut into bins of BIN_ROWS
// rows, and one mul! is two streaming passes, with no random reads of the
// matrix and no global atomics:
//   1. k_bin_products: each 4096-entry tile of the storage forms its products
//      a_ij * x_j (for CSC the entries of a column are consecutive, so x is
//      read nearly in order), stages them in LDS in bin order and writes each
//      bin's run to its place in the bin-major product array.  Where each
//      entry goes depends on the rows only, so k_bin_rank works it out once,
//      at plan time (a stable ranking by bin per tile), with the row within
//      the bin of every binned position;
//   2. k_bin_sum: one workgroup per bin.  Each wavefront adds a contiguous
//      share of the bin's pairs into its own Float64 accumulator in LDS, the
//      eight accumulators are added in a fixed order, and each row of y is
//      written once (beta, alpha applied there).
// Float32 products are accumulated in Float64.  The plan holds the per-tile
// offsets, slots and destinations (from the row indices only), the binned
// rows, and for CSC each entry's column; values and x are read at every call,
// and the binned products live in scratch of the call (the plan is read-only,
// so concurrent calls may share it).
#include "gsa_common.h"
#include "gsa_rows.h"
#include "gsa_scan.h"

namespace gsa {

constexpr int BIN_SHIFT = 11;
constexpr int BIN_ROWS = 1 << BIN_SHIFT;   // rows per bin
constexpr int BIN_WG = 256;
constexpr int BIN_IPT = 16;                // entries per thread in pass 1
constexpr int BIN_TILE = BIN_WG * BIN_IPT; // 4096 entries per tile
constexpr int BIN_WTILE = 64 * BIN_IPT;    // 1024 entries per wavefront
constexpr int BIN_SUM_WAVES = 8;
constexpr int BIN_SUM_U = 8;               // pairs per lane in flight in pass 2
constexpr int64_t BIN_MAX = 2048;          // bins (m <= 4,194,304 rows)

struct BinPlan {
    int itype = 0;
    int64_t m = 0, n = 0, nnz = 0;
    int64_t nbins = 0, ntiles = 0;
    int bins_cap = 0;                  // 512 or 2048: the pass-1 LDS layout
    int64_t* off = nullptr;            // nbins*ntiles + 1: start of (bin, tile), bin-major
    void* col = nullptr;               // CSC: each entry's one-based column (u32 or I)
    bool col32 = false;
    uint16_t* lrow = nullptr;          // nnz: row within the bin, by bin
    uint16_t* slot = nullptr;          // nnz: staging slot of each entry in its tile
    void* dstq = nullptr;              // nnz: binned position of each staging slot (u32 or i64)
    bool dst32 = false;
    const void* src[2] = {nullptr, nullptr};
};

template <typename I>
__global__ __launch_bounds__(BIN_WG) void k_bin_hist(int64_t m, int64_t nnz,
                                                     const I* __restrict__ rowidx,
                                                     int64_t nbins, int64_t ntiles,
                                                     int64_t* __restrict__ hist) {
    __shared__ int cnt[BIN_MAX];
    const int tid = threadIdx.x;
    for (int b = tid; b < nbins; b += BIN_WG) cnt[b] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * BIN_TILE;
    for (int i = tid; i < BIN_TILE; i += BIN_WG) {
        const int64_t j = base + i;
        if (j < nnz) {
            const int64_t r = dclamp((int64_t)rowidx[j] - 1, 0, m - 1);
            atomicAdd(&cnt[r >> BIN_SHIFT], 1);
        }
    }
    __syncthreads();
    for (int b = tid; b < nbins; b += BIN_WG) hist[(int64_t)b * ntiles + blockIdx.x] = cnt[b];
}

__global__ void k_bin_excl(int64_t n, const int64_t* __restrict__ hist,
                           const int64_t* __restrict__ incl, int64_t* __restrict__ off) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n;
         i += (int64_t)gridDim.x * blockDim.x)
        off[i] = i < n ? incl[i] - hist[i] : (n > 0 ? incl[n - 1] : 0);
}

template <typename I>
__global__ void k_to_u32(int64_t n, const I* __restrict__ in, uint32_t* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)in[i];
}

// Orders one wavefront's LDS writes before later LDS reads by its other lanes.
__device__ __forceinline__ void bin_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Plan time: the place of every stored entry in the binned arrays.  Each
// 4096-entry tile ranks its entries stably by bin (ballots, per-wavefront
// counters), as the products will be staged in LDS in bin order: slot[j] is
// entry j's position in that staging, dstq[tile*4096 + p] the binned position
// of staged slot p (runs of consecutive positions per bin), and blrow the row
// within the bin of every binned position.
template <typename I, typename D, int BINS>
__global__ __launch_bounds__(BIN_WG) void k_bin_rank(
    int64_t m, int64_t nnz, const I* __restrict__ rowidx, const int64_t* __restrict__ off,
    int64_t nbins, int64_t ntiles, uint16_t* __restrict__ slot, D* __restrict__ dstq,
    uint16_t* __restrict__ blrow) {
    constexpr int BITS = BINS == 512 ? 9 : 11;
    constexpr int PER = BINS / BIN_WG;  // bins per thread in the tile scan
    __shared__ int wcnt[4][BINS];       // per-wavefront bin counts, then offsets
    __shared__ int tstart[BINS];
    __shared__ int64_t gbase[BINS];
    __shared__ int wsum[4];
    __shared__ uint16_t s_lrow[BIN_TILE];
    __shared__ uint16_t s_bin[BIN_TILE];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t tile = xcd_tile(blockIdx.x, ntiles);
    if (tile >= ntiles) return;
    const int64_t base = tile * BIN_TILE;
    const int tile_n = (int)(nnz - base < BIN_TILE ? nnz - base : BIN_TILE);
    for (int d = tid; d < BINS; d += BIN_WG) {
        for (int w = 0; w < 4; ++w) wcnt[w][d] = 0;
        gbase[d] = d < nbins ? off[(int64_t)d * ntiles + tile] : 0;
    }
    // this wavefront's entries: tile positions wave*1024 + i*64 + lane
    I rr[BIN_IPT];
#pragma unroll
    for (int i = 0; i < BIN_IPT; ++i) {
        const int pos = wave * BIN_WTILE + i * 64 + lane;
        rr[i] = rowidx[base + (pos < tile_n ? pos : tile_n - 1)];
    }
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int* my = wcnt[wave];
    int rank[BIN_IPT], bin[BIN_IPT];
#pragma unroll
    for (int i = 0; i < BIN_IPT; ++i) {
        const int pos = wave * BIN_WTILE + i * 64 + lane;
        const bool valid = pos < tile_n;
        const int64_t r = dclamp((int64_t)rr[i] - 1, 0, m - 1);
        const int d = valid ? (int)(r >> BIN_SHIFT) : 0;
        bin[i] = d;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < BITS; ++b) {
            const uint64_t bal = __ballot(valid && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? bal : ~bal;
        }
        // every lane reads its bin's count before the round's leader adds to it
        const int before = valid ? my[d] : 0;
        rank[i] = before + __popcll(peers & lt_mask);
        const int leader = valid ? (__ffsll((long long)peers) - 1) : -1;
        bin_wave_sync();
        if (valid && lane == leader) my[d] = before + __popcll(peers);
        bin_wave_sync();
    }
    __syncthreads();
    // bins tid*PER .. tid*PER+PER-1: offsets of each wavefront's run, tile start
    int loc[PER];
    int tot = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int d = tid * PER + q;
        int t = 0;
        for (int w = 0; w < 4; ++w) {
            const int c = wcnt[w][d];
            wcnt[w][d] = t;
            t += c;
        }
        loc[q] = t;
        tot += t;
    }
    int all;
    int run = wg_exclusive_sum<int>(tot, wsum, &all);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        tstart[tid * PER + q] = run;
        run += loc[q];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BIN_IPT; ++i) {
        const int pos = wave * BIN_WTILE + i * 64 + lane;
        if (pos < tile_n) {
            const int d = bin[i];
            const int p = tstart[d] + my[d] + rank[i];
            const int64_t r = dclamp((int64_t)rr[i] - 1, 0, m - 1);
            slot[base + pos] = (uint16_t)p;
            s_lrow[p] = (uint16_t)(r & (BIN_ROWS - 1));
            s_bin[p] = (uint16_t)d;
        }
    }
    __syncthreads();
    for (int p = tid; p < tile_n; p += BIN_WG) {
        const int d = s_bin[p];
        const int64_t dst = dclamp(gbase[d] + (p - tstart[d]), 0, nnz - 1);
        dstq[base + p] = (D)dst;
        blrow[dst] = s_lrow[p];
    }
}

// Pass 1 of every call: the products of a tile, staged in LDS in bin order
// through the plan's slots, then written to their binned positions (runs of
// consecutive addresses).
template <typename T, typename CI, typename D>
__global__ __launch_bounds__(BIN_WG) void k_bin_products(
    int64_t n, int64_t nnz, const CI* __restrict__ colidx, const T* __restrict__ val,
    const T* __restrict__ x, const uint16_t* __restrict__ slot, const D* __restrict__ dstq,
    int64_t ntiles, T* __restrict__ bprod) {
    __shared__ T s_prod[BIN_TILE];
    const int tid = threadIdx.x;
    const int64_t tile = xcd_tile(blockIdx.x, ntiles);
    if (tile >= ntiles) return;
    const int64_t base = tile * BIN_TILE;
    const int tile_n = (int)(nnz - base < BIN_TILE ? nnz - base : BIN_TILE);
    CI cc[BIN_IPT];
    T vv[BIN_IPT];
    uint16_t q[BIN_IPT];
#pragma unroll
    for (int i = 0; i < BIN_IPT; ++i) {
        const int p = i * BIN_WG + tid;
        const int64_t j = base + (p < tile_n ? p : tile_n - 1);
        cc[i] = __builtin_nontemporal_load(colidx + j);
        vv[i] = __builtin_nontemporal_load(val + j);
        q[i] = __builtin_nontemporal_load(slot + j);
    }
#pragma unroll
    for (int i = 0; i < BIN_IPT; ++i)
        if (i * BIN_WG + tid < tile_n)
            s_prod[q[i] < tile_n ? q[i] : 0] = vv[i] * x[dclamp((int64_t)cc[i] - 1, 0, n - 1)];
    __syncthreads();
    for (int p = tid; p < tile_n; p += BIN_WG)
        bprod[dclamp((int64_t)__builtin_nontemporal_load(dstq + base + p), 0, nnz - 1)] = s_prod[p];
}

template <typename T>
__global__ __launch_bounds__(64 * BIN_SUM_WAVES) void k_bin_sum(
    int64_t m, int64_t nnz, int64_t nbins, int64_t ntiles, const int64_t* __restrict__ off,
    const uint16_t* __restrict__ blrow, const T* __restrict__ bprod, T* __restrict__ y, T alpha,
    T beta, int mode) {
    __shared__ double acc[BIN_SUM_WAVES][BIN_ROWS];
    const int64_t b = blockIdx.x;
    if (b >= nbins) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < BIN_SUM_WAVES * BIN_ROWS; i += blockDim.x)
        acc[i / BIN_ROWS][i % BIN_ROWS] = 0.0;
    __syncthreads();
    const int64_t s = off[b * ntiles];
    const int64_t e = b + 1 < nbins ? off[(b + 1) * ntiles] : nnz;
    const int64_t per = ((e - s + BIN_SUM_WAVES - 1) / BIN_SUM_WAVES + 63) / 64 * 64;
    const int64_t ws = s + wv * per < e ? s + wv * per : e;
    const int64_t we = ws + per < e ? ws + per : e;
    double* my = acc[wv];
    for (int64_t j0 = ws; j0 < we; j0 += 64 * BIN_SUM_U) {
        uint16_t lr[BIN_SUM_U];
        T pv[BIN_SUM_U];
#pragma unroll
        for (int u = 0; u < BIN_SUM_U; ++u) {
            const int64_t j = j0 + u * 64 + lane;
            if (j < we) {
                lr[u] = __builtin_nontemporal_load(blrow + j);
                pv[u] = __builtin_nontemporal_load(bprod + j);
            }
        }
#pragma unroll
        for (int u = 0; u < BIN_SUM_U; ++u)
            if (j0 + u * 64 + lane < we) atomicAdd(&my[lr[u]], (double)pv[u]);
    }
    __syncthreads();
    const int64_t r0 = b << BIN_SHIFT;
    for (int r = threadIdx.x; r < BIN_ROWS && r0 + r < m; r += blockDim.x) {
        double sum = acc[0][r];
#pragma unroll
        for (int w = 1; w < BIN_SUM_WAVES; ++w) sum = sum + acc[w][r];
        store_out(y, r0 + r, (T)sum, alpha, beta, mode);
    }
}

void release_bin_plan(BinPlan* p) {
    if (!p) return;
    for (void* q : {(void*)p->off, p->col, (void*)p->lrow, (void*)p->slot, p->dstq})
        if (q) (void)hipFree(q);
    delete p;
}

int build_bin_plan(hipStream_t s, int itype, int64_t m, int64_t n, int64_t nnz,
                   const void* rowidx, const void* colptr, BinPlan** out) {
    *out = nullptr;
    const int64_t nbins = cdiv(m, BIN_ROWS);
    if (m <= 0 || n <= 0 || nnz <= 0 || nbins > BIN_MAX) return GSA_OK;
    auto* p = new BinPlan();
    p->itype = itype;
    p->m = m;
    p->n = n;
    p->nnz = nnz;
    p->nbins = nbins;
    p->ntiles = cdiv(nnz, BIN_TILE);
    p->bins_cap = nbins <= 512 ? 512 : 2048;
    const int r = dispatch_i(itype, [&](auto ti) -> int {
        using I = decltype(ti);
        const int64_t cells = nbins * p->ntiles;
        GSA_HIP(hipMalloc(&p->off, sizeof(int64_t) * (cells + 1)));
        GSA_HIP(hipMalloc(&p->lrow, sizeof(uint16_t) * nnz));
        Scratch hist(s), incl(s);
        GSA_TRY(hist.alloc(sizeof(int64_t) * cells));
        GSA_TRY(incl.alloc(sizeof(int64_t) * cells));
        hipLaunchKernelGGL(k_bin_hist<I>, dim3((unsigned)p->ntiles), dim3(BIN_WG), 0, s, m, nnz,
                           (const I*)rowidx, nbins, p->ntiles, hist.as<int64_t>());
        GSA_TRY(inclusive_scan<int64_t>(s, hist.as<int64_t>(), incl.as<int64_t>(), cells, 0));
        hipLaunchKernelGGL(k_bin_excl, dim3(grid_for(cells + 1, 256)), dim3(256), 0, s, cells,
                           (const int64_t*)hist.as<int64_t>(), (const int64_t*)incl.as<int64_t>(),
                           p->off);
        GSA_HIP(hipGetLastError());
        GSA_HIP(hipMalloc(&p->slot, sizeof(uint16_t) * nnz));
        p->dst32 = nnz <= (int64_t)UINT32_MAX;
        GSA_HIP(hipMalloc(&p->dstq, (p->dst32 ? sizeof(uint32_t) : sizeof(int64_t)) * nnz));
        auto rank = [&](auto d_t) {
            using D = decltype(d_t);
            const unsigned g = (unsigned)xcd_tile_grid(p->ntiles);
            if (p->bins_cap == 512)
                hipLaunchKernelGGL((k_bin_rank<I, D, 512>), dim3(g), dim3(BIN_WG), 0, s, m, nnz,
                                   (const I*)rowidx, (const int64_t*)p->off, nbins, p->ntiles,
                                   p->slot, (D*)p->dstq, p->lrow);
            else
                hipLaunchKernelGGL((k_bin_rank<I, D, 2048>), dim3(g), dim3(BIN_WG), 0, s, m, nnz,
                                   (const I*)rowidx, (const int64_t*)p->off, nbins, p->ntiles,
                                   p->slot, (D*)p->dstq, p->lrow);
        };
        if (p->dst32) rank(uint32_t{});
        else rank(int64_t{});
        GSA_HIP(hipGetLastError());
        if (colptr) {  // CSC: the column of every stored entry
            Scratch cols(s);
            GSA_TRY(cols.alloc(sizeof(I) * nnz));
            GSA_TRY(expand_rows(s, itype, n, nnz, colptr, cols.p));
            p->col32 = n <= (int64_t)UINT32_MAX;
            if (p->col32) {
                GSA_HIP(hipMalloc(&p->col, sizeof(uint32_t) * nnz));
                hipLaunchKernelGGL(k_to_u32<I>, dim3(grid_for(nnz, 256)), dim3(256), 0, s, nnz,
                                   (const I*)cols.p, (uint32_t*)p->col);
            } else {
                GSA_HIP(hipMalloc(&p->col, sizeof(I) * nnz));
                GSA_HIP(hipMemcpyAsync(p->col, cols.p, sizeof(I) * nnz, hipMemcpyDeviceToDevice, s));
            }
            GSA_HIP(hipGetLastError());
        }
        GSA_HIP(hipStreamSynchronize(s));  // the scratch is released on return
        return GSA_OK;
    });
    if (r != GSA_OK) {
        release_bin_plan(p);
        return r;
    }
    *out = p;
    return GSA_OK;
}

bool bin_plan_matches(const BinPlan* p, int itype, int64_t m, int64_t n, int64_t nnz,
                      const void* a, const void* b) {
    return p && p->itype == itype && p->m == m && p->n == n && p->nnz == nnz && p->src[0] == a &&
           p->src[1] == b;
}

void bin_plan_set_sources(BinPlan* p, const void* a, const void* b) {
    p->src[0] = a;
    p->src[1] = b;
}

// y = beta y + alpha A x through a bin plan (built on the rows): colidx (COO)
// holds each entry's one-based column, or is null to use the plan's columns
// (CSC).  x has stride ldx.
int bin_spmv(hipStream_t s, const BinPlan& p, int vtype, const void* colidx,
             const void* nzval, const void* x, int64_t ldx, void* y, double alpha, double beta) {
    const int mode = beta == 0.0 ? 0 : (beta == 1.0 ? 1 : 2);
    return dispatch_vi(vtype, p.itype, [&](auto tv, auto ti) -> int {
        using T = decltype(tv);
        using I = decltype(ti);
        const T* xv = (const T*)x;
        Scratch xs(s);
        if (ldx != 1) {  // strided vector: compact it first
            GSA_TRY(xs.alloc(sizeof(T) * p.n));
            GSA_HIP(hipMemcpy2DAsync(xs.p, sizeof(T), x, sizeof(T) * ldx, sizeof(T), p.n,
                                     hipMemcpyDeviceToDevice, s));
            xv = xs.as<T>();
        }
        Scratch prod(s);  // the binned products of this call
        GSA_TRY(prod.alloc(sizeof(T) * p.nnz));
        const unsigned g1 = (unsigned)xcd_tile_grid(p.ntiles);
        auto products = [&](auto ci_t, auto d_t) {
            using CI = decltype(ci_t);
            using D = decltype(d_t);
            const CI* ci = colidx ? (const CI*)colidx : (const CI*)p.col;
            hipLaunchKernelGGL((k_bin_products<T, CI, D>), dim3(g1), dim3(BIN_WG), 0, s, p.n,
                               p.nnz, ci, (const T*)nzval, xv, (const uint16_t*)p.slot,
                               (const D*)p.dstq, p.ntiles, prod.as<T>());
        };
        auto by_col = [&](auto d_t) {
            if (colidx || !p.col32) products(I{}, d_t);
            else products(uint32_t{}, d_t);
        };
        if (p.dst32) by_col(uint32_t{});
        else by_col(int64_t{});
        hipLaunchKernelGGL(k_bin_sum<T>, dim3((unsigned)p.nbins), dim3(64 * BIN_SUM_WAVES), 0, s,
                           p.m, p.nnz, p.nbins, p.ntiles, (const int64_t*)p.off,
                           (const uint16_t*)p.lrow, (const T*)prod.as<T>(), (T*)y, (T)alpha, (T)beta,
                           mode);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    });
}

}  // namespace gsa
