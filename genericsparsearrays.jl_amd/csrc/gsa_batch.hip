// gsa_batch.hip — many uneven CSR SpMM instances in one persistent launch,
// the BF16-with-row-scale value format, and HIP-graph capture of a batch.
//
// Work unit: a tile = (instance, [r0, r1)) from a device table built once per
// batch, in 8 segments, one per XCD (an instance's tiles stay on one XCD, so
// its dense operand stays in that XCD's L2).  A persistent grid of workgroups
// takes tiles through one atomic counter per segment (reset by a small kernel
// launched just before, so the pair is graph-capturable: no memset).  The
// counters are scratch of the call, so a plan is read-only and may serve
// concurrent launches on several streams.  Inside a tile each row gets LPR
// lanes, each summing W consecutive output columns left to right (W elements
// of a row-major B per vector load, up to 16 bytes; W = 1 otherwise); rows over
// BATCH_LONG entries are reduced by the whole workgroup.
//
// Precision: Float32 and BF16 storage accumulate in Float32; rows over 256
// entries accumulate with TwoSum (an error-free sum and a running
// compensation, combined pairwise in the reduction).  Float64 accumulates in
// Float64.  BF16 values carry one Float32 scale per row (a power of two, so
// scaling is exact): a_ij = scale_i * bf16_ij.
#include "gsa_common.h"
#include <algorithm>
#include <vector>

namespace gsa {

constexpr int BWG = 256;
constexpr int BATCH_LONG = 256;   // rows above this: workgroup-cooperative, compensated
constexpr int BATCH_MAX_TILE_ROWS = 1024;
constexpr int BATCH_TCAP = 3072;   // entries of a tile staged in LDS at a time
constexpr int BATCH_CHUNK_LONG = BATCH_TCAP / (BATCH_LONG + 1) + 1;

// The tile table in segments: workgroup b starts on segment b % n, whose
// tiles are those of the instances given to XCD b % 8 (workgroups go to the
// XCDs round-robin), so each XCD's L2 holds the dense operands of its own
// instances only.  A workgroup that finds its segment used up takes tiles
// from the next ones.
constexpr int BATCH_SEGS = 8;
struct BatchSegs {
    int64_t start[BATCH_SEGS + 1];
    int n;
};

struct BatchPlan {
    int vtype = 0, itype = 0;
    int64_t count = 0;
    gsa_batch_desc* d_desc = nullptr;
    int32_t* d_tiles = nullptr;   // 3 per tile: instance, r0, r1
    int64_t ntiles = 0;
    BatchSegs segs{};
    int grid = 0;
};

__device__ __forceinline__ float bf16_to_f32(uint16_t v) {
    return __uint_as_float(((uint32_t)v) << 16);
}

// round-to-nearest-even Float32 -> BF16 (NaN kept quiet)
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Value access by storage type: V = double / float / uint16_t (BF16).
template <typename V> struct Acc;
template <> struct Acc<double> { using T = double; };
template <> struct Acc<float> { using T = float; };
template <> struct Acc<uint16_t> { using T = float; };

template <typename V>
__device__ __forceinline__ typename Acc<V>::T vload(const V* p) {
    if constexpr (sizeof(V) == 2) return bf16_to_f32(*p);
    else return *p;
}

// Knuth's TwoSum: s + e == a + b exactly (no FMA contraction in this file).
template <typename T>
__device__ __forceinline__ void two_sum(T a, T b, T& s, T& e) {
    s = a + b;
    T bp = s - a;
    T ap = s - bp;
    e = (a - ap) + (b - bp);
}

__global__ void k_reset_counter(int* c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

// Stored entries of every instance (rowptr[m] - 1), for the plan's balance.
template <typename I>
__global__ void k_inst_nnz(const gsa_batch_desc* __restrict__ desc, int64_t count,
                           int64_t* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    const gsa_batch_desc d = desc[i];
    out[i] = d.m > 0 ? (int64_t)((const I*)d.rowptr)[d.m] - 1 : 0;
}

// Shared-memory workspace of one workgroup processing tiles.
template <typename T>
struct TileLds {
    int nlong;
    int long_rows[BATCH_CHUNK_LONG];    // a chunk's rows over BATCH_LONG entries
    int off[BATCH_MAX_TILE_ROWS + 1];   // staged row offsets (relative to P0)
    int col[BATCH_TCAP];                // one chunk's column indices (0-based, clamped)
    T val[BATCH_TCAP];                  // and values (BF16 widened)
    T red[BWG];
    T cmp[BWG];
};

template <typename T, int W> struct VecT { typedef T type __attribute__((ext_vector_type(W))); };

// Elements of B one lane loads per stored entry: a 16-byte (or 8-byte) vector
// of a row-major B whose k, leading dimension and base allow it, else 1.  A
// row-major B row of k elements then takes k / W lanes, so one load
// instruction covers W times as many stored entries.
__host__ __device__ inline int batch_vw(const gsa_batch_desc& d, int esize) {
    if (!(d.flags & GSA_BATCH_B_ROWMAJOR) || d.k < 2) return 1;
    for (int w = 16 / esize; w >= 2; w >>= 1)
        if (d.k % w == 0 && d.ldb % w == 0 && ((uintptr_t)d.B) % (uintptr_t)(w * esize) == 0)
            return w;
    return 1;
}

// Lanes per row: next_pow2(k / W), at most 64.
__host__ __device__ inline int batch_lpr(const gsa_batch_desc& d, int esize) {
    const int64_t slots = (d.k + batch_vw(d, esize) - 1) / batch_vw(d, esize);
    int lpr = 1;
    while (lpr < slots && lpr < 64) lpr <<= 1;
    return lpr;
}

// One tile: rows [r0, r1) of instance d.  All BWG lanes call it together.
// The tile's row offsets are staged in LDS; its entries go through LDS in
// chunks of whole rows of at most BATCH_TCAP entries (coalesced loads).  In a
// chunk each row of at most BATCH_LONG entries gets LPR = next_pow2(k) lanes
// (one (row, kk) output per lane, summed left to right); longer rows are
// reduced by the whole workgroup with TwoSum, from LDS, or from HBM for a row
// longer than a chunk.
template <typename V, typename I, int VW>
__device__ __forceinline__ void process_tile(const gsa_batch_desc& d, int64_t r0, int64_t r1,
                                             typename Acc<V>::T alpha, typename Acc<V>::T beta,
                                             int mode, TileLds<typename Acc<V>::T>& sh) {
    using T = typename Acc<V>::T;
    const int tid = threadIdx.x;
    const I* __restrict__ rowptr = (const I*)d.rowptr;
    const I* __restrict__ colval = (const I*)d.colval;
    const V* __restrict__ nzval = (const V*)d.nzval;
    const float* __restrict__ scale = (const float*)d.row_scale;
    const T* __restrict__ B = (const T*)d.B;
    T* __restrict__ C = (T*)d.C;
    const int k = d.k;
    const int64_t n = d.n > 0 ? d.n : 1;  // clamp bound for column indices
    // layout flags: B[i,kk] at i + kk*ldb (column-major) or i*ldb + kk
    // (row-major); the same for C with ldc
    const bool brow = (d.flags & GSA_BATCH_B_ROWMAJOR) != 0;
    const bool crow = (d.flags & GSA_BATCH_C_ROWMAJOR) != 0;
    const int64_t bsr = brow ? d.ldb : 1, bsc = brow ? 1 : d.ldb;
    const int64_t csr = crow ? d.ldc : 1, csc = crow ? 1 : d.ldc;
    const int nrows = (int)(r1 - r0);
    const int64_t P0 = (int64_t)rowptr[r0] - 1;
    for (int i = tid; i <= nrows; i += BWG) sh.off[i] = (int)((int64_t)rowptr[r0 + i] - 1 - P0);
    __syncthreads();
    const int lpr = batch_lpr(d, (int)sizeof(T));  // VW == batch_vw(d, sizeof(T))
    const int groups = BWG / lpr;
    const int g = tid / lpr, lane = tid % lpr;
    auto emit = [&](int64_t r, int kk, T acc) {
        if (scale) acc = acc * (T)scale[r];
        T* out = C + r * csr + kk * csc;
        T tt = acc * alpha;
        T base = mode == 0 ? T(0) : (mode == 1 ? *out : *out * beta);
        *out = base + tt;
    };
    // one row by the whole workgroup: the groups stride over its entries, each
    // lane owns one kk (all k columns in one pass), TwoSum per lane, then an
    // error-free combine of the groups' partial sums through LDS.  `at(j)` gives
    // entry j's (column, value), j in [s, e).
    auto wg_row = [&](int64_t r, int64_t s, int64_t e, auto at) {
        for (int k0 = 0; k0 < k; k0 += lpr) {
            const int kk = k0 + lane;
            const bool act = kk < k;
            const int kc = act ? kk : 0;
            T sum = T(0), cmp = T(0);
            for (int64_t j = s + g; j < e; j += 4 * (int64_t)groups) {
                T pv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t j2 = j + u * (int64_t)groups;
                    const int64_t jj = j2 < e ? j2 : e - 1;
                    int64_t c;
                    T v;
                    at(jj, c, v);
                    pv[u] = (j2 < e) ? v * B[c * bsr + kc * bsc] : T(0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    T ns, er;
                    two_sum(sum, pv[u], ns, er);
                    sum = ns;
                    cmp = cmp + er;
                }
            }
            sh.red[tid] = sum;
            sh.cmp[tid] = cmp;
            __syncthreads();
            for (int h = groups / 2; h > 0; h >>= 1) {
                if (g < h) {
                    const int o = tid + h * lpr;
                    T ns, er;
                    two_sum(sh.red[tid], sh.red[o], ns, er);
                    sh.red[tid] = ns;
                    sh.cmp[tid] = sh.cmp[tid] + sh.cmp[o] + er;
                }
                __syncthreads();
            }
            if (g == 0 && act) emit(r, kk, sh.red[tid] + sh.cmp[tid]);
            __syncthreads();
        }
    };
    for (int ca = 0; ca < nrows;) {
        // the chunk: rows [ca, cb) with at most BATCH_TCAP entries (at least one row)
        const int ea = sh.off[ca];
        int cb;
        {
            int lo = ca, hi = nrows;  // the last cb with off[cb] - ea <= BATCH_TCAP
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sh.off[mid] - ea <= BATCH_TCAP) lo = mid;
                else hi = mid - 1;
            }
            cb = lo;
        }
        if (cb == ca) {  // a row longer than a chunk, from HBM
            wg_row(r0 + ca, P0 + sh.off[ca], P0 + sh.off[ca + 1],
                   [&](int64_t j, int64_t& c, T& v) {
                       c = (int64_t)colval[j] - 1;
                       c = c < 0 ? 0 : (c >= n ? n - 1 : c);
                       v = vload(nzval + j);
                   });
            ++ca;
            continue;
        }
        const int eb = sh.off[cb];
        if (tid == 0) sh.nlong = 0;
        for (int i = tid; i < eb - ea; i += BWG) {
            int64_t c = (int64_t)colval[P0 + ea + i] - 1;
            sh.col[i] = (int)(c < 0 ? 0 : (c >= n ? n - 1 : c));
            sh.val[i] = vload(nzval + P0 + ea + i);
        }
        __syncthreads();
        for (int rr = ca + g; rr < cb; rr += groups) {
            const int64_t r = r0 + rr;
            const int s = sh.off[rr] - ea, e = sh.off[rr + 1] - ea;
            if (e - s > BATCH_LONG) {
                if (lane == 0) {
                    int q = atomicAdd(&sh.nlong, 1);
                    if (q < BATCH_CHUNK_LONG) sh.long_rows[q] = rr;
                }
                continue;
            }
            if constexpr (VW == 1) {
                for (int kk = lane; kk < k; kk += lpr) {
                    // 8 gathers in flight per lane; the sum stays left to right
                    T acc = T(0);
                    for (int j = s; j < e; j += 8) {
                        T pv[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const int jj = j + u < e ? j + u : e - 1;
                            pv[u] = sh.val[jj] * B[(int64_t)sh.col[jj] * bsr + kk * bsc];
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            if (j + u < e) acc = acc + pv[u];
                    }
                    emit(r, kk, acc);
                }
            } else {
                // row-major B: VW consecutive columns kk0.. per lane, one vector load
                // per stored entry, 8 in flight; each column's sum left to right
                using VT = typename VecT<T, VW>::type;
                for (int kk0 = lane * VW; kk0 < k; kk0 += lpr * VW) {
                    T acc[VW];
#pragma unroll
                    for (int c = 0; c < VW; ++c) acc[c] = T(0);
                    for (int j = s; j < e; j += 8) {
                        VT bv[8];
                        T av[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const int jj = j + u < e ? j + u : e - 1;
                            av[u] = sh.val[jj];
                            bv[u] = *reinterpret_cast<const VT*>(B + (int64_t)sh.col[jj] * bsr + kk0);
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            if (j + u < e) {
#pragma unroll
                                for (int c = 0; c < VW; ++c) acc[c] = acc[c] + av[u] * bv[u][c];
                            }
                    }
#pragma unroll
                    for (int c = 0; c < VW; ++c) emit(r, kk0 + c, acc[c]);
                }
            }
        }
        __syncthreads();
        // the chunk's long rows, from LDS
        const int nl = sh.nlong < BATCH_CHUNK_LONG ? sh.nlong : BATCH_CHUNK_LONG;
        for (int q = 0; q < nl; ++q) {
            const int rr = sh.long_rows[q];
            wg_row(r0 + rr, sh.off[rr] - ea, sh.off[rr + 1] - ea,
                   [&](int64_t j, int64_t& c, T& v) {
                       c = sh.col[j];
                       v = sh.val[j];
                   });
        }
        __syncthreads();  // the staging is reused by the next chunk
        ca = cb;
    }
}

// Persistent: tiles from the table's segments, each taken through its atomic
// counter (one per segment); a workgroup leaves when every segment is used up.
template <typename V, typename I>
__global__ __launch_bounds__(BWG) void k_batch_spmm(const gsa_batch_desc* __restrict__ desc,
                                                    const int32_t* __restrict__ tiles,
                                                    BatchSegs segs, int* __restrict__ counter,
                                                    typename Acc<V>::T alpha,
                                                    typename Acc<V>::T beta, int mode) {
    using T = typename Acc<V>::T;
    __shared__ int s_tile;
    __shared__ TileLds<T> sh;
    const int nseg = segs.n > 1 ? (segs.n < BATCH_SEGS ? segs.n : BATCH_SEGS) : 1;
    int seg = (int)(blockIdx.x % (unsigned)nseg), tried = 0;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(counter + seg, 1);
        __syncthreads();
        int64_t t = s_tile;
        __syncthreads();  // every lane has read s_tile before lane 0 takes the next
        if (t >= segs.start[seg + 1] - segs.start[seg]) {  // this segment is used up
            if (++tried >= nseg) break;
            seg = seg + 1 < nseg ? seg + 1 : 0;
            continue;
        }
        t += segs.start[seg];
        const gsa_batch_desc d = desc[tiles[3 * t]];
        const int64_t ta = tiles[3 * t + 1] > 0 ? tiles[3 * t + 1] : 0;
        const int64_t tb = tiles[3 * t + 2] < d.m ? tiles[3 * t + 2] : d.m;
        const I* __restrict__ rp = (const I*)d.rowptr;
        // pieces of at most BATCH_MAX_TILE_ROWS rows whose entries an int counts;
        // failing that, one row at a time (a row has at most n < 2^31 entries)
        for (int64_t a = ta; a < tb; a += BATCH_MAX_TILE_ROWS) {
            const int64_t b = a + BATCH_MAX_TILE_ROWS < tb ? a + BATCH_MAX_TILE_ROWS : tb;
            const int vw = batch_vw(d, (int)sizeof(T));
            auto run = [&](int64_t x, int64_t y) {
                if (vw == 4) process_tile<V, I, 4>(d, x, y, alpha, beta, mode, sh);
                else if (vw == 2) process_tile<V, I, 2>(d, x, y, alpha, beta, mode, sh);
                else process_tile<V, I, 1>(d, x, y, alpha, beta, mode, sh);
            };
            if ((int64_t)rp[b] - (int64_t)rp[a] <= (int64_t)INT32_MAX) {
                run(a, b);
            } else {
                for (int64_t r = a; r < b; ++r) run(r, r + 1);
            }
        }
    }
}

template <typename I>
__global__ void k_row_bf16(int64_t m, const I* __restrict__ rowptr, const void* __restrict__ nz,
                           int is_f64, uint16_t* __restrict__ out, float* __restrict__ scale) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = (int64_t)rowptr[r] - 1, e = (int64_t)rowptr[r + 1] - 1;
        double mx = 0;
        for (int64_t j = s; j < e; ++j) {
            double v = is_f64 ? ((const double*)nz)[j] : (double)((const float*)nz)[j];
            v = v < 0 ? -v : v;
            mx = v > mx ? v : mx;
        }
        // power-of-two scale >= max |a|: the division below is exact
        int ex = 0;
        if (mx > 0) (void)frexp(mx, &ex);
        const double sc = mx > 0 ? ldexp(1.0, ex) : 1.0;
        scale[r] = (float)sc;
        for (int64_t j = s; j < e; ++j) {
            double v = is_f64 ? ((const double*)nz)[j] : (double)((const float*)nz)[j];
            out[j] = f32_to_bf16((float)(v / sc));
        }
    }
}

// d_counter: segs.n ints of scratch
static int launch_batch(hipStream_t s, int vtype, int itype, const gsa_batch_desc* d_desc,
                        const int32_t* d_tiles, const BatchSegs& segs, int* d_counter,
                        double alpha, double beta, int grid) {
    const int mode = beta == 0.0 ? 0 : (beta == 1.0 ? 1 : 2);
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, d_counter, segs.n);
    if (segs.start[segs.n] == 0) return GSA_OK;
    if (grid <= 0) grid = 256 * 8;
    auto go = [&](auto vz, auto iz) {
        using V = decltype(vz);
        using I = decltype(iz);
        using T = typename Acc<V>::T;
        hipLaunchKernelGGL((k_batch_spmm<V, I>), dim3(grid), dim3(BWG), 0, s, d_desc, d_tiles, segs,
                           d_counter, (T)alpha, (T)beta, mode);
        return (int)GSA_OK;
    };
    int r;
    if (vtype == GSA_BF16) r = dispatch_i(itype, [&](auto iz) { return go(uint16_t{}, iz); });
    else r = dispatch_vi(vtype, itype, [&](auto vz, auto iz) { return go(vz, iz); });
    if (r != GSA_OK) return r;
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

// Tiles for `count` instances: tile rows = max(16, rows_per_tile / lanes per
// row).  The instances are dealt to the BATCH_SEGS segments largest first,
// each to the least loaded (work ~ entries weighted by k, plus rows), and keep
// their order inside a segment.
static void make_tiles(const gsa_batch_desc* h, const int64_t* nnz, int64_t count,
                       int rows_per_tile, int esize, std::vector<int32_t>& tiles,
                       BatchSegs& segs) {
    std::vector<int64_t> idx((size_t)count);
    std::vector<double> work((size_t)count);
    for (int64_t i = 0; i < count; ++i) {
        idx[(size_t)i] = i;
        work[(size_t)i] = (double)nnz[i] * (8.0 + h[i].k) / 8.0 + (double)h[i].m;
    }
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int64_t a, int64_t b) { return work[(size_t)a] > work[(size_t)b]; });
    std::vector<int> bin((size_t)count, 0);
    double load[BATCH_SEGS] = {0};
    for (int64_t i : idx) {
        int x = 0;
        for (int y = 1; y < BATCH_SEGS; ++y)
            if (load[y] < load[x]) x = y;
        bin[(size_t)i] = x;
        load[x] += work[(size_t)i];
    }
    segs.n = BATCH_SEGS;
    for (int x = 0; x < BATCH_SEGS; ++x) {
        segs.start[x] = (int64_t)tiles.size() / 3;
        for (int64_t i = 0; i < count; ++i) {
            if (bin[(size_t)i] != x) continue;
            const int lpr = batch_lpr(h[i], esize);
            int64_t rows = rows_per_tile / lpr;
            if (rows < 16) rows = 16;
            if (rows > BATCH_MAX_TILE_ROWS) rows = BATCH_MAX_TILE_ROWS;
            for (int64_t r = 0; r < h[i].m; r += rows) {
                tiles.push_back((int32_t)i);
                tiles.push_back((int32_t)r);
                tiles.push_back((int32_t)(r + rows < h[i].m ? r + rows : h[i].m));
            }
        }
    }
    segs.start[BATCH_SEGS] = (int64_t)tiles.size() / 3;
}

void batch_plan_release(gsa_plan_s* plan) {
    auto* bp = (BatchPlan*)plan->batch;
    if (!bp) return;
    (void)hipDeviceSynchronize();
    (void)hipFree(bp->d_desc);
    (void)hipFree(bp->d_tiles);
    delete bp;
    plan->batch = nullptr;
}

}  // namespace gsa

using namespace gsa;

extern "C" {

int gsa_batch_plan_create(void* stream, int vtype, int itype, int64_t count,
                          const gsa_batch_desc* host_desc, int rows_per_tile, gsa_plan* plan) {
    GSA_TRY(ensure_init());
    GSA_CHECK(plan && count >= 0 && (host_desc || count == 0), GSA_ERR_INVALID,
              "gsa_batch_plan_create: bad arguments");
    GSA_CHECK(vtype == GSA_F64 || vtype == GSA_F32 || vtype == GSA_BF16, GSA_ERR_UNSUPPORTED,
              "bad vtype %d", vtype);
    GSA_CHECK(itype == GSA_I32 || itype == GSA_I64, GSA_ERR_UNSUPPORTED, "bad itype %d", itype);
    if (rows_per_tile <= 0) rows_per_tile = 512;
    for (int64_t i = 0; i < count; ++i) {
        const gsa_batch_desc& d = host_desc[i];
        GSA_CHECK(d.m >= 0 && d.n >= 0 && d.k >= 0, GSA_ERR_INVALID, "instance %lld: negative size",
                  (long long)i);
        GSA_CHECK(d.m == 0 || (d.rowptr && d.C), GSA_ERR_INVALID, "instance %lld: null pointer",
                  (long long)i);
        const bool brow = (d.flags & GSA_BATCH_B_ROWMAJOR) != 0;
        const bool crow = (d.flags & GSA_BATCH_C_ROWMAJOR) != 0;
        GSA_CHECK(d.k < 1 || ((brow ? d.ldb >= d.k : (d.k == 1 || d.ldb >= d.n)) &&
                              (crow ? d.ldc >= d.k : (d.k == 1 || d.ldc >= d.m))),
                  GSA_ERR_INVALID, "instance %lld: leading dimension too small", (long long)i);
        GSA_CHECK(vtype != GSA_BF16 || d.row_scale || d.m == 0, GSA_ERR_INVALID,
                  "instance %lld: BF16 values need row_scale", (long long)i);
    }
    hipStream_t s = as_stream(stream);
    std::vector<int32_t> tiles;
    auto* bp = new BatchPlan();
    bp->vtype = vtype; bp->itype = itype; bp->count = count;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    bp->grid = cus * 8;
    std::vector<int64_t> nnz((size_t)count);
    int64_t* d_nnz = nullptr;
    hipError_t e = hipMalloc(&bp->d_desc, sizeof(gsa_batch_desc) * (count > 0 ? count : 1));
    if (e == hipSuccess && count > 0)
        e = hipMemcpyAsync(bp->d_desc, host_desc, sizeof(gsa_batch_desc) * count,
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess && count > 0) e = hipMalloc(&d_nnz, sizeof(int64_t) * count);
    if (e == hipSuccess && count > 0) {
        if (itype == GSA_I32)
            hipLaunchKernelGGL(k_inst_nnz<int32_t>, dim3((unsigned)cdiv(count, 256)), dim3(256), 0, s,
                               bp->d_desc, count, d_nnz);
        else
            hipLaunchKernelGGL(k_inst_nnz<int64_t>, dim3((unsigned)cdiv(count, 256)), dim3(256), 0, s,
                               bp->d_desc, count, d_nnz);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(nnz.data(), d_nnz, sizeof(int64_t) * count, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (d_nnz) (void)hipFree(d_nnz);
    if (e == hipSuccess) {
        make_tiles(host_desc, nnz.data(), count, rows_per_tile, vtype == GSA_F64 ? 8 : 4, tiles,
                   bp->segs);
        bp->ntiles = (int64_t)tiles.size() / 3;
        e = hipMalloc(&bp->d_tiles, sizeof(int32_t) * (tiles.size() + 3));
    }
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpyAsync(bp->d_tiles, tiles.data(), sizeof(int32_t) * tiles.size(),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // host vectors go out of scope
    if (e != hipSuccess) {
        (void)hipFree(bp->d_desc); (void)hipFree(bp->d_tiles);
        delete bp;
        return hip_fail(e, "gsa_batch_plan_create");
    }
    auto* p = new gsa_plan_s();
    p->kind = 1;
    p->batch = bp;
    *plan = p;
    return GSA_OK;
}

int gsa_batch_spmm(void* stream, gsa_plan plan, float alpha, float beta) {
    GSA_CHECK(plan && plan->kind == 1 && plan->batch, GSA_ERR_INVALID, "gsa_batch_spmm: not a batch plan");
    auto* bp = (BatchPlan*)plan->batch;
    hipStream_t s = as_stream(stream);
    Scratch counter(s);  // this launch's tile counters, one per segment
    GSA_TRY(counter.alloc(sizeof(int) * BATCH_SEGS));
    return launch_batch(s, bp->vtype, bp->itype, bp->d_desc, bp->d_tiles, bp->segs,
                        counter.as<int>(), alpha, beta, bp->grid);
}

int gsa_batch_spmm_tables(void* stream, int vtype, int itype, const gsa_batch_desc* d_desc,
                          const int32_t* d_tiles, int64_t ntiles, int* d_counter, float alpha,
                          float beta, int grid) {
    GSA_CHECK(ntiles >= 0 && d_counter && (ntiles == 0 || (d_desc && d_tiles)), GSA_ERR_INVALID,
              "gsa_batch_spmm_tables: bad arguments");
    BatchSegs one{};
    one.n = 1;
    one.start[0] = 0;
    one.start[1] = ntiles;
    return launch_batch(as_stream(stream), vtype, itype, d_desc, d_tiles, one, d_counter, alpha,
                        beta, grid);
}

int gsa_batch_graph_create(void* stream, gsa_plan plan, float alpha, float beta, void** graph_exec) {
    GSA_CHECK(plan && plan->kind == 1 && graph_exec, GSA_ERR_INVALID, "gsa_batch_graph_create: bad arguments");
    hipStream_t s = as_stream(stream);
    hipGraph_t g = nullptr;
    GSA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int r = gsa_batch_spmm(stream, plan, alpha, beta);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (r != GSA_OK) { if (g) (void)hipGraphDestroy(g); return r; }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *graph_exec = ex;
    return GSA_OK;
}

int gsa_graph_launch(void* graph_exec, void* stream) {
    GSA_CHECK(graph_exec, GSA_ERR_INVALID, "gsa_graph_launch: null graph");
    GSA_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, as_stream(stream)));
    return GSA_OK;
}

int gsa_graph_destroy(void* graph_exec) {
    if (graph_exec) GSA_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return GSA_OK;
}

int gsa_csr_to_bf16(void* stream, int vtype, int itype, int64_t m, const void* rowptr,
                    const void* nzval, uint16_t* nzval_bf16, float* row_scale) {
    GSA_TRY(ensure_init());
    GSA_CHECK(m >= 0 && (vtype == GSA_F64 || vtype == GSA_F32), GSA_ERR_INVALID,
              "gsa_csr_to_bf16: bad arguments");
    if (m == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_i(itype, [&](auto iz) {
        using I = decltype(iz);
        hipLaunchKernelGGL(k_row_bf16<I>, dim3(grid_for(m, 256)), dim3(256), 0, s, m, (const I*)rowptr,
                           nzval, vtype == GSA_F64 ? 1 : 0, nzval_bf16, row_scale);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    });
}

}  // extern "C"
