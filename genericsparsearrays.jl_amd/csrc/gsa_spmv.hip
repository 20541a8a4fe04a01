// gsa_spmv.hip — mul! for CSR / CSC / COO on gfx950.
//
// CSR gather SpMV (CSR-N, CSC-T, and COO-N / CSC-N through a row view):
// "flat windows".  Wavefront w reduces exactly the stored entries
// [576w, 576(w+1)): it streams their indices and values with non-temporal
// loads, gathers x, stages the products in LDS and sums the rows that start
// in its window (the plan: one lower bound per window, built in O(m) on the
// device).  Rows of at most 32 entries are summed left to right on one lane —
// the summation order of kernel_spmatmul_csr_N!
// (src/matrix_csr/matrix_csr_kernels.jl:15-20), so they are bit-identical to
// the reference, including the ones that cross a window end (finished from 32
// entries staged past it).  Longer rows are reduced by the whole wavefront;
// those that run past the window leave partial sums that k_spmv_fold adds in
// entry order.  Every output element is written once: no atomics, and the
// result does not depend on scheduling.  A plan given its column array
// launches the windows inside long sorted rows on the XCD whose L2 holds the
// eighth of x they gather from.
//
// SpMM (k > 1) uses 2048-entry row blocks per workgroup; the dense operand is
// read as a row-major n x k array (transposed once into scratch when it is
// column-major), so the k values a stored entry needs are one contiguous
// segment, and each (row, k) output is summed on one lane in storage order.
//
// Scatter-shaped products without a row view (CSR-T, CSC-N, COO-T, COO SpMM)
// add into C with hardware float atomics, as the reference's @atomic kernels
// do (matrix_csr_kernels.jl:24-45, src/matrix_csc/matrix_csc_kernels.jl:1-22,
// src/matrix_coo/matrix_coo_kernels.jl:1-43); in the COO kernel, runs of
// equal output index are pre-reduced in the wavefront, one atomic per run.
#include "gsa_common.h"
#include "gsa_rows.h"
#include "gsa_scan.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

// The scratch of gsa_spmv_csr_noalloc: the window table and the head/tail
// partial sums of one SpMV, in a fixed device array (one copy per GPU, at its
// own address on each).  Host code never passes this name to a kernel: it is
// a host-side stand-in; the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) uint8_t scratch[GSA_NOALLOC_SCRATCH_BYTES];

namespace gsa {

constexpr int WG = 256;

__device__ __forceinline__ int64_t clamp64(int64_t v, int64_t lo, int64_t hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Streamed-once operands (indices, values, row pointers) are read with
// non-temporal loads so they do not evict the gathered x from L2.
template <typename U>
__device__ __forceinline__ U ldnt(const U* p) {
    if constexpr (VT<U>::cx) {  // complex: the two parts
        using R = decltype(p->re);
        const R* q = reinterpret_cast<const R*>(p);
        return U(__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1));
    } else {
        return __builtin_nontemporal_load(p);
    }
}

// Stored value -> arithmetic type: identity, or BF16 bits -> Float32.
template <typename T, typename V>
__device__ __forceinline__ T widen(V v) {
    if constexpr (sizeof(V) == 2) return __uint_as_float(((uint32_t)v) << 16);
    else return (T)v;
}

// Stored value j, or, through a row view of a COO/CSC matrix, the value at
// storage position perm[j].
template <typename PI, typename V>
__device__ __forceinline__ V load_val(const V* __restrict__ nz, const PI* __restrict__ perm,
                                      int64_t j) {
    if constexpr (std::is_void_v<PI>) return ldnt(nz + j);
    else return nz[(int64_t)__builtin_nontemporal_load(perm + j)];
}

// Row sum times the row's scale (BF16 storage), or unchanged.
template <typename T>
__device__ __forceinline__ T row_scaled(T acc, const float* __restrict__ rscale, int64_t r) {
    if constexpr (std::is_floating_point_v<T>) return rscale ? acc * (T)rscale[r] : acc;
    else return acc;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = v + lane_xor(v, off);
    return v;
}

// Knuth's TwoSum: s + e == a + b exactly (this library is built without FMA
// contraction).
template <typename T>
__device__ __forceinline__ void two_sum(T a, T b, T& s, T& e) {
    s = a + b;
    T bp = s - a;
    T ap = s - bp;
    e = (a - ap) + (b - bp);
}

// Float32 rows reduced across lanes are summed with TwoSum: each lane keeps a
// (sum, compensation) pair, pairs are combined error-free in the butterfly.
// Float64, complex and integer values keep the plain sum.
template <typename T>
struct CompSum {
    T s = T(0), c = T(0);
    __device__ __forceinline__ void add(T p) {
        if constexpr (VT<T>::f32) {
            T ns, er;
            two_sum(s, p, ns, er);
            s = ns;
            c = c + er;
        } else {
            s = s + p;
        }
    }
    // this lane's compensated total
    __device__ __forceinline__ T total() const {
        if constexpr (VT<T>::f32) return s + c;
        else return s;
    }
    __device__ __forceinline__ T wave_total() {
        if constexpr (VT<T>::f32) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                T os = lane_xor(s, off), oc = lane_xor(c, off);
                T ns, er;
                two_sum(s, os, ns, er);
                s = ns;
                c = c + oc + er;
            }
            return s + c;
        } else {
            return wave_sum(s);
        }
    }
};

// ------------------------------------------------------------ plan kernels --
template <typename I>
__global__ void k_csr_plan(int64_t m, int64_t nnz, const I* __restrict__ ptr,
                           int64_t* __restrict__ blk_row, int64_t nblocks, int nb) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r > m) return;
    int64_t cur = (r == m) ? nnz : clamp64((int64_t)ptr[r] - 1, 0, nnz);
    int64_t bstart = 0;
    if (r > 0) bstart = clamp64((int64_t)ptr[r - 1] - 1, 0, nnz) / nb + 1;
    int64_t bend = cur / nb;
    for (int64_t b = bstart; b <= bend && b < nblocks; ++b) blk_row[b] = r;
    if (r == m) blk_row[nblocks] = m;
}

// Per window: P0, Ps, Pend, spill (see PLAN_POS).  The last owned row is
// "spilled" (reduced from HBM by the whole workgroup) when the window's rows
// do not fit the LDS staging capacity.
template <typename I>
__global__ void k_csr_plan_pos(int64_t m, int64_t nnz, const I* __restrict__ ptr,
                               const int64_t* __restrict__ blk_row, int64_t nblocks,
                               int64_t* __restrict__ pos, int cap) {
    int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    int64_t r0 = clamp64(blk_row[b], 0, m), r1 = clamp64(blk_row[b + 1], r0, m);
    int64_t p0 = (r0 < m) ? clamp64((int64_t)ptr[r0] - 1, 0, nnz) : nnz;
    int64_t pend = clamp64((int64_t)ptr[r1] - 1, p0, nnz);
    int64_t ps = pend, spill = 0;
    if (r1 > r0 && pend - p0 > cap) {
        ps = clamp64((int64_t)ptr[r1 - 1] - 1, p0, p0 + cap);
        spill = 1;
    }
    pos[PLAN_POS * b + 0] = p0;
    pos[PLAN_POS * b + 1] = ps;
    pos[PLAN_POS * b + 2] = pend;
    pos[PLAN_POS * b + 3] = spill;
}

int build_csr_plan(hipStream_t s, int itype, int64_t m, int64_t nnz, const void* ptr,
                   int64_t* blk_row, int64_t nblocks, int64_t* blk_pos, int nb, int cap) {
    GSA_TRY(zero_fill(s, blk_row, sizeof(int64_t) * (nblocks + 1)));
    return dispatch_i(itype, [&](auto ti) {
        using I = decltype(ti);
        unsigned g = (unsigned)cdiv(m + 1, 256);
        hipLaunchKernelGGL(k_csr_plan<I>, dim3(g), dim3(256), 0, s, m, nnz,
                           (const I*)ptr, blk_row, nblocks, nb);
        if (blk_pos)
            hipLaunchKernelGGL(k_csr_plan_pos<I>, dim3((unsigned)cdiv(nblocks, 256)), dim3(256), 0, s,
                               m, nnz, (const I*)ptr, (const int64_t*)blk_row, nblocks, blk_pos, cap);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    });
}

// Window of block b under the plan.
struct Window {
    int64_t r0, r1, P0, Ps, Pend, rs_end;
    bool spill;
};

__device__ __forceinline__ Window load_window(int64_t b, int64_t m, int64_t nnz,
                                              const int64_t* __restrict__ blk_row,
                                              const int64_t* __restrict__ pos,
                                              int cap = SPMV_CAP) {
    Window w;
    w.r0 = clamp64(blk_row[b], 0, m);
    w.r1 = clamp64(blk_row[b + 1], w.r0, m);
    w.P0 = clamp64(pos[PLAN_POS * b + 0], 0, nnz);
    w.Pend = clamp64(pos[PLAN_POS * b + 2], w.P0, nnz);
    w.Ps = clamp64(pos[PLAN_POS * b + 1], w.P0, w.P0 + cap);
    if (w.Ps > w.Pend) w.Ps = w.Pend;
    w.spill = pos[PLAN_POS * b + 3] != 0 && w.r1 > w.r0;
    w.rs_end = w.spill ? w.r1 - 1 : w.r1;
    return w;
}

// ------------------------------------------------------------ SpMV kernel --
// Flat windows: wavefront w reduces exactly the entries [w*SPMV_WIN,
// (w+1)*SPMV_WIN), so every wavefront streams and gathers a full window
// whatever the row lengths (the gathers of x are the bound; a half-empty
// wavefront costs nearly as much as a full one).  The wavefronts of a
// workgroup are independent (no workgroup barrier).
//
// Window w owns the rows that start in it.  A row is summed where it lies:
//   * an owned row that ends inside the window (or a short row, at most
//     SPMV_WX entries, that ends within the SPMV_WX entries staged past the
//     window) is written directly.  Rows of at most W_LONG entries are summed
//     left to right on one lane, the order of the reference loop, so they are
//     bit-identical to it; longer ones by the whole wavefront;
//   * a longer row that runs past the window leaves partial sums: its owner's
//     part ("tail") and each later window's part ("head").  k_spmv_fold adds
//     them in entry order and writes the row.
constexpr int WAVES = 4;
constexpr int WU = SPMV_WIN / 64;        // window entries per lane
constexpr int WS = SPMV_WIN + SPMV_WX;   // staged entries per window
constexpr int WRP = 256;                 // row offsets staged per wavefront
constexpr int WRPL = WRP / 64;
constexpr int W_LONG = SPMV_WX;          // rows longer than this: the whole wavefront

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Zero-based [start, end) of row r, clamped to [0, nnz].
template <typename I>
__device__ __forceinline__ void row_span(const I* __restrict__ ptr, int64_t r, int64_t nnz,
                                         int64_t& a, int64_t& b) {
    a = clamp64((int64_t)ptr[r] - 1, 0, nnz);
    b = clamp64((int64_t)ptr[r + 1] - 1, a, nnz);
}

// Does window w's last owned row (r1-1) leave a tail: longer than W_LONG and
// running past the window?  Shared by the SpMV kernel and the fold.
template <typename I>
__device__ __forceinline__ bool window_tail(const I* __restrict__ ptr, int64_t r0, int64_t r1,
                                            int64_t e1, int64_t nnz, int64_t& end) {
    if (r1 <= r0) return false;
    int64_t a;
    row_span(ptr, r1 - 1, nnz, a, end);
    return end > e1 && end - a > W_LONG;
}

template <typename T, typename I, typename V = T, typename PI = void, typename CI = I>
__global__ __launch_bounds__(64 * WAVES) void k_csr_spmv(
    int64_t m, int64_t n, int64_t nnz, const I* __restrict__ rowptr,
    const CI* __restrict__ colval, const V* __restrict__ nzval, const PI* __restrict__ perm,
    const float* __restrict__ rscale, const T* __restrict__ x, T* __restrict__ y, T alpha,
    T beta, int mode, int cj, const int64_t* __restrict__ win_row, int64_t nwin,
    const int64_t* __restrict__ order, T* __restrict__ head, T* __restrict__ tail) {
    __shared__ T s_prod[WAVES][WS];
    __shared__ int s_off[WAVES][WRP];

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int64_t slot = (int64_t)blockIdx.x * WAVES + wv;
    const int64_t w = order ? order[slot] : slot;
    if (w < 0 || w >= nwin) return;
    auto emit = [&](int64_t r, T acc) {
        store_out(y, r, row_scaled(acc, rscale, r), alpha, beta, mode);
    };
    const int64_t E0 = w * SPMV_WIN < nnz ? w * SPMV_WIN : nnz;
    const int64_t E1 = E0 + SPMV_WIN < nnz ? E0 + SPMV_WIN : nnz;
    const int cnt = (int)(E1 - E0);
    const int64_t r0 = clamp64(win_row[w], 0, m);
    const int64_t r1 = clamp64(win_row[w + 1], r0, m);
    const int64_t nrows = r1 - r0;
    const int nst = (int)(nrows < WRP - 1 ? nrows : WRP - 1);  // offsets r0..r0+nst staged
    const bool rp_in_lds = nrows + 1 <= WRP;
    T* prod = s_prod[wv];
    int* off = s_off[wv];

    // the last owned row: a tail, or a short row finished from the extension
    int64_t last_end = 0;
    const bool has_tail = window_tail(rowptr, r0, r1, E1, nnz, last_end);
    const int need = (nrows > 0 && !has_tail && last_end > E1) ? (int)(last_end - E1) : 0;
    const int staged = cnt + need;  // <= SPMV_WIN + SPMV_WX

    // 1. row offsets, indices and values in flight together, then the gathers
    I rpv[WRPL];
#pragma unroll
    for (int k = 0; k < WRPL; ++k) {
        int i = lane + 64 * k;
        i = i <= nst ? i : nst;
        rpv[k] = ldnt(rowptr + r0 + i);
    }
    if (staged > 0) {
        CI ci[WU + 1];
        T vv[WU + 1], xv[WU + 1];
#pragma unroll
        for (int u = 0; u <= WU; ++u) {
            int i = u * 64 + lane;
            i = i < staged ? i : staged - 1;
            ci[u] = ldnt(colval + E0 + i);
            vv[u] = conj_if(widen<T>(load_val(nzval, perm, E0 + i)), cj & 1);
        }
#pragma unroll
        for (int u = 0; u <= WU; ++u)
            xv[u] = conj_if(x[clamp64((int64_t)ci[u] - 1, 0, n - 1)], (cj & 2) != 0);
#pragma unroll
        for (int u = 0; u <= WU; ++u) {
            int i = u * 64 + lane;
            if (i < staged) prod[i] = vv[u] * xv[u];
        }
    }
    const int64_t EX = E0 + staged;
#pragma unroll
    for (int k = 0; k < WRPL; ++k) {
        int i = lane + 64 * k;
        if (i <= nst) off[i] = (int)(clamp64((int64_t)rpv[k] - 1, E0, EX) - E0);
    }
    wave_lds_sync();

    // 2. head: entries [0, hend) belong to a row started in an earlier window;
    //    short ones were finished by their owner
    const int hend = off[0] < cnt ? off[0] : cnt;
    if (hend > 0 && r0 > 0) {
        int64_t a, b;
        row_span(rowptr, r0 - 1, nnz, a, b);
        if (b - a > W_LONG) {
            CompSum<T> cs;
            for (int j = lane; j < hend; j += 64) cs.add(prod[j]);
            const T acc = cs.wave_total();
            if (lane == 0) head[w] = acc;
        }
    }

    // 3. owned rows: one lane per short row (left to right); longer rows and
    //    the tail by the whole wavefront, one at a time
    for (int64_t base = r0; base < r1; base += 64) {
        const int64_t r = base + lane;
        const bool act = r < r1;
        int s = 0, e = 0;
        if (act) {
            if (rp_in_lds) {
                s = off[r - r0];
                e = off[r - r0 + 1];
            } else {
                s = (int)(clamp64((int64_t)rowptr[r] - 1, E0, EX) - E0);
                e = (int)(clamp64((int64_t)rowptr[r + 1] - 1, E0, EX) - E0);
            }
        }
        const bool is_tail = act && has_tail && r == r1 - 1;
        if (is_tail) e = cnt;  // this window's part only
        const bool lng = act && (is_tail || e - s > W_LONG);
        if (act && !lng) {
            T acc = T(0);
            for (int j = s; j < e; ++j) acc = acc + prod[j];
            emit(r, acc);
        }
        uint64_t lm = __ballot(lng);
        while (lm) {
            const int src = __ffsll((long long)lm) - 1;
            lm &= lm - 1;
            const int ls = __shfl(s, src, 64), le = __shfl(e, src, 64);  // int lanes
            CompSum<T> cs;
            for (int j = ls + lane; j < le; j += 64) cs.add(prod[j]);
            const T acc = cs.wave_total();
            if (lane == 0) {
                if (has_tail && base + src == r1 - 1) tail[w] = acc;
                else emit(base + src, acc);
            }
        }
    }
}

// Rows that ran past their window: the owner's tail then every later
// window's head, in entry order (compensated for Float32).  One thread per
// window.
template <typename T, typename I>
__global__ void k_spmv_fold(int64_t m, int64_t nnz, const I* __restrict__ rowptr,
                            const int64_t* __restrict__ win_row, int64_t nwin,
                            const T* __restrict__ head, const T* __restrict__ tail,
                            const float* __restrict__ rscale, T* __restrict__ y, T alpha, T beta,
                            int mode) {
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (w >= nwin) return;
    const int64_t E0 = w * SPMV_WIN < nnz ? w * SPMV_WIN : nnz;
    const int64_t E1 = E0 + SPMV_WIN < nnz ? E0 + SPMV_WIN : nnz;
    const int64_t r0 = clamp64(win_row[w], 0, m);
    const int64_t r1 = clamp64(win_row[w + 1], r0, m);
    int64_t end = 0;
    if (!window_tail(rowptr, r0, r1, E1, nnz, end)) return;
    int64_t we = (end - 1) / SPMV_WIN;
    we = we < nwin - 1 ? we : nwin - 1;
    CompSum<T> cs;
    cs.add(tail[w]);
    // heads in order; eight loads in flight at a time
    for (int64_t v = w + 1; v <= we; v += 8) {
        T hv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) hv[u] = v + u <= we ? head[v + u] : T(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (v + u <= we) cs.add(hv[u]);
    }
    const int64_t r = r1 - 1;
    store_out(y, r, row_scaled(cs.total(), rscale, r), alpha, beta, mode);
}

// ---- launch order of the SpMV windows ----------------------------------------
// A window whose middle entry lies in a row longer than SPMV_LOCAL_ROW entries
// gathers from a narrow band of x (the row's columns ascend), so it is given
// to the XCD whose eighth of the columns holds that entry.  Workgroups go to
// the 8 XCDs round-robin: workgroup b runs on XCD b % 8, so XCD x's windows go
// to slots of workgroups 8j + x, and its L2 keeps that eighth of x.  The other
// windows fill every XCD's list up to an even share.
constexpr int SPMV_XCDS = 8;
constexpr int64_t SPMV_LOCAL_ROW = SPMV_WIN;

template <typename I>
__global__ void k_win_bucket(int64_t m, int64_t n, int64_t nnz, const I* __restrict__ rowptr,
                             const I* __restrict__ colval, const int64_t* __restrict__ win_row,
                             int64_t nwin, int* __restrict__ bucket) {
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (w >= nwin) return;
    const int64_t E0 = w * SPMV_WIN < nnz ? w * SPMV_WIN : nnz;
    const int64_t E1 = E0 + SPMV_WIN < nnz ? E0 + SPMV_WIN : nnz;
    int b = SPMV_XCDS;  // any XCD
    if (E1 > E0 && m > 0) {
        const int64_t mid = E0 + (E1 - E0) / 2;
        const int64_t r0 = clamp64(win_row[w], 0, m), r1 = clamp64(win_row[w + 1], r0, m);
        // the row holding `mid`: the last row in [r0-1, r1-1] starting at or before it
        int64_t lo = r0 > 0 ? r0 - 1 : 0, hi = r1 > lo ? r1 - 1 : lo;
        while (lo < hi) {
            const int64_t md = (lo + hi + 1) / 2;
            if ((int64_t)rowptr[md] - 1 <= mid) lo = md;
            else hi = md - 1;
        }
        const int64_t len = (int64_t)rowptr[lo + 1] - rowptr[lo];
        if (len > SPMV_LOCAL_ROW && n > 0) {
            const int64_t c = clamp64((int64_t)colval[mid] - 1, 0, n - 1);
            b = (int)(c * SPMV_XCDS / n);
        }
    }
    bucket[w] = b;
}

// ------------------------------------------------------------ SpMM kernel --
constexpr int SPMM_LONG = 64;  // staged rows up to this length are summed by one group
constexpr int SPMM_RP = 1024;  // row offsets staged in LDS per workgroup
// outputs (m·k) from which a column-major C goes through row-major scratch
constexpr int64_t SPMM_VIA_ROWS_MIN = int64_t(1) << 21;

// Dense operand read as BT[col*ldbt + kk] (row-major).  LPR lanes per row.
template <typename T, typename I, typename CI, int LPR, typename V = T, typename PI = void,
          typename SC = I>
__global__ __launch_bounds__(WG) void k_csr_spmm(
    int64_t m, int64_t n, int64_t k, int64_t nnz, const I* __restrict__ rowptr,
    const SC* __restrict__ colval, const V* __restrict__ nzval, const PI* __restrict__ perm,
    const float* __restrict__ rscale,
    const T* __restrict__ BT,
    int64_t ldbt, T* __restrict__ C, int64_t ldc, int64_t ldk, T alpha, T beta, int mode,
    int cj, const int64_t* __restrict__ blk_row, const int64_t* __restrict__ blk_pos) {
    const bool cjb = (cj & 2) != 0;
    // mode < 0: the value itself (row-major scratch that k_from_rowmajor finishes)
    auto out = [&](int64_t idx, T v) {
        if (mode < 0) C[idx] = v;
        else store_out(C, idx, v, alpha, beta, mode);
    };
    // output (r, kk) at C[r * ldc + kk * ldk]: column-major (1, ld) or row-major (ld, 1)
    __shared__ CI s_col[SPMV_CAP];
    __shared__ T s_val[SPMV_CAP];
    __shared__ T red[WG];
    __shared__ int s_off[SPMM_RP];
    // per group: the row whose head part it left in red, -1 for none, -2 for an empty share
    __shared__ int64_t s_hrow[WG / LPR];
    // share q: first entry, first row at or after it, the row it starts inside (or -1)
    __shared__ int s_bnd[WG / LPR + 1];
    __shared__ int64_t s_first[WG / LPR + 1], s_head[WG / LPR + 1];

    const int tid = threadIdx.x;
    const Window w = load_window(blockIdx.x, m, nnz, blk_row, blk_pos);
    if (w.r0 >= w.r1) return;
    const int64_t P0 = w.P0, Ps = w.Ps, Pend = w.Pend, r0 = w.r0, r1 = w.r1;
    const int64_t rs_end = w.rs_end;
    const int cnt = (int)(Ps - P0);
    const int nst = (int)(rs_end - r0 < SPMM_RP - 1 ? rs_end - r0 : SPMM_RP - 1);
    const bool rp_in_lds = rs_end - r0 + 1 <= SPMM_RP;

    // stage row offsets, column indices and values (one round trip)
    I rpv[SPMM_RP / WG];
#pragma unroll
    for (int q = 0; q < SPMM_RP / WG; ++q) {
        int i = tid + q * WG;
        i = i <= nst ? i : nst;
        rpv[q] = ldnt(rowptr + r0 + i);
    }
    constexpr int U = 8;
    for (int base = 0; base < cnt; base += WG * U) {
        SC ci[U];
        T vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int i = base + u * WG + tid;
            i = i < cnt ? i : cnt - 1;
            ci[u] = ldnt(colval + P0 + i);
            vv[u] = conj_if(widen<T>(load_val(nzval, perm, P0 + i)), cj & 1);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int i = base + u * WG + tid;
            if (i < cnt) {
                s_col[i] = (CI)clamp64((int64_t)ci[u] - 1, 0, n - 1);
                s_val[i] = vv[u];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < SPMM_RP / WG; ++q) {
        int i = tid + q * WG;
        if (i <= nst) s_off[i] = (int)(clamp64((int64_t)rpv[q] - 1, P0, Ps) - P0);
    }
    __syncthreads();

    constexpr int GROUPS = WG / LPR;
    const int g = tid / LPR, lane = tid % LPR;
    // staged rows 0..nrows-1 (relative to r0) cover entries [off(0), off(nrows)) of the
    // LDS stage; each group of LPR lanes takes a contiguous share of those entries
    const int64_t nrows = rs_end - r0;
    auto off = [&](int64_t i) -> int {
        return rp_in_lds ? s_off[i]
                         : (int)(clamp64((int64_t)rowptr[r0 + i] - 1, P0, Ps) - P0);
    };
    const int begin = nrows > 0 ? off(0) : 0, end = nrows > 0 ? off(nrows) : 0;
    // Share boundaries: the even split of [begin, end) at p_q = begin + D*q/GROUPS,
    // moved back to the start of the row it falls in when that row has at most
    // SPMM_LONG entries, so such a row is summed by one group, left to right, as
    // the reference loop sums it.  Each row writes the boundaries that fall in it
    // (q with off(i) <= p_q < off(i+1), found by division), with the first row at
    // or after the boundary and the row it cuts (whose "head" part the share
    // takes over), so no group searches for them.
    const int D = end - begin;
    if (tid <= GROUPS && (tid == 0 || tid == GROUPS || D == 0)) {
        s_bnd[tid] = tid == GROUPS ? end : begin;
        s_first[tid] = 0;
        s_head[tid] = -1;
    }
    if (D > 0) {
        for (int64_t i = tid; i < nrows; i += WG) {
            const int a = off(i), b = off(i + 1);
            if (b <= a) continue;  // an empty row holds no position
            int q = (int)(((int64_t)(a - begin) * GROUPS + D - 1) / D);
            const int qe = (int)(((int64_t)(b - begin) * GROUPS + D - 1) / D);  // q < qe
            q = q > 1 ? q : 1;
            for (; q < qe && q < GROUPS; ++q) {
                const int p = begin + (int)((int64_t)D * q / GROUPS);
                if (b - a <= SPMM_LONG || p == a) {
                    int64_t j = i;  // empty rows just before i start there too
                    while (j > 0 && off(j - 1) == a) --j;
                    s_bnd[q] = a;
                    s_first[q] = j;
                    s_head[q] = -1;
                } else {
                    s_bnd[q] = p;
                    s_first[q] = i + 1;
                    s_head[q] = i;
                }
            }
        }
    }
    __syncthreads();
    const int js = s_bnd[g], je = s_bnd[g + 1];
    // rows are walked in order; the first row of a share may have started in an
    // earlier share (its "head" part is left in LDS for the row's owner), the last
    // may run on into later shares (its owner adds their head parts after a barrier)
    const int64_t first = s_first[g];
    const bool has_head = js < je && s_head[g] >= 0;
    const int64_t head_row = has_head ? s_head[g] : -1;
    for (int64_t k0 = 0; k0 < k; k0 += LPR) {
        const int64_t kk = k0 + lane;
        const bool kin = kk < k;
        int64_t cur = has_head ? head_row : first;
        int cend = (js < je) ? off(cur + 1) : 0;
        T acc = T(0);
        auto put = [&](int64_t r, T v) {  // the finished row r (absolute)
            if (kin) out(r * ldc + kk * ldk, v);
        };
        // row `cur` is finished: a head part to LDS, an owned row to C
        auto flush = [&]() {
            if (cur == head_row) {
                red[g * LPR + lane] = acc;
            } else {
                const int64_t r = r0 + cur;
                put(r, row_scaled(acc, rscale, r));
            }
            acc = T(0);
            ++cur;
        };
        for (int j = js; j < je; j += 8) {
            T av[8], bv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {  // entries past the share's end issue no load
                const int jj = j + u < je ? j + u : je - 1;
                av[u] = s_val[jj];
                bv[u] = (j + u < je && kin) ? conj_if(BT[(int64_t)s_col[jj] * ldbt + kk], cjb)
                                            : T(0);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {  // in entry order, closing the rows that end first
                if (j + u < je) {
                    while (cend <= j + u) {  // empty rows too
                        flush();
                        cend = off(cur + 1);
                    }
                    acc = acc + av[u] * bv[u];
                }
            }
        }
        int64_t tail_row = -1;
        if (js < je) {
            if (cend == je) {  // the share ends with its row
                flush();
            } else if (cur == head_row) {  // the share lies inside one row owned earlier
                red[g * LPR + lane] = acc;
                ++cur;
            } else {
                tail_row = cur++;
            }
            // empty rows at the end of the stage belong to the share that reaches it
            if (je == end)
                for (; cur < nrows; ++cur) put(r0 + cur, row_scaled(T(0), rscale, r0 + cur));
        } else if (g == 0 && begin == end) {  // a stage of empty rows only
            for (int64_t i = 0; i < nrows; ++i)
                if (kin) {
                    const int64_t r = r0 + i;
                    out(r * ldc + kk * ldk, row_scaled(T(0), rscale, r));
                }
        }
        if (lane == 0) s_hrow[g] = js == je ? -2 : head_row;
        __syncthreads();
        if (tail_row >= 0) {  // later shares' parts of this row, in entry order
            for (int h = g + 1; h < GROUPS; ++h) {
                const int64_t hr = s_hrow[h];
                if (hr == -2) continue;
                if (hr != tail_row) break;
                acc = acc + red[h * LPR + lane];
            }
            if (kin) {
                const int64_t r = r0 + tail_row;
                out(r * ldc + kk * ldk, row_scaled(acc, rscale, r));
            }
        }
        __syncthreads();
    }

    if (rs_end < r1) {  // spill row: staged in LDS chunk by chunk, groups split each chunk
        const int64_t r = r1 - 1;
        for (int64_t k0 = 0; k0 < k; k0 += LPR) {
            const int64_t kk = k0 + lane;
            T acc = T(0);
            for (int64_t c0 = Ps; c0 < Pend; c0 += SPMV_CAP) {
                const int cn = (int)(Pend - c0 < SPMV_CAP ? Pend - c0 : SPMV_CAP);
                __syncthreads();  // previous users of s_col/s_val are done
                for (int i = tid; i < cn; i += WG) {
                    s_col[i] = (CI)clamp64((int64_t)ldnt(colval + c0 + i) - 1, 0, n - 1);
                    s_val[i] = conj_if(widen<T>(load_val(nzval, perm, c0 + i)), cj & 1);
                }
                __syncthreads();
                if (kk < k) {
                    const int per = (cn + GROUPS - 1) / GROUPS;
                    const int js = g * per, je = (js + per < cn) ? js + per : cn;
                    for (int j = js; j < je; j += 8) {
                        T av[8], bv[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const int jj = j + u < je ? j + u : je - 1;
                            av[u] = s_val[jj];
                            bv[u] = conj_if(BT[(int64_t)s_col[jj] * ldbt + kk], cjb);
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            if (j + u < je) acc = acc + av[u] * bv[u];
                    }
                }
            }
            __syncthreads();
            T v = acc;
#pragma unroll
            for (int o = LPR; o < 64; o <<= 1) v = v + lane_xor(v, o);
            if ((tid & 63) < LPR) red[(tid >> 6) * LPR + lane] = v;
            __syncthreads();
            if (tid < LPR && kk < k) {
                T sum = red[lane];
                for (int wq = 1; wq < WG / 64; ++wq) sum = sum + red[wq * LPR + lane];
                out(r * ldc + kk * ldk, row_scaled(sum, rscale, r));
            }
        }
    }
}

// ------------------------------------------------------ helpers / scatter --
template <typename T>
__global__ void k_beta(int64_t rows, int64_t k, T* __restrict__ C, int64_t ldc, T beta,
                       int mode) {
    int64_t total = rows * k;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = t % rows, j = t / rows;
        T* p = C + i + j * ldc;
        *p = (mode == 0) ? T(0) : mul_scalar(*p, beta);
    }
}

// B (rows x cols, column-major, ld) -> BT row-major (ld = cols), in tiles of
// 64 rows x 32 columns: each wavefront issues its eight column reads (64
// consecutive rows each) before the first LDS write, then the tile leaves as
// contiguous row-major runs.
template <typename T>
__global__ __launch_bounds__(256) void k_to_rowmajor(int64_t rows, int64_t cols,
                                                     const T* __restrict__ B, int64_t ldb,
                                                     T* __restrict__ BT) {
    constexpr int RB = 64, CB = 32, Q = CB / 4;
    __shared__ T tile[CB][RB + 1];  // [column][row]
    const int64_t r0 = blockIdx.x * (int64_t)RB;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rn = (int)(rows - r0 < RB ? rows - r0 : RB);
    for (int64_t c0 = 0; c0 < cols; c0 += CB) {
        const int cw = (int)(cols - c0 < CB ? cols - c0 : CB);
        T v[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int cc = wv + 4 * q;
            v[q] = (cc < cw && lane < rn) ? B[(r0 + lane) + (c0 + cc) * ldb] : T(0);
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) tile[wv + 4 * q][lane] = v[q];
        __syncthreads();
        if (cw == CB && cols == CB) {  // the tile is one contiguous run of rn rows
            T* out = BT + r0 * cols;
#pragma unroll
            for (int q = 0; q < RB * CB / 256; ++q) {
                const int e = threadIdx.x + 256 * q;
                const int rr = e / CB, cc = e % CB;
                if (rr < rn) out[e] = tile[cc][rr];
            }
        } else {
            for (int e = threadIdx.x; e < rn * cw; e += 256) {
                const int rr = e / cw, cc = e % cw;
                BT[(r0 + rr) * cols + c0 + cc] = tile[cc][rr];
            }
        }
        __syncthreads();
    }
}

// CT (rows x cols, row-major, ld = cols) -> C (column-major, ldc) through store_out,
// in tiles of 64 rows x 32 columns: the tile arrives as contiguous row-major runs
// and leaves as 64-row column segments (512 contiguous bytes per wavefront for
// Float64), so the SpMM kernel can write its rows whole and C still receives
// whole lines.  The value in CT is the row's sum before α and β.
template <typename T>
__global__ __launch_bounds__(256) void k_from_rowmajor(int64_t rows, int64_t cols,
                                                       const T* __restrict__ CT,
                                                       T* __restrict__ C, int64_t ldc, T alpha,
                                                       T beta, int mode) {
    constexpr int RB = 64, CB = 32, Q = CB / 4;
    __shared__ T tile[CB][RB + 1];  // [column][row]
    const int64_t r0 = blockIdx.x * (int64_t)RB;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rn = (int)(rows - r0 < RB ? rows - r0 : RB);
    for (int64_t c0 = 0; c0 < cols; c0 += CB) {
        const int cw = (int)(cols - c0 < CB ? cols - c0 : CB);
        if (cw == CB && cols == CB) {  // the tile is one contiguous run of rn rows
            const T* in = CT + r0 * cols;
            T v[RB * CB / 256];
#pragma unroll
            for (int q = 0; q < RB * CB / 256; ++q) {
                const int e = threadIdx.x + 256 * q;
                v[q] = e / CB < rn ? in[e] : T(0);
            }
#pragma unroll
            for (int q = 0; q < RB * CB / 256; ++q) {
                const int e = threadIdx.x + 256 * q;
                tile[e % CB][e / CB] = v[q];
            }
        } else {
            for (int e = threadIdx.x; e < rn * cw; e += 256) {
                const int rr = e / cw, cc = e % cw;
                tile[cc][rr] = CT[(r0 + rr) * cols + c0 + cc];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int cc = wv + 4 * q;
            if (cc < cw && lane < rn)
                store_out(C, (r0 + lane) + (c0 + cc) * ldc, tile[cc][lane], alpha, beta, mode);
        }
        __syncthreads();
    }
}

template <typename T>
__device__ __forceinline__ void atomic_add(T* p, T v) {
    atomic_add_v(p, v);
}

// Scatter: for every stored (row, col, v) of the CSR arrays,
// C[col, kk] += v * (B[row, kk] * α)  (kernel_spmatmul_csr_T!).
// Scatter on CSR arrays (CSR-T / Adjoint, CSC-N): C[col, kk] += v * (B[row, kk] * α)
// with hardware float atomics, as the reference's @atomic kernels
// (matrix_csr_kernels.jl:24-45, src/matrix_csc/matrix_csc_kernels.jl:1-22).
// It runs on the SpMV windows: wavefront w streams the entries [576w, 576(w+1))
// with coalesced loads, and each lane finds its entry's row by a binary search
// over the window's row offsets (in LDS): the row of the entries before the
// first owned row is the one started in an earlier window.
template <typename T, typename I>
__global__ __launch_bounds__(64 * WAVES) void k_csr_scatter(
    int64_t m, int64_t n, int64_t k, int64_t nnz, const I* __restrict__ rowptr,
    const I* __restrict__ colval, const T* __restrict__ nzval, const T* __restrict__ B,
    int64_t ldb, int opB, T* __restrict__ C, int64_t ldc, T alpha, int cj,
    const int64_t* __restrict__ win_row, int64_t nwin) {
    __shared__ int s_off[WAVES][WRP];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int64_t w = (int64_t)blockIdx.x * WAVES + wv;
    if (w >= nwin) return;
    const int64_t E0 = w * SPMV_WIN < nnz ? w * SPMV_WIN : nnz;
    const int64_t E1 = E0 + SPMV_WIN < nnz ? E0 + SPMV_WIN : nnz;
    const int cnt = (int)(E1 - E0);
    if (cnt <= 0) return;
    const int64_t r0 = clamp64(win_row[w], 0, m);
    const int64_t r1 = clamp64(win_row[w + 1], r0, m);
    const int64_t nrows = r1 - r0;
    const bool rp_in_lds = nrows + 1 <= WRP;
    int* off = s_off[wv];
    if (rp_in_lds) {
        for (int i = lane; i <= (int)nrows; i += 64)
            off[i] = (int)(clamp64((int64_t)rowptr[r0 + i] - 1, E0, E1) - E0);
    }
    wave_lds_sync();
    // zero-based row of window entry i (E0 + i)
    auto row_of = [&](int i) -> int64_t {
        int64_t lo = 0, hi = nrows;  // largest t in [0, nrows] with start(r0 + t) <= E0 + i
        const int64_t first = rp_in_lds ? off[0]
                                        : clamp64((int64_t)rowptr[r0] - 1, E0, E1) - E0;
        if (nrows == 0 || i < first) return r0 - 1;  // started in an earlier window
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            const int64_t st = rp_in_lds ? off[mid]
                                         : clamp64((int64_t)rowptr[r0 + mid] - 1, E0, E1) - E0;
            if (st <= i) lo = mid;
            else hi = mid - 1;
        }
        return r0 + lo;
    };
#pragma unroll
    for (int u = 0; u < WU; ++u) {
        const int i = u * 64 + lane;
        if (i >= cnt) continue;
        const int64_t j = E0 + i;
        const int64_t c = clamp64((int64_t)ldnt(colval + j) - 1, 0, n - 1);
        const T v = conj_if(ldnt(nzval + j), cj & 1);
        const int64_t r = clamp64(row_of(i), 0, m - 1);
        for (int64_t kk = 0; kk < k; ++kk) {
            const T xb = conj_if(opB ? B[kk + r * ldb] : B[r + kk * ldb], (cj & 2) != 0);
            atomic_add(&C[c + kk * ldc], v * mul_scalar(xb, alpha));
        }
    }
}

// COO scatter: C[r, kk] += v * (B[c, kk] * α); trans swaps the roles of r and c.
// Consecutive lanes adding into the same output (a row-major COO gives runs of
// equal rows) are first summed across the wavefront by a segmented scan, so
// one atomic per run reaches memory instead of one per entry.
template <typename T, typename I>
__global__ void k_coo_scatter(int64_t rows_out, int64_t cols_in, int64_t k, int64_t nnz,
                              const I* __restrict__ rowind, const I* __restrict__ colind,
                              const T* __restrict__ nzval, int trans,
                              const T* __restrict__ B, int64_t ldb, int opB,
                              T* __restrict__ C, int64_t ldc, T alpha, int cj) {
    const int lane = threadIdx.x & 63;
    const int64_t total = nnz * k;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // whole wavefronts iterate together (the scan needs every lane)
    const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63);
    for (int64_t tb = t0; tb < total; tb += stride) {
        const int64_t t = tb + lane;
        const bool valid = t < total;
        int64_t key = -1 - lane;  // distinct for invalid lanes: never joins a run
        T val = T(0);
        if (valid) {
            const int64_t i = t % nnz, kk = t / nnz;
            int64_t r = (int64_t)(trans ? colind[i] : rowind[i]) - 1;
            int64_t c = (int64_t)(trans ? rowind[i] : colind[i]) - 1;
            r = clamp64(r, 0, rows_out - 1);
            c = clamp64(c, 0, cols_in - 1);
            const T xb = conj_if(opB ? B[kk + c * ldb] : B[c + kk * ldb], (cj & 2) != 0);
            val = conj_if(nzval[i], cj & 1) * mul_scalar(xb, alpha);
            key = r + kk * ldc;
        }
        // segmented inclusive scan over runs of equal keys (head = key differs
        // from the previous lane's)
        const int64_t kprev = __shfl_up(key, 1, 64);
        int head = (lane == 0 || kprev != key) ? 1 : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const T vu = shfl_up_v(val, o);
            const int hu = __shfl_up(head, o, 64);
            if (lane >= o && !head) {
                val = val + vu;
                head = hu;
            }
        }
        const int64_t knext = __shfl_down(key, 1, 64);
        const bool tail = lane == 63 || knext != key;
        if (valid && tail) atomic_add(&C[key], val);
    }
}

// ---------------------------------------------------------------- drivers --
template <typename T>
static int apply_beta(hipStream_t s, int64_t rows, int64_t k, T* C, int64_t ldc, T beta) {
    int mode = beta_mode_of(beta);
    if (mode == 1 || rows == 0 || k == 0) return GSA_OK;
    hipLaunchKernelGGL(k_beta<T>, dim3(grid_for(rows * k, 256)), dim3(256), 0, s, rows, k, C,
                       ldc, beta, mode);
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

// Row-block plan for one call: use the caller's plan when it matches, else build
// a temporary one in scratch (device work only, no synchronisation).
struct PlanUse {
    const int64_t* blk_row = nullptr;
    const int64_t* blk_pos = nullptr;
    int64_t nblocks = 0;
    const int64_t* win_row = nullptr;
    int64_t nwin = 0;
    const int64_t* order = nullptr;  // window launch order, when built for these columns
    int64_t nslot = 0;
    const void* order_col = nullptr;
    const uint32_t* col32 = nullptr;  // 32-bit copy of those columns (Int64 indices)
    CsrPlanData* owner = nullptr;
    Scratch tmp, tmp_part;
    explicit PlanUse(hipStream_t s) : tmp(s), tmp_part(s) {}
    // wave = true: the SpMV windows; else the workgroup row blocks.
    int get(hipStream_t s, gsa_plan plan, int itype, int64_t m, int64_t nnz, const void* ptr,
            bool wave) {
        if (plan && plan->kind == 0 && plan->csr.m == m && plan->csr.nnz == nnz &&
            plan->csr.itype == itype && plan->csr.ptr == ptr) {
            owner = &plan->csr;
            blk_row = owner->blk_row;
            blk_pos = owner->blk_pos;
            nblocks = owner->nblocks;
            win_row = owner->win_row;
            nwin = owner->nwin;
            order = owner->order;
            nslot = owner->nslot;
            order_col = owner->order_col;
            col32 = owner->col32;
            return GSA_OK;
        }
        if (wave) {
            const int64_t cnt = nnz / SPMV_WIN + 1;
            GSA_TRY(tmp.alloc(sizeof(int64_t) * (cnt + 1)));
            GSA_TRY(build_csr_plan(s, itype, m, nnz, ptr, tmp.as<int64_t>(), cnt, nullptr,
                                   SPMV_WIN));
            win_row = tmp.as<int64_t>();
            nwin = cnt;
            return GSA_OK;
        }
        const int64_t cnt = nnz / SPMV_NB + 1;
        GSA_TRY(tmp.alloc(sizeof(int64_t) * ((cnt + 1) + PLAN_POS * cnt)));
        int64_t* br = tmp.as<int64_t>();
        int64_t* bp = br + (cnt + 1);
        GSA_TRY(build_csr_plan(s, itype, m, nnz, ptr, br, cnt, bp, SPMV_NB, SPMV_CAP));
        blk_row = br; blk_pos = bp; nblocks = cnt;
        return GSA_OK;
    }
    // Head/tail partial sums of the SpMV windows: scratch of this call, taken
    // on its stream, so a plan is never written by a product and one plan
    // serves any number of concurrent calls on any streams.
    int part_buffer(void** out) {
        GSA_TRY(tmp_part.alloc(sizeof(double) * 2 * nwin));
        *out = tmp_part.p;
        return GSA_OK;
    }
};

// Gather-shaped product on CSR arrays: C (rows m) = β C + α A B, B has n rows.
// With `perm` (SpMV through a row view), value j is nzval[perm[j]].
template <typename T, typename I, typename V = T, typename PI = void>
static int gather_csr(hipStream_t s, int64_t m, int64_t n, int64_t k, const I* rowptr,
                      const I* colval, const V* nzval, int64_t nnz, const T* B, int64_t ldb,
                      int opB, T* C, int64_t ldc, T alpha, T beta, gsa_plan plan,
                      int itype, const float* rscale = nullptr, const PI* perm = nullptr,
                      bool c_rowmajor = false, int cj = 0) {
    if (m == 0 || k == 0) return GSA_OK;
    if (n == 0 || nnz == 0) {  // A*B contributes nothing: only β applies
        return c_rowmajor ? apply_beta<T>(s, k, m, C, ldc, beta)
                          : apply_beta<T>(s, m, k, C, ldc, beta);
    }
    GSA_CHECK(!c_rowmajor || k > 1 || ldc == 1, GSA_ERR_INVALID,
              "a row-major C with one column needs ldc = 1");
    if constexpr (!std::is_void_v<PI>)  // values are read through perm: it must exist
        GSA_CHECK(perm != nullptr || nnz == 0, GSA_ERR_INVALID, "value permutation is null");
    PlanUse pu(s);
    GSA_TRY(pu.get(s, plan, itype, m, nnz, rowptr, k == 1));
    const int mode = beta_mode_of(beta);
    const unsigned grid = (unsigned)pu.nblocks;
    if (k == 1) {
        const T* x = B;
        Scratch xs(s);
        if (opB && ldb != 1) {  // strided vector: compact it first
            GSA_TRY(xs.alloc(sizeof(T) * n));
            GSA_HIP(hipMemcpy2DAsync(xs.p, sizeof(T), B, sizeof(T) * ldb, sizeof(T), n,
                                     hipMemcpyDeviceToDevice, s));
            x = xs.as<T>();
        }
        void* buf = nullptr;
        GSA_TRY(pu.part_buffer(&buf));
        T* head = static_cast<T*>(buf);
        T* tail = head + pu.nwin;
        // the XCD launch order only for the column array it was built from
        const int64_t* order = (pu.order && pu.order_col == (const void*)colval) ? pu.order : nullptr;
        const int64_t grid1 = order ? pu.nslot / WAVES : cdiv(pu.nwin, WAVES);
        GSA_CHECK(grid1 <= (int64_t)UINT32_MAX, GSA_ERR_INVALID, "SpMV grid too large");
        // the plan's 32-bit copy of these Int64 columns: 4 bytes less per entry
        const uint32_t* c32 = (order && sizeof(I) == 8) ? pu.col32 : nullptr;
        if (c32)
            hipLaunchKernelGGL((k_csr_spmv<T, I, V, PI, uint32_t>), dim3((unsigned)grid1),
                               dim3(64 * WAVES), 0, s, m, n, nnz, rowptr, c32, nzval, perm,
                               rscale, x, C, alpha, beta, mode, cj, pu.win_row, pu.nwin, order,
                               head, tail);
        else
            hipLaunchKernelGGL((k_csr_spmv<T, I, V, PI>), dim3((unsigned)grid1), dim3(64 * WAVES),
                               0, s, m, n, nnz, rowptr, colval, nzval, perm, rscale, x, C, alpha,
                               beta, mode, cj, pu.win_row, pu.nwin, order, head, tail);
        hipLaunchKernelGGL((k_spmv_fold<T, I>), dim3((unsigned)cdiv(pu.nwin, 256)), dim3(256), 0, s,
                           m, nnz, rowptr, pu.win_row, pu.nwin, (const T*)head, (const T*)tail,
                           rscale, C, alpha, beta, mode);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    }
    const T* BT = B;
    int64_t ldbt = ldb;
    Scratch bt(s);
    if (!opB) {
        GSA_TRY(bt.alloc(sizeof(T) * n * k));
        hipLaunchKernelGGL(k_to_rowmajor<T>, dim3((unsigned)cdiv(n, 64)), dim3(256), 0, s, n, k,
                           B, ldb, bt.as<T>());
        GSA_HIP(hipGetLastError());
        BT = bt.as<T>();
        ldbt = k;
    }
    // a column-major C of 2^21 outputs or more is written through row-major scratch:
    // the kernel then stores each row's k sums as one contiguous segment, and
    // k_from_rowmajor turns the tiles into whole column segments of C (one 8-byte
    // store per (row, column) into k different lines costs more than the extra pass)
    const bool via_rows = !c_rowmajor && k > 1 && m * k >= SPMM_VIA_ROWS_MIN;
    Scratch ct(s);
    T* Cw = C;
    int kmode = mode;
    int64_t c_ld = c_rowmajor ? ldc : 1, c_lk = c_rowmajor ? 1 : ldc;
    if (via_rows) {
        GSA_TRY(ct.alloc(sizeof(T) * m * k));
        Cw = ct.as<T>();
        c_ld = k;
        c_lk = 1;
        kmode = -1;
    }
#define GSA_SPMM_LAUNCH(L)                                                                  \
    do {                                                                                    \
        if (n <= INT32_MAX)                                                                 \
            hipLaunchKernelGGL((k_csr_spmm<T, I, int32_t, L, V, PI>), dim3(grid), dim3(WG), 0, s, m, \
                               n, k, nnz, rowptr, colval, nzval, perm, rscale, BT, ldbt, Cw, \
                               c_ld, c_lk, alpha, beta, kmode, cj, pu.blk_row, pu.blk_pos);  \
        else                                                                                \
            hipLaunchKernelGGL((k_csr_spmm<T, I, I, L, V, PI>), dim3(grid), dim3(WG), 0, s, m, n, k, \
                               nnz, rowptr, colval, nzval, perm, rscale, BT, ldbt, Cw, c_ld, \
                               c_lk, alpha, beta, kmode, cj, pu.blk_row, pu.blk_pos);        \
    } while (0)
    if (k <= 2) GSA_SPMM_LAUNCH(2);
    else if (k <= 4) GSA_SPMM_LAUNCH(4);
    else if (k <= 8) GSA_SPMM_LAUNCH(8);
    else if (k <= 16) GSA_SPMM_LAUNCH(16);
    else if (k <= 32) GSA_SPMM_LAUNCH(32);
    else GSA_SPMM_LAUNCH(64);
#undef GSA_SPMM_LAUNCH
    GSA_HIP(hipGetLastError());
    if (via_rows) {
        hipLaunchKernelGGL(k_from_rowmajor<T>, dim3((unsigned)cdiv(m, 64)), dim3(256), 0, s, m, k,
                           (const T*)Cw, C, ldc, alpha, beta, mode);
        GSA_HIP(hipGetLastError());
    }
    return GSA_OK;
}

// Scatter-shaped product on CSR arrays (m rows stored): C (n rows) += α Aᵀ B.
template <typename T, typename I>
static int scatter_csr(hipStream_t s, int64_t m, int64_t n, int64_t k, const I* rowptr,
                       const I* colval, const T* nzval, int64_t nnz, const T* B, int64_t ldb,
                       int opB, T* C, int64_t ldc, T alpha, T beta, gsa_plan plan,
                       int itype, int cj = 0) {
    if (n == 0 || k == 0) return GSA_OK;
    GSA_TRY(apply_beta<T>(s, n, k, C, ldc, beta));
    if (m == 0 || nnz == 0) return GSA_OK;
    PlanUse pu(s);
    GSA_TRY(pu.get(s, plan, itype, m, nnz, rowptr, true));
    hipLaunchKernelGGL((k_csr_scatter<T, I>), dim3((unsigned)cdiv(pu.nwin, WAVES)), dim3(64 * WAVES),
                       0, s, m, n, k, nnz, rowptr, colval, nzval, B, ldb, opB, C, ldc, alpha,
                       cj, pu.win_row, pu.nwin);
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

// SpMV y = β y + α A x through a row view (m rows, x of length n): the CSR
// window kernel over the view's row pointer, columns from the view (or the
// caller's `colind` when the storage is row-ordered), values through the
// view's storage positions.
template <typename T, typename I>
static int view_spmv(hipStream_t s, const RowView& v, int64_t n, const I* colind, const T* nzval,
                     const T* B, int64_t ldb, int opB, T* C, int64_t ldc, T alpha,
                     T beta, int itype, int64_t k = 1, int cj = 0) {
    const I* col = v.col ? (const I*)v.col : colind;
    if (!v.perm)
        return gather_csr<T, I>(s, v.m, n, k, (const I*)v.rowptr, col, nzval, v.nnz, B, ldb, opB,
                                C, ldc, alpha, beta, v.win, itype, nullptr, (const void*)nullptr,
                                false, cj);
    if (v.perm64)
        return gather_csr<T, I, T, uint64_t>(s, v.m, n, k, (const I*)v.rowptr, col, nzval, v.nnz,
                                             B, ldb, opB, C, ldc, alpha, beta, v.win, itype,
                                             nullptr, (const uint64_t*)v.perm, false, cj);
    return gather_csr<T, I, T, uint32_t>(s, v.m, n, k, (const I*)v.rowptr, col, nzval, v.nnz, B,
                                         ldb, opB, C, ldc, alpha, beta, v.win, itype, nullptr,
                                         (const uint32_t*)v.perm, false, cj);
}

// A plan's gather view (the output rows of a scatter-shaped product, with each
// entry's B row and storage position) when it was built on these arrays.
static const RowView* gview_for(const RowView* v, int itype, int64_t rows, int64_t nnz,
                                const void* a, const void* b) {
    if (!v || v->itype != itype || v->m != rows || v->nnz != nnz || v->src[0] != a ||
        v->src[1] != b)
        return nullptr;
    return v;
}

// The plan's row view when it was built on exactly these arrays.
static const RowView* view_for(gsa_plan plan, int itype, int64_t m, int64_t nnz, const void* a,
                               const void* b) {
    if (!plan || !plan->view) return nullptr;
    const RowView* v = plan->view;
    if (v->itype != itype || v->m != m || v->nnz != nnz || v->src[0] != a || v->src[1] != b)
        return nullptr;
    return v;
}

static int check_common(int opA, int opB, int64_t m, int64_t n, int64_t k, int64_t nnz,
                        int64_t ldb, int64_t ldc) {
    GSA_CHECK(opA >= GSA_OP_N && opA <= GSA_OP_C, GSA_ERR_INVALID, "invalid opA %d", opA);
    GSA_CHECK(opB >= GSA_OP_N && opB <= GSA_OP_C, GSA_ERR_INVALID, "invalid opB %d", opB);
    GSA_CHECK(m >= 0 && n >= 0 && k >= 0 && nnz >= 0, GSA_ERR_INVALID,
              "negative dimension (m=%lld n=%lld k=%lld nnz=%lld)", (long long)m, (long long)n,
              (long long)k, (long long)nnz);
    const int64_t rows_b = (opA == GSA_OP_N) ? n : m;
    const int64_t rows_c = (opA == GSA_OP_N) ? m : n;
    if (opB == GSA_OP_N) {
        GSA_CHECK(ldb >= (rows_b > 1 ? rows_b : 1) || k <= 1, GSA_ERR_INVALID,
                  "ldb %lld < rows of B %lld", (long long)ldb, (long long)rows_b);
    } else {
        GSA_CHECK(ldb >= (k > 1 ? k : 1), GSA_ERR_INVALID, "ldb %lld < k %lld for transposed B",
                  (long long)ldb, (long long)k);
    }
    GSA_CHECK(ldc >= (rows_c > 1 ? rows_c : 1) || k <= 1, GSA_ERR_INVALID,
              "ldc %lld < rows of C %lld", (long long)ldc, (long long)rows_c);
    return GSA_OK;
}

// ---- allocation-free SpMV ----------------------------------------------------
// Device address of `scratch` on each GPU, looked up once per device (with that
// device current) on its first allocation-free call.
constexpr int NOALLOC_DEVICES = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[NOALLOC_DEVICES] = {};

static int noalloc_scratch(uint8_t** out) {
    int dev = 0;
    GSA_HIP(hipGetDevice(&dev));
    GSA_CHECK(dev >= 0 && dev < NOALLOC_DEVICES, GSA_ERR_INVALID,
              "device ordinal %d beyond the %d the allocation-free path serves", dev,
              NOALLOC_DEVICES);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        GSA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        GSA_CHECK(p != nullptr, GSA_ERR_HIP, "no device address for the allocation-free scratch");
        g_noalloc_ptr[dev] = p;
    }
    *out = static_cast<uint8_t*>(g_noalloc_ptr[dev]);
    return GSA_OK;
}

// Bytes of `scratch` one SpMV of nnz entries uses: the window table (nwin + 1
// row indices), then head and tail partial sums (nwin values each).
static int64_t noalloc_bytes(int64_t nnz, size_t value_bytes, int64_t* head_off) {
    const int64_t nwin = nnz / SPMV_WIN + 1;
    const int64_t table = (sizeof(int64_t) * (nwin + 1) + 255) / 256 * 256;
    *head_off = table;
    return table + 2 * (int64_t)value_bytes * nwin;
}

// The k = 1, no-plan path of gather_csr (window table built for the call, the
// window kernel, the fold), with every buffer taken from `ws`.
template <typename T, typename I>
static int spmv_noalloc(hipStream_t s, int itype, int64_t m, int64_t n, const I* rowptr,
                        const I* colval, const T* nzval, int64_t nnz, const T* x, T* y,
                        T alpha, T beta, uint8_t* ws) {
    if (m == 0) return GSA_OK;
    if (n == 0 || nnz == 0) return apply_beta<T>(s, m, 1, y, m, beta);
    int64_t head_off = 0;
    const int64_t need = noalloc_bytes(nnz, sizeof(T), &head_off);
    GSA_CHECK(need <= (int64_t)GSA_NOALLOC_SCRATCH_BYTES, GSA_ERR_INVALID,
              "gsa_spmv_csr_noalloc: %lld bytes of scratch needed, %d available",
              (long long)need, (int)GSA_NOALLOC_SCRATCH_BYTES);
    const int64_t nwin = nnz / SPMV_WIN + 1;
    int64_t* win_row = reinterpret_cast<int64_t*>(ws);
    T* head = reinterpret_cast<T*>(ws + head_off);
    T* tail = head + nwin;
    // the table is cleared and filled here; every head and tail the fold reads is
    // written by the window kernel of this call first (see window_tail)
    GSA_TRY(build_csr_plan(s, itype, m, nnz, rowptr, win_row, nwin, nullptr, SPMV_WIN));
    const int mode = beta_mode_of(beta);
    const int64_t grid = cdiv(nwin, WAVES);
    GSA_CHECK(grid <= (int64_t)UINT32_MAX, GSA_ERR_INVALID, "SpMV grid too large");
    hipLaunchKernelGGL((k_csr_spmv<T, I>), dim3((unsigned)grid), dim3(64 * WAVES), 0, s, m, n, nnz,
                       rowptr, colval, nzval, (const void*)nullptr, (const float*)nullptr, x, y,
                       alpha, beta, mode, 0, (const int64_t*)win_row, nwin,
                       (const int64_t*)nullptr, head, tail);
    hipLaunchKernelGGL((k_spmv_fold<T, I>), dim3((unsigned)cdiv(nwin, 256)), dim3(256), 0, s, m,
                       nnz, rowptr, (const int64_t*)win_row, nwin, (const T*)head,
                       (const T*)tail, (const float*)nullptr, y, alpha, beta, mode);
    GSA_HIP(hipGetLastError());
    return GSA_OK;
}

static int plan_release(CsrPlanData& d) {
    hipError_t e = hipSuccess, f;
    for (void* q : {(void*)d.blk_row, (void*)d.order, (void*)d.col32})
        if (q && (f = hipFree(q)) != hipSuccess && e == hipSuccess) e = f;
    d.blk_row = d.blk_pos = d.win_row = nullptr;
    d.col32 = nullptr;
    d.order = nullptr;
    d.nslot = 0;
    d.order_col = nullptr;
    if (e != hipSuccess) return hip_fail(e, "hipFree(plan)");
    return GSA_OK;
}

}  // namespace gsa

using namespace gsa;

extern "C" {

int gsa_plan_create(void* stream, int itype, int64_t m, int64_t nnz, const void* ptr,
                    gsa_plan* plan) {
    GSA_TRY(ensure_init());
    GSA_CHECK(plan && m >= 0 && nnz >= 0 && ptr, GSA_ERR_INVALID,
              "gsa_plan_create: bad arguments (the row pointer has m+1 >= 1 entries)");
    GSA_CHECK(itype == GSA_I32 || itype == GSA_I64, GSA_ERR_UNSUPPORTED, "bad itype %d", itype);
    hipStream_t s = as_stream(stream);
    auto* p = new gsa_plan_s();
    p->kind = 0;
    p->csr.itype = itype;
    p->csr.m = m;
    p->csr.nnz = nnz;
    p->csr.ptr = ptr;
    p->csr.nblocks = nnz / SPMV_NB + 1;
    p->csr.nwin = nnz / SPMV_WIN + 1;
    const int64_t words = (p->csr.nblocks + 1) + PLAN_POS * p->csr.nblocks + (p->csr.nwin + 1);
    hipError_t e = hipMalloc(&p->csr.blk_row, sizeof(int64_t) * words);
    if (e != hipSuccess) {
        (void)plan_release(p->csr);
        delete p;
        return hip_fail(e, "hipMalloc(plan)");
    }
    p->csr.blk_pos = p->csr.blk_row + (p->csr.nblocks + 1);
    p->csr.win_row = p->csr.blk_pos + PLAN_POS * p->csr.nblocks;
    int r = build_csr_plan(s, itype, m, nnz, ptr, p->csr.blk_row, p->csr.nblocks, p->csr.blk_pos);
    if (r == GSA_OK)
        r = build_csr_plan(s, itype, m, nnz, ptr, p->csr.win_row, p->csr.nwin, nullptr, SPMV_WIN);
    if (r != GSA_OK) { (void)plan_release(p->csr); delete p; return r; }
    *plan = p;
    return GSA_OK;
}

// Slots of the windows (host side, plan time): bucket b < SPMV_XCDS puts a
// window on XCD b, the others top every XCD's list up to an even share; XCD
// x's j-th window goes to slot 4*(8*(j/4) + x) + j%4 (workgroup 8*(j/4) + x,
// which runs on XCD x).  Unused slots hold -1.
static std::vector<int64_t> window_slots(const std::vector<int>& bucket) {
    const int64_t nwin = (int64_t)bucket.size();
    int64_t cnt[SPMV_XCDS + 1] = {0};
    for (int b : bucket) cnt[(b >= 0 && b < SPMV_XCDS) ? b : SPMV_XCDS]++;
    const int64_t share = (nwin + SPMV_XCDS - 1) / SPMV_XCDS;
    int64_t extra[SPMV_XCDS], len[SPMV_XCDS], most = 0, left = cnt[SPMV_XCDS];
    for (int x = 0; x < SPMV_XCDS; ++x) {
        extra[x] = std::min<int64_t>(cnt[x] < share ? share - cnt[x] : 0, left);
        left -= extra[x];
    }
    for (int x = 0; left > 0; x = (x + 1) % SPMV_XCDS) { ++extra[x]; --left; }
    for (int x = 0; x < SPMV_XCDS; ++x) {
        len[x] = cnt[x] + extra[x];
        most = std::max(most, len[x]);
    }
    const int64_t rounds = (most + WAVES - 1) / WAVES;
    std::vector<int64_t> order((size_t)(rounds * SPMV_XCDS * WAVES), -1);
    int64_t pos[SPMV_XCDS] = {0};
    int xs = 0;  // next XCD taking shared windows
    for (int64_t w = 0; w < nwin; ++w) {
        int x = bucket[w];
        if (x < 0 || x >= SPMV_XCDS) {
            while (extra[xs] == 0) ++xs;  // the extras sum to the shared count
            --extra[xs];
            x = xs;
        }
        const int64_t j = pos[x]++;
        order[(size_t)(((j / WAVES) * SPMV_XCDS + x) * WAVES + (j % WAVES))] = w;
    }
    return order;
}

// The XCD launch order of a plan's SpMV windows for column array `colval`
// (n columns).  Synchronises once.
__global__ void k_cols_to_u32(int64_t nnz, const int64_t* __restrict__ col,
                              uint32_t* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)col[i];
}

static int build_order(hipStream_t s, CsrPlanData& d, const void* colval, int64_t n) {
    if (d.order || d.col32) {
        GSA_HIP(hipStreamSynchronize(s));
        if (d.order) GSA_HIP(hipFree(d.order));
        if (d.col32) GSA_HIP(hipFree(d.col32));
        d.order = nullptr;
        d.col32 = nullptr;
        d.nslot = 0;
        d.order_col = nullptr;
    }
    if (d.nwin == 0 || d.nnz == 0 || n <= 0) return GSA_OK;
    Scratch bucket(s);
    GSA_TRY(bucket.alloc(sizeof(int) * d.nwin));
    GSA_TRY(dispatch_i(d.itype, [&](auto ti) {
        using I = decltype(ti);
        hipLaunchKernelGGL(k_win_bucket<I>, dim3((unsigned)cdiv(d.nwin, 256)), dim3(256), 0, s, d.m, n,
                           d.nnz, (const I*)d.ptr, (const I*)colval, (const int64_t*)d.win_row,
                           d.nwin, bucket.as<int>());
        GSA_HIP(hipGetLastError());
        return (int)GSA_OK;
    }));
    std::vector<int> hb((size_t)d.nwin);
    GSA_HIP(hipMemcpyAsync(hb.data(), bucket.p, sizeof(int) * d.nwin, hipMemcpyDeviceToHost, s));
    GSA_HIP(hipStreamSynchronize(s));
    const std::vector<int64_t> order = window_slots(hb);
    if (d.itype == GSA_I64 && n <= (int64_t)UINT32_MAX) {
        // structure only, like the order: a 32-bit copy of the one-based columns
        GSA_HIP(hipMalloc(&d.col32, sizeof(uint32_t) * d.nnz));
        hipLaunchKernelGGL(k_cols_to_u32, dim3(grid_for(d.nnz, 256)), dim3(256), 0, s, d.nnz,
                           (const int64_t*)colval, d.col32);
        GSA_HIP(hipGetLastError());
    }
    GSA_HIP(hipMalloc(&d.order, sizeof(int64_t) * order.size()));
    GSA_HIP(hipMemcpyAsync(d.order, order.data(), sizeof(int64_t) * order.size(),
                           hipMemcpyHostToDevice, s));
    GSA_HIP(hipStreamSynchronize(s));  // before the host vector goes
    d.nslot = (int64_t)order.size();
    d.order_col = colval;
    return GSA_OK;
}

int gsa_plan_set_columns(void* stream, gsa_plan plan, const void* colval, int64_t n) {
    GSA_TRY(ensure_init());
    GSA_CHECK(plan && plan->kind == 0 && n >= 0 && (colval || plan->csr.nnz == 0), GSA_ERR_INVALID,
              "gsa_plan_set_columns: bad arguments");
    return build_order(as_stream(stream), plan->csr, colval, n);
}

int gsa_plan_create_coo(void* stream, int itype, int64_t m, int64_t n, int64_t nnz,
                        const void* rowind, const void* colind, gsa_plan* plan) {
    GSA_TRY(ensure_init());
    GSA_CHECK(plan && m >= 0 && n >= 0 && nnz >= 0 && (nnz == 0 || (rowind && colind)),
              GSA_ERR_INVALID, "gsa_plan_create_coo: bad arguments");
    GSA_CHECK(itype == GSA_I32 || itype == GSA_I64, GSA_ERR_UNSUPPORTED, "bad itype %d", itype);
    hipStream_t s = as_stream(stream);
    auto* p = new gsa_plan_s();
    p->kind = 2;
    // entries not in row order: row bins (two streaming passes per SpMV)
    int sorted = 1;
    int r0 = indices_sorted(s, itype, nnz, rowind, &sorted);
    if (r0 == GSA_OK && !sorted) r0 = build_bin_plan(s, itype, m, n, nnz, rowind, nullptr, &p->bins);
    if (r0 != GSA_OK) { delete p; return r0; }
    if (p->bins) {
        bin_plan_set_sources(p->bins, rowind, colind);
        *plan = p;
        return GSA_OK;
    }
    p->view = new RowView();
    int r = build_row_view(s, itype, m, nnz, rowind, colind, false, *p->view);
    if (r == GSA_OK) r = gsa_plan_create(stream, itype, m, nnz, p->view->rowptr, &p->view->win);
    if (r == GSA_OK)
        r = build_order(s, p->view->win->csr, p->view->col ? p->view->col : colind, n);
    if (r != GSA_OK) { release_row_view(p->view); delete p; return r; }
    p->view->src[0] = rowind;
    p->view->src[1] = colind;
    *plan = p;
    return GSA_OK;
}

int gsa_plan_create_csc(void* stream, int itype, int64_t m, int64_t n, int64_t nnz,
                        const void* colptr, const void* rowval, gsa_plan* plan) {
    GSA_CHECK(plan && m >= 0 && n >= 0 && nnz >= 0 && colptr && (nnz == 0 || rowval),
              GSA_ERR_INVALID, "gsa_plan_create_csc: bad arguments");
    GSA_TRY(gsa_plan_create(stream, itype, n, nnz, colptr, plan));  // column blocks (Aᵀ x)
    hipStream_t s = as_stream(stream);
    // A x: row bins when m allows them, else the row view
    int rb = build_bin_plan(s, itype, m, n, nnz, rowval, colptr, &(*plan)->bins);
    if (rb != GSA_OK) { gsa_plan_destroy(*plan); *plan = nullptr; return rb; }
    if ((*plan)->bins) {
        bin_plan_set_sources((*plan)->bins, colptr, rowval);
        return GSA_OK;
    }
    int r = dispatch_i(itype, [&](auto ti) -> int {
        using I = decltype(ti);
        Scratch cols(s);  // the column of every stored entry
        if (nnz > 0) {
            GSA_TRY(cols.alloc(sizeof(I) * nnz));
            GSA_TRY(expand_rows(s, itype, n, nnz, colptr, cols.p));
        }
        (*plan)->view = new RowView();
        GSA_TRY(build_row_view(s, itype, m, nnz, rowval, cols.p, true, *(*plan)->view));
        GSA_TRY(gsa_plan_create(stream, itype, m, nnz, (*plan)->view->rowptr, &(*plan)->view->win));
        GSA_TRY(build_order(s, (*plan)->view->win->csr, (*plan)->view->col, n));
        GSA_HIP(hipStreamSynchronize(s));  // the column scratch is released on return
        return GSA_OK;
    });
    if (r != GSA_OK) { gsa_plan_destroy(*plan); *plan = nullptr; return r; }
    (*plan)->view->src[0] = colptr;
    (*plan)->view->src[1] = rowval;
    return GSA_OK;
}

int gsa_plan_add_transpose(void* stream, gsa_plan plan, int format, int itype, int64_t m,
                           int64_t n, int64_t nnz, const void* a, const void* b) {
    GSA_TRY(ensure_init());
    GSA_CHECK(plan && (format == GSA_FMT_CSR || format == GSA_FMT_COO) && m >= 0 && n >= 0 &&
                  nnz >= 0 && a && (nnz == 0 || b),
              GSA_ERR_INVALID, "gsa_plan_add_transpose: bad arguments");
    if (bin_plan_matches(plan->bins_t, itype, n, m, nnz, a, b)) return GSA_OK;
    if (plan->bins_t) {
        GSA_HIP(hipDeviceSynchronize());  // queued products may still read the old bins
        release_bin_plan(plan->bins_t);
        plan->bins_t = nullptr;
    }
    hipStream_t s = as_stream(stream);
    // Aᵀ x scatters into the n columns: bins over the column indices; the sources
    // are the rows (CSR: from the row pointer; COO: rowind, given at each call)
    GSA_TRY(build_bin_plan(s, itype, n, m, nnz, b, format == GSA_FMT_CSR ? a : nullptr,
                           &plan->bins_t));
    if (plan->bins_t) bin_plan_set_sources(plan->bins_t, a, b);
    return GSA_OK;
}

int gsa_plan_add_gather_view(void* stream, gsa_plan plan, int format, int trans, int itype,
                             int64_t m, int64_t n, int64_t nnz, const void* a, const void* b) {
    GSA_TRY(ensure_init());
    GSA_CHECK(plan && (format == GSA_FMT_CSR || format == GSA_FMT_CSC || format == GSA_FMT_COO) &&
                  m >= 0 && n >= 0 && nnz >= 0 && a && (nnz == 0 || b),
              GSA_ERR_INVALID, "gsa_plan_add_gather_view: bad arguments");
    // CSR serves transpose(A)*B, CSC serves A*B; COO either
    GSA_CHECK(format == GSA_FMT_COO || (format == GSA_FMT_CSR) == (trans != 0), GSA_ERR_INVALID,
              "gsa_plan_add_gather_view: a CSR view is for transpose(A)*B, a CSC one for A*B");
    const int64_t rows = trans ? n : m;  // output rows
    RowView*& slot = trans ? plan->gview_t : plan->gview;
    if (gview_for(slot, itype, rows, nnz, a, b)) return GSA_OK;
    if (slot) {
        GSA_HIP(hipDeviceSynchronize());  // queued products may still read the old view
        release_row_view(slot);
        slot = nullptr;
    }
    hipStream_t s = as_stream(stream);
    return dispatch_i(itype, [&](auto ti) -> int {
        using I = decltype(ti);
        // each entry's output row and B row
        const void* out_row = nullptr;
        const void* b_row = nullptr;
        Scratch expanded(s);
        if (format == GSA_FMT_COO) {
            out_row = trans ? b : a;
            b_row = trans ? a : b;
        } else {  // the compressed index is the output row; the other comes from the pointer
            if (nnz > 0) {
                GSA_TRY(expanded.alloc(sizeof(I) * nnz));
                GSA_TRY(expand_rows(s, itype, format == GSA_FMT_CSR ? m : n, nnz, a, expanded.p));
            }
            out_row = b;
            b_row = expanded.p;
        }
        auto* v = new RowView();
        int r = build_row_view(s, itype, rows, nnz, out_row, b_row, true, *v);
        if (r == GSA_OK) r = gsa_plan_create(stream, itype, rows, nnz, v->rowptr, &v->win);
        if (r != GSA_OK) {
            release_row_view(v);
            return r;
        }
        v->src[0] = a;
        v->src[1] = b;
        slot = v;
        GSA_HIP(hipStreamSynchronize(s));  // the expanded rows are released on return
        return GSA_OK;
    });
}

int gsa_plan_describe(gsa_plan plan, int* flags) {
    GSA_CHECK(plan && flags, GSA_ERR_INVALID, "gsa_plan_describe: null argument");
    int f = 0;
    if (plan->kind == 0) f |= GSA_PLAN_WINDOWS;
    if (plan->kind == 0 && plan->csr.order) f |= GSA_PLAN_XCD_ORDER;
    if (plan->view) f |= GSA_PLAN_ROW_VIEW;
    if (plan->bins) f |= GSA_PLAN_ROW_BINS;
    if (plan->bins_t) f |= GSA_PLAN_TRANSPOSE_BINS;
    if (plan->gview || plan->gview_t) f |= GSA_PLAN_GATHER_VIEW;
    if (plan->kind == 1) f |= GSA_PLAN_BATCH;
    *flags = f;
    return GSA_OK;
}

int gsa_plan_destroy(gsa_plan plan) {
    if (!plan) return GSA_OK;
    // the plan may still be in use by queued work on any stream
    if (plan->kind == 0 || plan->view || plan->bins || plan->bins_t || plan->gview ||
        plan->gview_t)
        GSA_HIP(hipDeviceSynchronize());
    if (plan->kind == 0) GSA_TRY(plan_release(plan->csr));
    if (plan->view) release_row_view(plan->view);
    if (plan->gview) release_row_view(plan->gview);
    if (plan->gview_t) release_row_view(plan->gview_t);
    if (plan->bins) release_bin_plan(plan->bins);
    if (plan->bins_t) release_bin_plan(plan->bins_t);
    if (plan->kind == 1) batch_plan_release(plan);
    delete plan;
    return GSA_OK;
}

int gsa_spmm_csr(void* stream, int vtype, int itype, int opA, int64_t m, int64_t n, int64_t k,
                 const void* rowptr, const void* colval, const void* nzval, int64_t nnz,
                 const void* B, int64_t ldb, int opB, void* C, int64_t ldc, const void* alpha,
                 const void* beta, gsa_plan plan) {
    GSA_TRY(ensure_init());
    GSA_TRY(check_common(opA, opB, m, n, k, nnz, ldb, ldc));
    hipStream_t s = as_stream(stream);
    const int bt = opB != GSA_OP_N;                        // B read transposed (TRANSB)
    const int cjb = opB == GSA_OP_C ? 2 : 0;               // CONJB
    const int cj = (opA == GSA_OP_C ? 1 : 0) | cjb;        // CONJA | CONJB
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        const T a = host_scalar<T>(alpha, T(1)), b = host_scalar<T>(beta, T(0));
        if (opA == GSA_OP_N)
            return gather_csr<T, I>(s, m, n, k, (const I*)rowptr, (const I*)colval,
                                    (const T*)nzval, nnz, (const T*)B, ldb, bt, (T*)C, ldc, a, b,
                                    plan, itype, nullptr, (const void*)nullptr, false, cjb);
        if constexpr (std::is_floating_point_v<T>) {  // the row bins hold real products
            if (k == 1 && n > 0 && plan &&
                bin_plan_matches(plan->bins_t, itype, n, m, nnz, rowptr, colval))
                return bin_spmv(s, *plan->bins_t, vtype, nullptr, nzval, B, bt ? ldb : 1, C,
                                (double)a, (double)b);
        }
        if ((k > 1 || !std::is_floating_point_v<T>) && plan) {  // gather view
            if (const RowView* v = gview_for(plan->gview_t, itype, n, nnz, rowptr, colval))
                return view_spmv<T, I>(s, *v, m, (const I*)nullptr, (const T*)nzval,
                                       (const T*)B, ldb, bt, (T*)C, ldc, a, b, itype, k, cj);
        }
        return scatter_csr<T, I>(s, m, n, k, (const I*)rowptr, (const I*)colval, (const T*)nzval,
                                 nnz, (const T*)B, ldb, bt, (T*)C, ldc, a, b, plan, itype, cj);
    });
}

int gsa_spmv_csr_noalloc(void* stream, int vtype, int itype, int64_t m, int64_t n,
                         const void* rowptr, const void* colval, const void* nzval, int64_t nnz,
                         const void* x, void* y, double alpha, double beta) {
    GSA_TRY(ensure_init());
    GSA_CHECK(m >= 0 && n >= 0 && nnz >= 0, GSA_ERR_INVALID,
              "gsa_spmv_csr_noalloc: negative dimension (m=%lld n=%lld nnz=%lld)", (long long)m,
              (long long)n, (long long)nnz);
    GSA_CHECK(nnz <= GSA_NOALLOC_MAX_NNZ, GSA_ERR_INVALID,
              "gsa_spmv_csr_noalloc: %lld stored entries, at most %lld", (long long)nnz,
              (long long)GSA_NOALLOC_MAX_NNZ);
    GSA_CHECK(m == 0 || (rowptr && y), GSA_ERR_INVALID, "gsa_spmv_csr_noalloc: null pointer");
    GSA_CHECK(m == 0 || n == 0 || nnz == 0 || (colval && nzval && x), GSA_ERR_INVALID,
              "gsa_spmv_csr_noalloc: null pointer");
    uint8_t* ws = nullptr;
    GSA_TRY(noalloc_scratch(&ws));
    hipStream_t s = as_stream(stream);
    return dispatch_vi(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        return spmv_noalloc<T, I>(s, itype, m, n, (const I*)rowptr, (const I*)colval,
                                  (const T*)nzval, nnz, (const T*)x, (T*)y, (T)alpha, (T)beta,
                                  ws);
    });
}

int gsa_spmm_csc(void* stream, int vtype, int itype, int opA, int64_t m, int64_t n, int64_t k,
                 const void* colptr, const void* rowval, const void* nzval, int64_t nnz,
                 const void* B, int64_t ldb, int opB, void* C, int64_t ldc, const void* alpha,
                 const void* beta, gsa_plan plan) {
    GSA_TRY(ensure_init());
    GSA_TRY(check_common(opA, opB, m, n, k, nnz, ldb, ldc));
    hipStream_t s = as_stream(stream);
    const int bt = opB != GSA_OP_N;
    const int cjb = opB == GSA_OP_C ? 2 : 0;
    const int cj = (opA == GSA_OP_C ? 1 : 0) | cjb;
    // A (m x n) in CSC is Aᵀ (n x m) in CSR: A x is the scatter, Aᵀ x the gather.
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        const T a = host_scalar<T>(alpha, T(1)), b = host_scalar<T>(beta, T(0));
        if (opA == GSA_OP_N) {
            if constexpr (std::is_floating_point_v<T>) {
                if (k == 1 && m > 0 && plan &&
                    bin_plan_matches(plan->bins, itype, m, n, nnz, colptr, rowval))
                    return bin_spmv(s, *plan->bins, vtype, nullptr, nzval, B, bt ? ldb : 1, C,
                                    (double)a, (double)b);
            }
            const RowView* v = k == 1 ? view_for(plan, itype, m, nnz, colptr, rowval) : nullptr;
            if (v)  // sorted-key form: rows gathered through the plan's row view
                return view_spmv<T, I>(s, *v, n, (const I*)nullptr, (const T*)nzval, (const T*)B,
                                       ldb, bt, (T*)C, ldc, a, b, itype, 1, cjb);
            if ((k > 1 || !std::is_floating_point_v<T>) && plan) {  // gather view
                if (const RowView* g = gview_for(plan->gview, itype, m, nnz, colptr, rowval))
                    return view_spmv<T, I>(s, *g, n, (const I*)nullptr, (const T*)nzval,
                                           (const T*)B, ldb, bt, (T*)C, ldc, a, b, itype, k, cjb);
            }
            return scatter_csr<T, I>(s, n, m, k, (const I*)colptr, (const I*)rowval,
                                     (const T*)nzval, nnz, (const T*)B, ldb, bt, (T*)C, ldc, a, b,
                                     plan, itype, cjb);
        }
        return gather_csr<T, I>(s, n, m, k, (const I*)colptr, (const I*)rowval, (const T*)nzval,
                                nnz, (const T*)B, ldb, bt, (T*)C, ldc, a, b, plan, itype, nullptr,
                                (const void*)nullptr, false, cj);
    });
}

// opC of the row-major-capable entries: N = column-major C (ldc >= m), T = row-major
// C (ldc >= k).
static int check_c_layout(int opC, int64_t k, int64_t ldc) {
    GSA_CHECK(opC == GSA_OP_N || opC == GSA_OP_T, GSA_ERR_INVALID, "invalid opC %d", opC);
    if (opC == GSA_OP_T)
        GSA_CHECK(ldc >= (k > 1 ? k : 1), GSA_ERR_INVALID, "ldc %lld < k %lld for a row-major C",
                  (long long)ldc, (long long)k);
    return GSA_OK;
}

int gsa_spmm_csr_bf16(void* stream, int itype, int64_t m, int64_t n, int64_t k,
                      const void* rowptr, const void* colval, const uint16_t* nzval_bf16,
                      const float* row_scale, int64_t nnz, const float* B, int64_t ldb, int opB,
                      float* C, int64_t ldc, int opC, float alpha, float beta, gsa_plan plan) {
    GSA_TRY(ensure_init());
    GSA_TRY(check_c_layout(opC, k, ldc));
    GSA_TRY(check_common(GSA_OP_N, opB, m, n, k, nnz, ldb, opC == GSA_OP_T ? m : ldc));
    GSA_CHECK(m == 0 || row_scale, GSA_ERR_INVALID, "gsa_spmm_csr_bf16: row_scale is null");
    hipStream_t s = as_stream(stream);
    return dispatch_i(itype, [&](auto ti) {
        using I = decltype(ti);
        return gather_csr<float, I, uint16_t, void>(s, m, n, k, (const I*)rowptr,
                                                    (const I*)colval, nzval_bf16, nnz, B, ldb,
                                                    opB != GSA_OP_N, C, ldc, alpha, beta, plan,
                                                    itype, row_scale, nullptr, opC == GSA_OP_T);
    });
}

int gsa_spmm_csr_ex(void* stream, int vtype, int itype, int64_t m, int64_t n, int64_t k,
                    const void* rowptr, const void* colval, const void* nzval, int64_t nnz,
                    const void* B, int64_t ldb, int opB, void* C, int64_t ldc, int opC,
                    double alpha, double beta, gsa_plan plan) {
    GSA_TRY(ensure_init());
    GSA_TRY(check_c_layout(opC, k, ldc));
    GSA_TRY(check_common(GSA_OP_N, opB, m, n, k, nnz, ldb, opC == GSA_OP_T ? m : ldc));
    hipStream_t s = as_stream(stream);
    return dispatch_vi(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        return gather_csr<T, I, T, void>(s, m, n, k, (const I*)rowptr, (const I*)colval,
                                         (const T*)nzval, nnz, (const T*)B, ldb, opB != GSA_OP_N,
                                         (T*)C, ldc, (T)alpha, (T)beta, plan, itype, nullptr,
                                         nullptr, opC == GSA_OP_T);
    });
}

int gsa_spmm_coo(void* stream, int vtype, int itype, int opA, int64_t m, int64_t n, int64_t k,
                 const void* rowind, const void* colind, const void* nzval, int64_t nnz,
                 const void* B, int64_t ldb, int opB, void* C, int64_t ldc, const void* alpha,
                 const void* beta, gsa_plan plan) {
    GSA_TRY(ensure_init());
    GSA_TRY(check_common(opA, opB, m, n, k, nnz, ldb, ldc));
    hipStream_t s = as_stream(stream);
    const int trans = opA != GSA_OP_N;
    const int64_t rows_out = trans ? n : m, cols_in = trans ? m : n;
    const int bt = opB != GSA_OP_N;
    const int cjb = opB == GSA_OP_C ? 2 : 0;
    const int cj = (opA == GSA_OP_C ? 1 : 0) | cjb;
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) {
        using T = decltype(tv);
        using I = decltype(ti);
        const T a = host_scalar<T>(alpha, T(1)), b = host_scalar<T>(beta, T(0));
        if (rows_out == 0 || k == 0) return (int)GSA_OK;
        if constexpr (std::is_floating_point_v<T>) {  // the row bins hold real products
            if (!trans && k == 1 && plan &&
                bin_plan_matches(plan->bins, itype, m, n, nnz, rowind, colind))
                return bin_spmv(s, *plan->bins, vtype, colind, nzval, B, bt ? ldb : 1, C,
                                (double)a, (double)b);
            if (trans && k == 1 && plan &&
                bin_plan_matches(plan->bins_t, itype, n, m, nnz, rowind, colind))
                return bin_spmv(s, *plan->bins_t, vtype, rowind, nzval, B, bt ? ldb : 1, C,
                                (double)a, (double)b);
        }
        const RowView* v = (!trans && k == 1) ? view_for(plan, itype, m, nnz, rowind, colind)
                                              : nullptr;
        if (v)  // sorted-key form: rows gathered through the plan's row view
            return view_spmv<T, I>(s, *v, n, (const I*)colind, (const T*)nzval, (const T*)B, ldb,
                                   bt, (T*)C, ldc, a, b, itype, 1, cj);
        if ((k > 1 || !std::is_floating_point_v<T>) && plan) {  // gather view
            if (const RowView* g = gview_for(trans ? plan->gview_t : plan->gview, itype, rows_out,
                                             nnz, rowind, colind))
                return view_spmv<T, I>(s, *g, cols_in, (const I*)nullptr, (const T*)nzval,
                                       (const T*)B, ldb, bt, (T*)C, ldc, a, b, itype, k, cj);
        }
        GSA_TRY(apply_beta<T>(s, rows_out, k, (T*)C, ldc, b));
        if (nnz == 0 || cols_in == 0) return (int)GSA_OK;
        hipLaunchKernelGGL((k_coo_scatter<T, I>), dim3(grid_for(nnz * k, 256)), dim3(256), 0, s,
                           rows_out, cols_in, k, nnz, (const I*)rowind, (const I*)colind,
                           (const T*)nzval, trans, (const T*)B, ldb, bt, (T*)C, ldc, a, cj);
        GSA_HIP(hipGetLastError());
        return (int)GSA_OK;
    });
}

}  // extern "C"
