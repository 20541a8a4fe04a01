// gsa_reduce.hip — traversals of the stored entries: tr, sum, dot(x, A, y),
// issymmetric / ishermitian, and A + λI for CSR / CSC.
//
// "ptr form": a compressed matrix seen as `major` runs of entries (CSR: rows
// with their column indices; CSC: columns with their row indices).  The
// diagonal test (index == major) and the symmetry test read the same in both.
//
// Sums (tr, sum, dot) are accumulated in Float64 (real values), ComplexF64
// (complex values) or 64-bit wrapping integers (integer values, sign-extended:
// the low 32 bits are Int32's wrapping sum), with a fixed assignment of entries
// to workgroups and a fixed order inside each, then one workgroup adds the
// partials in order: the result does not change from run to run.  The
// reference adds with @atomic (tr) or per work-item partials (dot,
// src/matrix_csr/matrix_csr_kernels.jl:47-131), so its order is undefined.
#include "gsa_common.h"
#include "gsa_rows.h"
#include "gsa_scan.h"

#include <algorithm>
#include <cmath>
#include <limits>

namespace gsa {

constexpr int RD_PARTS = 1024;  // workgroups (partial sums) of the entry-parallel reductions

// Accumulator of each value type: Float64 for real values, ComplexF64 for
// complex values, wrapping 64-bit integers for integer values.
template <typename T> struct RAcc { using A = double; };
template <typename R> struct RAcc<cplx<R>> { using A = cplx<double>; };
template <> struct RAcc<uint32_t> { using A = uint64_t; };
template <> struct RAcc<uint64_t> { using A = uint64_t; };

template <typename T>
__device__ __forceinline__ typename RAcc<T>::A to_acc(T v) {
    if constexpr (VT<T>::cx) return cplx<double>((double)v.re, (double)v.im);
    else if constexpr (std::is_same_v<T, uint32_t>) return (uint64_t)(int64_t)(int32_t)v;
    else if constexpr (std::is_same_v<T, uint64_t>) return v;
    else return (double)v;
}

// Sum over the workgroup's 256 lanes in a fixed order; lane 0 gets the total.
template <typename A>
__device__ __forceinline__ A wg_sum_fixed(A v, A* sh) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = v + lane_xor(v, o);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    A t = A(0);
    if (threadIdx.x == 0)
        for (int w = 0; w < ROW_WG / 64; ++w) t = t + sh[w];
    __syncthreads();
    return t;
}

template <typename A>
__global__ void k_sum_parts(int64_t n, const A* __restrict__ part, A* __restrict__ out) {
    __shared__ A sh[ROW_WG / 64];
    A acc = A(0);
    for (int64_t i = threadIdx.x; i < n; i += ROW_WG) acc = acc + part[i];
    const A t = wg_sum_fixed(acc, sh);
    if (threadIdx.x == 0) *out = t;
}

// ---- tr ----------------------------------------------------------------------
// Entries whose index equals their major (kernel_tr, src/matrix_csr/matrix_csr.jl:118-142;
// src/matrix_csc/matrix_csc.jl:117-140): every such entry counts, duplicates too.
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_trace_ptr(int64_t n, int64_t nnz,
                                                      const I* __restrict__ ptr,
                                                      const I* __restrict__ idx,
                                                      const T* __restrict__ val,
                                                      const int64_t* __restrict__ blk_row,
                                                      typename RAcc<T>::A* __restrict__ part) {
    using A = typename RAcc<T>::A;
    __shared__ A sh[ROW_WG / 64];
    A acc = A(0);
    // the index and the value, read with the plan (the value of every entry:
    // a diagonal entry's value line is nearly always fetched anyway)
    struct E {
        I c;
        T v;
    };
    auto pre = [&](int64_t j) { return E{idx[j], val[j]}; };
    auto load = [](int64_t, int64_t, const E& x) { return x; };
    auto none = [](int64_t, int64_t, E&) {};
    auto use = [&](int64_t r, int64_t, const E& x) {
        if ((int64_t)x.c - 1 == r) acc = acc + to_acc(x.v);
    };
    visit_owned_entries<I>(n, nnz, ptr, blk_row, [](int64_t, int64_t, int64_t) {}, pre,
                           lane_rounds<ENT_PER_LANE>(load, none, use));
    const A t = wg_sum_fixed(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// COO: entries with rowind == colind (src/matrix_coo/matrix_coo.jl:131-150);
// with `colind` null, every stored value (sum(A) = sum(nonzeros(A)),
// src/core.jl:24).
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_trace_coo(int64_t nnz, const I* __restrict__ rowind,
                                                      const I* __restrict__ colind,
                                                      const T* __restrict__ val,
                                                      typename RAcc<T>::A* __restrict__ part) {
    using A = typename RAcc<T>::A;
    __shared__ A sh[ROW_WG / 64];
    A acc = A(0);
    const int64_t per = (nnz + gridDim.x - 1) / gridDim.x;  // contiguous share per workgroup
    const int64_t b0 = blockIdx.x * per, b1 = b0 + per < nnz ? b0 + per : nnz;
    for (int64_t j = b0 + threadIdx.x; j < b1; j += ROW_WG)
        if (!colind || rowind[j] == colind[j]) acc = acc + to_acc(val[j]);
    const A t = wg_sum_fixed(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---- dot(x, A, y) ------------------------------------------------------------
// Each stored entry a (major r, index i) adds conj(x[p]) * a * y[q], left to
// right (dot(x, a, y) of scalars; a conjugated under Adjoint, CONJA;
// kernel_workgroup_dot_csr_N!/T!, src/matrix_csr/matrix_csr_kernels.jl:47-131),
// with (p, q) = (r, i) when major_left, else (i, r).
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_dot3_ptr(int64_t nmaj, int64_t nidx, int64_t nnz,
                                                     const I* __restrict__ ptr,
                                                     const I* __restrict__ idx,
                                                     const T* __restrict__ val,
                                                     const T* __restrict__ x,
                                                     const T* __restrict__ y, int major_left,
                                                     int cja,
                                                     const int64_t* __restrict__ blk_row,
                                                     typename RAcc<T>::A* __restrict__ part) {
    using A = typename RAcc<T>::A;
    __shared__ A sh[ROW_WG / 64];
    A acc = A(0);
    // the index and the value, read with the plan; then both vectors' entries
    // (the major's, and the other's gathered through the index) in one round
    struct P {
        I c;
        T a;
    };
    struct E {
        T a, wr, wi;
    };
    auto pre = [&](int64_t j) { return P{idx[j], val[j]}; };
    auto load = [&](int64_t r, int64_t, const P& p) {
        const int64_t i = dclamp((int64_t)p.c - 1, 0, nidx - 1);
        return E{p.a, major_left ? x[r] : y[r], major_left ? y[i] : x[i]};
    };
    auto none = [](int64_t, int64_t, E&) {};
    auto use = [&](int64_t, int64_t, const E& e) {
        const T xv = conj_v(major_left ? e.wr : e.wi);
        const T yv = major_left ? e.wi : e.wr;
        const T t = (xv * conj_if(e.a, cja != 0)) * yv;
        acc = acc + to_acc(t);
    };
    visit_owned_entries<I>(nmaj, nnz, ptr, blk_row, [](int64_t, int64_t, int64_t) {}, pre,
                           lane_rounds<ENT_PER_LANE / 2>(load, none, use));
    const A t = wg_sum_fixed(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// sum of w[i] * t[i] per workgroup (double), grid-stride, fixed order.
template <typename T>
__global__ __launch_bounds__(ROW_WG) void k_dot2(int64_t n, const T* __restrict__ w,
                                                 const T* __restrict__ t,
                                                 double* __restrict__ part) {
    __shared__ double sh[ROW_WG / 64];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)ROW_WG + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * ROW_WG)
        acc += (double)(w[i] * t[i]);
    const double r = wg_sum_fixed<double>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// COO: x[p] * a * y[q] with (p, q) = (rowind, colind), or swapped for opA = T.
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_dot3_coo(int64_t m, int64_t n, int64_t nnz,
                                                     const I* __restrict__ rowind,
                                                     const I* __restrict__ colind,
                                                     const T* __restrict__ val,
                                                     const T* __restrict__ x,
                                                     const T* __restrict__ y, int trans,
                                                     int cja,
                                                     typename RAcc<T>::A* __restrict__ part) {
    using A = typename RAcc<T>::A;
    __shared__ A sh[ROW_WG / 64];
    A acc = A(0);
    const int64_t per = (nnz + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = blockIdx.x * per, b1 = b0 + per < nnz ? b0 + per : nnz;
    for (int64_t j = b0 + threadIdx.x; j < b1; j += ROW_WG) {
        const int64_t r = dclamp((int64_t)rowind[j] - 1, 0, m - 1);
        const int64_t c = dclamp((int64_t)colind[j] - 1, 0, n - 1);
        const T a = conj_if(val[j], cja != 0);
        const T t = trans ? (conj_v(x[c]) * a) * y[r] : (conj_v(x[r]) * a) * y[c];
        acc = acc + to_acc(t);
    }
    const A t = wg_sum_fixed(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---- issymmetric / ishermitian -------------------------------------------------
// For every stored entry (major r, index c != r): a binary search of major c
// for index r, exactly as kernel_check_symmetry_csr!/csc!
// (src/matrix_csr/matrix_csr_kernels.jl:348-407, src/matrix_csc/matrix_csc_kernels.jl:350-405)
// does it; the entry must be found and equal (Hermitian: equal to the
// conjugate, and every diagonal value equal to its own conjugate).  flag = 0 on
// the first failure seen.
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_symmetric_ptr(int64_t n, int64_t nnz,
                                                          const I* __restrict__ ptr,
                                                          const I* __restrict__ idx,
                                                          const T* __restrict__ val,
                                                          int hermitian,
                                                          const int64_t* __restrict__ blk_row,
                                                          int* __restrict__ flag) {
    // G entries per lane searched in lock step: every round of the binary
    // searches issues its G index loads together.  Each search probes the
    // same positions as a search of its own would, so it finds the same entry.
    constexpr int G = 2;  // 53 us on the 1M-row Laplacian; 54 with 4, 59 with 1, 68 with 8
    struct E {
        I c;
        T v;
    };
    auto pre = [&](int64_t j) { return E{idx[j], val[j]}; };
    auto batch = [&](const int64_t* rr, const int64_t*, const bool* okk, const E* pe, int64_t) {
        bool fail = false;
#pragma unroll
        for (int g = 0; g < ENT_PER_LANE; g += G) {
            const int64_t* r = rr + g;
            const bool* ok = okk + g;
            const E* p = pe + g;
            int64_t lo[G], hi[G], pos[G];
            bool act[G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const int64_t cc = (int64_t)p[u].c - 1;
                if (ok[u] && cc == r[u] && hermitian && p[u].v != conj_v(p[u].v)) fail = true;
                const bool off = ok[u] && cc != r[u];
                act[u] = off && cc >= 0 && cc < n;
                if (off && !act[u]) fail = true;  // an index outside the matrix: no mirror
                const int64_t cs = act[u] ? cc : 0;
                lo[u] = (int64_t)ptr[cs];  // one-based, inclusive
                hi[u] = (int64_t)ptr[cs + 1] - 1;
                pos[u] = -1;
            }
            bool any = false;
#pragma unroll
            for (int u = 0; u < G; ++u) {
                act[u] = act[u] && lo[u] <= hi[u];
                any |= act[u];
            }
            for (int round = 0; any && round < 64; ++round) {
                int64_t mid[G];
                I mc[G];
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    mid[u] = lo[u] + (hi[u] - lo[u]) / 2;
                    mc[u] = idx[act[u] ? dclamp(mid[u] - 1, 0, nnz - 1) : 0];
                }
                any = false;
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    if (!act[u]) continue;
                    const int64_t m1 = (int64_t)mc[u] - 1;
                    if (m1 == r[u]) {
                        pos[u] = dclamp(mid[u] - 1, 0, nnz - 1);
                        act[u] = false;
                    } else {
                        if (m1 < r[u]) lo[u] = mid[u] + 1;
                        else hi[u] = mid[u] - 1;
                        act[u] = lo[u] <= hi[u];
                    }
                    any |= act[u];
                }
            }
            T tv[G];
#pragma unroll
            for (int u = 0; u < G; ++u) tv[u] = val[pos[u] >= 0 ? pos[u] : 0];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const int64_t cc = (int64_t)p[u].c - 1;
                if (!ok[u] || cc == r[u] || cc < 0 || cc >= n) continue;
                if (pos[u] < 0 || p[u].v != (hermitian ? conj_v(tv[u]) : tv[u])) fail = true;
            }
        }
        if (fail) *flag = 0;
    };
    visit_owned_entries<I>(n, nnz, ptr, blk_row, [](int64_t, int64_t, int64_t) {}, pre, batch);
}

// ---- A + λI ------------------------------------------------------------------
// Per major: one more entry when none of its entries sits on the diagonal
// (kernel_count_diag_entries_csr!, src/matrix_csr/matrix_csr_kernels.jl:445-463).
// cnt[r + 1] starts at the largest value (k_count_fill); the major's lane then
// takes it down to length + 1 and any diagonal entry to length, with atomic
// minima: the entries of a major can fall in other workgroups than its own,
// and the result does not depend on the order the writes arrive in.
template <typename I>
__global__ void k_count_fill(int64_t n, I* __restrict__ cnt) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x)
        cnt[r + 1] = std::numeric_limits<I>::max();
}

template <typename I>
__device__ __forceinline__ void atomic_min_i(I* p, I v) {
    if constexpr (sizeof(I) == 8) atomicMin(reinterpret_cast<long long*>(p), (long long)v);
    else atomicMin(reinterpret_cast<int*>(p), (int)v);
}

template <typename I>
__global__ __launch_bounds__(ROW_WG) void k_diag_missing(int64_t n, int64_t nnz,
                                                         const I* __restrict__ ptr,
                                                         const I* __restrict__ idx,
                                                         const int64_t* __restrict__ blk_row,
                                                         I* __restrict__ cnt) {
    // a diagonal entry reads its major's length (a second round, taken by the
    // diagonal entries alone)
    struct E {
        I c, len;
    };
    auto pre = [&](int64_t j) { return E{idx[j], I(0)}; };
    auto load = [](int64_t, int64_t, const E& x) { return x; };
    auto fetch = [&](int64_t r, int64_t, E& x) {
        if ((int64_t)x.c - 1 == r) x.len = (I)((int64_t)ptr[r + 1] - (int64_t)ptr[r]);
    };
    auto use = [&](int64_t r, int64_t, const E& x) {
        if ((int64_t)x.c - 1 == r) atomic_min_i(cnt + r + 1, x.len);
    };
    visit_owned_entries<I>(
        n, nnz, ptr, blk_row,
        [&](int64_t r, int64_t s, int64_t e) { atomic_min_i(cnt + r + 1, (I)((e - s) + 1)); }, pre,
        lane_rounds<ENT_PER_LANE>(load, fetch, use));
}

// The merge of kernel_add_uniformscaling_csr! (matrix_csr_kernels.jl:466-509)
// for rows whose indices ascend: the entries in storage order with λ added to
// every diagonal entry; when the major has no diagonal entry, λ inserted before
// its first index past the diagonal (or appended).  Entry j of the major moves
// to ptr_out + (j - start) + (1 if λ was inserted before it); whether λ is
// inserted is ptr_out's row length against ptr's.  λ itself is written by the
// first entry past the diagonal, else by the last entry, else (an empty
// major) by the major's lane.  Writes stay inside the major's new range.
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_add_diag(int64_t n, int64_t nnz,
                                                     const I* __restrict__ ptr,
                                                     const I* __restrict__ idx,
                                                     const T* __restrict__ val, T lambda,
                                                     const int64_t* __restrict__ blk_row,
                                                     const I* __restrict__ ptr_out,
                                                     I* __restrict__ idx_out,
                                                     T* __restrict__ val_out) {
    auto put = [&](int64_t w, int64_t wend, I c, T v) {
        if (w < wend) {
            idx_out[w] = c;
            val_out[w] = v;
        }
    };
    // the entry read with the plan; then its major's pointers (old and new);
    // then the previous index, for an entry past the diagonal of a major
    // missing one
    struct P {
        I c;
        T v;
    };
    struct E {
        I c, ps, pe, os, oe, prev;
        T v;
    };
    auto pre = [&](int64_t j) { return P{idx[j], val[j]}; };
    auto load = [&](int64_t r, int64_t, const P& p) {
        return E{p.c, ptr[r], ptr[r + 1], ptr_out[r], ptr_out[r + 1], I(0), p.v};
    };
    auto fetch = [&](int64_t r, int64_t j, E& x) {
        const int64_t s = (int64_t)x.ps - 1, e = (int64_t)x.pe - 1;
        const bool missing = (int64_t)x.oe - (int64_t)x.os > e - s;
        if (missing && (int64_t)x.c - 1 > r && j > s) x.prev = idx[j - 1];
    };
    auto use = [&](int64_t r, int64_t j, const E& x) {
        const int64_t s = (int64_t)x.ps - 1, e = (int64_t)x.pe - 1;
        const int64_t base = (int64_t)x.os - 1, wend = (int64_t)x.oe - 1;
        const bool missing = wend - base > e - s;
        const int64_t c = (int64_t)x.c - 1;
        const bool past = missing && c > r;
        const int64_t w = base + (j - s) + (past ? 1 : 0);
        put(w, wend, x.c, c == r ? x.v + lambda : x.v);
        if (past && (j == s || (int64_t)x.prev - 1 < r)) put(w - 1, wend, (I)(r + 1), lambda);
        if (missing && !past && j == e - 1) put(w + 1, wend, (I)(r + 1), lambda);
    };
    visit_owned_entries<I>(
        n, nnz, ptr, blk_row,
        [&](int64_t r, int64_t s, int64_t e) {  // an empty major gets λ alone
            if (s == e && ptr_out[r + 1] > ptr_out[r])
                put((int64_t)ptr_out[r] - 1, (int64_t)ptr_out[r + 1] - 1, (I)(r + 1), lambda);
        },
        pre, lane_rounds<1>(load, fetch, use));  // one entry per round: 75 us, 80 with two
}

__global__ void k_set_flag(int* f) { *f = 1; }

// The n diagonal entries (i, i, d_i) of a COO matrix, d_i = dvals[i] or λ
// (kernel_fill_diagonal_coo!, src/matrix_coo/matrix_coo_kernels.jl:312-325).
template <typename T, typename I>
__global__ void k_diag_coo(int64_t n, const T* __restrict__ dvals, T lambda, I* __restrict__ row,
                           I* __restrict__ col, T* __restrict__ val) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        row[i] = (I)(i + 1);
        col[i] = (I)(i + 1);
        val[i] = dvals ? dvals[i] : lambda;
    }
}

// ---- GenericSparseVector: dot(x, y) and norms of the stored values ---------------
// dot(x::GenericSparseVector, y::DenseVector) (src/vector/vector.jl:102-145):
// sum over stored entries of dot(nzval[i], y[nzind[i]]) = conj(nzval[i]) * y[..],
// the reference's @atomic sum replaced by fixed-order per-workgroup partials.
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_spvec_dot(int64_t n, int64_t nnz,
                                                      const I* __restrict__ nzind,
                                                      const T* __restrict__ nzval,
                                                      const T* __restrict__ y,
                                                      typename RAcc<T>::A* __restrict__ part) {
    using A = typename RAcc<T>::A;
    __shared__ A sh[ROW_WG / 64];
    A acc = A(0);
    const int64_t per = (nnz + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = blockIdx.x * per, b1 = b0 + per < nnz ? b0 + per : nnz;
    for (int64_t j = b0 + threadIdx.x; j < b1; j += ROW_WG) {
        const int64_t i = dclamp((int64_t)nzind[j] - 1, 0, n - 1);
        acc = acc + to_acc(conj_v(nzval[j]) * y[i]);
    }
    const A t = wg_sum_fixed(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// |v| in Float64 (complex: the modulus, by hypot).
template <typename T>
__device__ __forceinline__ double abs_d(T v) {
    if constexpr (VT<T>::cx) return hypot((double)v.re, (double)v.im);
    else if constexpr (std::is_same_v<T, uint32_t>) return fabs((double)(int32_t)v);
    else if constexpr (std::is_same_v<T, uint64_t>) return fabs((double)(int64_t)v);
    else return fabs((double)v);
}

// Per workgroup: max |v| and min |v| over a contiguous share of the values.
template <typename T>
__global__ __launch_bounds__(ROW_WG) void k_abs_extrema(int64_t nnz, const T* __restrict__ val,
                                                        double* __restrict__ part) {
    __shared__ double smax[ROW_WG / 64], smin[ROW_WG / 64];
    double mx = 0.0, mn = INFINITY;
    const int64_t per = (nnz + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = blockIdx.x * per, b1 = b0 + per < nnz ? b0 + per : nnz;
    for (int64_t j = b0 + threadIdx.x; j < b1; j += ROW_WG) {
        const double a = abs_d(val[j]);
        mx = (a > mx || a != a) ? a : mx;  // NaN propagates to the maximum
        mn = a < mn ? a : mn;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double u = __shfl_xor(mx, o, 64), w = __shfl_xor(mn, o, 64);
        mx = (u > mx || u != u) ? u : mx;
        mn = w < mn ? w : mn;
    }
    if ((threadIdx.x & 63) == 0) {
        smax[threadIdx.x >> 6] = mx;
        smin[threadIdx.x >> 6] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < ROW_WG / 64; ++w) {
            mx = (smax[w] > mx || smax[w] != smax[w]) ? smax[w] : mx;
            mn = smin[w] < mn ? smin[w] : mn;
        }
        part[2 * blockIdx.x] = mx;
        part[2 * blockIdx.x + 1] = mn;
    }
}

__global__ void k_extrema_parts(int64_t n, const double* __restrict__ part,
                                double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double mx = 0.0, mn = INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        const double a = part[2 * i], b = part[2 * i + 1];
        mx = (a > mx || a != a) ? a : mx;
        mn = b < mn ? b : mn;
    }
    out[0] = mx;
    out[1] = mn;
}

// Per workgroup: sum of (|v| / scale)^p, or of (v != 0) for p = 0.
template <typename T>
__global__ __launch_bounds__(ROW_WG) void k_pow_sum(int64_t nnz, const T* __restrict__ val,
                                                    double p, double scale,
                                                    double* __restrict__ part) {
    __shared__ double sh[ROW_WG / 64];
    double acc = 0.0;
    const int64_t per = (nnz + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = blockIdx.x * per, b1 = b0 + per < nnz ? b0 + per : nnz;
    for (int64_t j = b0 + threadIdx.x; j < b1; j += ROW_WG) {
        const double a = abs_d(val[j]);
        if (p == 0.0) acc += a != 0.0 ? 1.0 : 0.0;
        else if (p == 1.0) acc += a / scale;
        else if (p == 2.0) acc += (a / scale) * (a / scale);
        else acc += pow(a / scale, p);
    }
    const double t = wg_sum_fixed<double>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---- sparse + dense: C[i, j] += a_ij into a column-major dense C -----------
// (kernel_add_sparse_to_dense_{csr,csc,coo}!, src/matrix_csr/matrix_csr_kernels.jl:134-145,
// src/matrix_csc/matrix_csc_kernels.jl:134-145, src/matrix_coo/matrix_coo_kernels.jl:132-142).
// Every stored entry adds into its element with an atomic add: without
// duplicate indices each element gets at most one add, so the result is
// C = B + a_ij rounded once, whatever the order; duplicates add in an order the
// hardware picks (the reference's COO kernel races on them unprotected).
// Compressed formats: `cm` says whether the major index is C's column (CSC,
// or transpose(CSR)); entries are visited entry-parallel per row block.
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_add_dense_ptr(int64_t nmaj, int64_t nmin,
                                                          int64_t nnz, const I* __restrict__ ptr,
                                                          const I* __restrict__ idx,
                                                          const T* __restrict__ val, int cm,
                                                          int cj, const int64_t* __restrict__ blk_row,
                                                          T* __restrict__ C, int64_t ldc) {
    struct E {
        I c;
        T v;
    };
    auto pre = [&](int64_t j) { return E{idx[j], val[j]}; };
    auto load = [](int64_t, int64_t, const E& x) { return x; };
    auto none = [](int64_t, int64_t, E&) {};
    auto add = [&](int64_t r, int64_t, const E& x) {
        const int64_t q = dclamp((int64_t)x.c - 1, 0, nmin - 1);
        const int64_t at = cm ? q + r * ldc : r + q * ldc;
        atomic_add_v(C + at, conj_if(x.v, cj != 0));
    };
    visit_owned_entries<I>(nmaj, nnz, ptr, blk_row, [](int64_t, int64_t, int64_t) {}, pre,
                           lane_rounds<ENT_PER_LANE>(load, none, add));
}

template <typename T, typename I>
__global__ void k_add_dense_coo(int64_t m, int64_t n, int64_t nnz, const I* __restrict__ row,
                                const I* __restrict__ col, const T* __restrict__ val, int cj,
                                T* __restrict__ C, int64_t ldc) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nnz;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = dclamp((int64_t)row[j] - 1, 0, m - 1);
        const int64_t k = dclamp((int64_t)col[j] - 1, 0, n - 1);
        atomic_add_v(C + i + k * ldc, conj_if(val[j], cj != 0));
    }
}

// Row-block plan over `ptr` for the visit_owned_entries kernels: the caller's
// plan when it was built on this pointer array, else one built into `tmp`.
static int row_blocks(hipStream_t s, gsa_plan plan, int itype, int64_t n, int64_t nnz,
                      const void* ptr, Scratch& tmp, const int64_t** blk_row, int64_t* nblocks) {
    if (plan && plan->kind == 0 && plan->csr.m == n && plan->csr.nnz == nnz &&
        plan->csr.itype == itype && plan->csr.ptr == ptr) {
        *blk_row = plan->csr.blk_row;
        *nblocks = plan->csr.nblocks;
        return GSA_OK;
    }
    *nblocks = nnz / SPMV_NB + 1;
    GSA_TRY(tmp.alloc(sizeof(int64_t) * (*nblocks + 1)));
    *blk_row = tmp.as<int64_t>();
    return build_csr_plan(s, itype, n, nnz, ptr, tmp.as<int64_t>(), *nblocks);
}

// Parts -> one host accumulator value (synchronises).
template <typename A>
static int finish_sum(hipStream_t s, const A* part, int64_t nparts, void* out) {
    Scratch res(s);
    GSA_TRY(res.alloc(sizeof(A)));
    hipLaunchKernelGGL(k_sum_parts<A>, dim3(1), dim3(ROW_WG), 0, s, nparts, part, res.as<A>());
    GSA_HIP(hipGetLastError());
    GSA_HIP(hipMemcpyAsync(out, res.p, sizeof(A), hipMemcpyDeviceToHost, s));
    GSA_HIP(hipStreamSynchronize(s));
    return GSA_OK;
}

// Bytes of the accumulator a reduction of vtype hands back.
static size_t acc_bytes(int vtype) {
    return (vtype == GSA_C64 || vtype == GSA_C128) ? 16 : 8;
}

}  // namespace gsa

using namespace gsa;

extern "C" {

int gsa_trace_ptr(void* stream, int vtype, int itype, int64_t n, int64_t nnz, const void* ptr,
                  const void* idx, const void* val, gsa_plan plan, void* out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(out && n >= 0 && nnz >= 0 && (n == 0 || ptr), GSA_ERR_INVALID,
              "gsa_trace_ptr: bad arguments");
    memset(out, 0, acc_bytes(vtype));
    if (n == 0 || nnz == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) -> int {
        using T = decltype(tv);
        using I = decltype(ti);
        using A = typename RAcc<T>::A;
        Scratch tmp(s), part(s);
        const int64_t* br = nullptr;
        int64_t nb = 0;
        GSA_TRY(row_blocks(s, plan, itype, n, nnz, ptr, tmp, &br, &nb));
        GSA_TRY(part.alloc(sizeof(A) * nb));
        hipLaunchKernelGGL((k_trace_ptr<T, I>), dim3((unsigned)nb), dim3(ROW_WG), 0, s, n, nnz,
                           (const I*)ptr, (const I*)idx, (const T*)val, br, part.as<A>());
        GSA_HIP(hipGetLastError());
        return finish_sum<A>(s, part.as<A>(), nb, out);
    });
}

int gsa_trace_coo(void* stream, int vtype, int itype, int64_t nnz, const void* rowind,
                  const void* colind, const void* val, void* out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(out && nnz >= 0, GSA_ERR_INVALID, "gsa_trace_coo: bad arguments");
    memset(out, 0, acc_bytes(vtype));
    if (nnz == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) -> int {
        using T = decltype(tv);
        using I = decltype(ti);
        using A = typename RAcc<T>::A;
        Scratch part(s);
        GSA_TRY(part.alloc(sizeof(A) * RD_PARTS));
        hipLaunchKernelGGL((k_trace_coo<T, I>), dim3(RD_PARTS), dim3(ROW_WG), 0, s, nnz,
                           (const I*)rowind, (const I*)colind, (const T*)val, part.as<A>());
        GSA_HIP(hipGetLastError());
        return finish_sum<A>(s, part.as<A>(), RD_PARTS, out);
    });
}

int gsa_sum_values(void* stream, int vtype, int64_t nnz, const void* val, void* out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(out && nnz >= 0, GSA_ERR_INVALID, "gsa_sum_values: bad arguments");
    memset(out, 0, acc_bytes(vtype));
    if (nnz == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_va(vtype, [&](auto tv) -> int {
        using T = decltype(tv);
        using A = typename RAcc<T>::A;
        Scratch part(s);
        GSA_TRY(part.alloc(sizeof(A) * RD_PARTS));
        hipLaunchKernelGGL((k_trace_coo<T, int64_t>), dim3(RD_PARTS), dim3(ROW_WG), 0, s, nnz,
                           (const int64_t*)nullptr, (const int64_t*)nullptr, (const T*)val,
                           part.as<A>());
        GSA_HIP(hipGetLastError());
        return finish_sum<A>(s, part.as<A>(), RD_PARTS, out);
    });
}

int gsa_dot3_ptr(void* stream, int vtype, int itype, int64_t nmaj, int64_t nidx, int64_t nnz,
                 const void* ptr, const void* idx, const void* val, const void* x, const void* y,
                 int major_left, int conj_a, gsa_plan plan, void* out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(out && nmaj >= 0 && nidx >= 0 && nnz >= 0 && (nmaj == 0 || ptr), GSA_ERR_INVALID,
              "gsa_dot3_ptr: bad arguments");
    memset(out, 0, acc_bytes(vtype));
    if (nmaj == 0 || nidx == 0 || nnz == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    // Float64 with the matrix's SpMV plan: t = the SpMV of the stored entries
    // with the vector their index selects (the flat windows gather it), then
    // sum of w[r] * t[r] with the vector the major index selects.  Each row's
    // sum of a*v is rounded before its weight multiplies it, where the
    // reference multiplies term by term; both stay within the same normalised
    // bound, and the reference's own summation order depends on its launch.
    if (vtype == GSA_F64 && plan && plan->kind == 0 && plan->csr.m == nmaj &&
        plan->csr.nnz == nnz && plan->csr.itype == itype && plan->csr.ptr == ptr &&
        plan->csr.win_row) {
        Scratch t(s), part(s);
        GSA_TRY(t.alloc(sizeof(double) * nmaj));
        GSA_TRY(part.alloc(sizeof(double) * RD_PARTS));
        const void* v = major_left ? y : x;
        const double* w = (const double*)(major_left ? x : y);
        GSA_TRY(gsa_spmm_csr(stream, GSA_F64, itype, GSA_OP_N, nmaj, nidx, 1, ptr, idx, val, nnz,
                             v, nidx, GSA_OP_N, t.p, nmaj, nullptr, nullptr, plan));
        hipLaunchKernelGGL((k_dot2<double>), dim3(RD_PARTS), dim3(ROW_WG), 0, s, nmaj, w,
                           (const double*)t.p, part.as<double>());
        GSA_HIP(hipGetLastError());
        return finish_sum<double>(s, part.as<double>(), RD_PARTS, out);
    }
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) -> int {
        using T = decltype(tv);
        using I = decltype(ti);
        using A = typename RAcc<T>::A;
        Scratch tmp(s), part(s);
        const int64_t* br = nullptr;
        int64_t nb = 0;
        GSA_TRY(row_blocks(s, plan, itype, nmaj, nnz, ptr, tmp, &br, &nb));
        GSA_TRY(part.alloc(sizeof(A) * nb));
        hipLaunchKernelGGL((k_dot3_ptr<T, I>), dim3((unsigned)nb), dim3(ROW_WG), 0, s, nmaj, nidx,
                           nnz, (const I*)ptr, (const I*)idx, (const T*)val, (const T*)x,
                           (const T*)y, major_left, conj_a, br, part.as<A>());
        GSA_HIP(hipGetLastError());
        return finish_sum<A>(s, part.as<A>(), nb, out);
    });
}

int gsa_dot3_coo(void* stream, int vtype, int itype, int64_t m, int64_t n, int64_t nnz,
                 const void* rowind, const void* colind, const void* val, const void* x,
                 const void* y, int trans, int conj_a, void* out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(out && m >= 0 && n >= 0 && nnz >= 0, GSA_ERR_INVALID, "gsa_dot3_coo: bad arguments");
    memset(out, 0, acc_bytes(vtype));
    if (m == 0 || n == 0 || nnz == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) -> int {
        using T = decltype(tv);
        using I = decltype(ti);
        using A = typename RAcc<T>::A;
        Scratch part(s);
        GSA_TRY(part.alloc(sizeof(A) * RD_PARTS));
        hipLaunchKernelGGL((k_dot3_coo<T, I>), dim3(RD_PARTS), dim3(ROW_WG), 0, s, m, n, nnz,
                           (const I*)rowind, (const I*)colind, (const T*)val, (const T*)x,
                           (const T*)y, trans, conj_a, part.as<A>());
        GSA_HIP(hipGetLastError());
        return finish_sum<A>(s, part.as<A>(), RD_PARTS, out);
    });
}

int gsa_symmetric_ptr(void* stream, int vtype, int itype, int64_t n, int64_t nnz,
                      const void* ptr, const void* idx, const void* val, int hermitian,
                      gsa_plan plan, int* out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(out && n >= 0 && nnz >= 0 && (n == 0 || ptr), GSA_ERR_INVALID,
              "gsa_symmetric_ptr: bad arguments");
    *out = 1;
    if (n == 0 || nnz == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) -> int {
        using T = decltype(tv);
        using I = decltype(ti);
        Scratch tmp(s), flag(s);
        const int64_t* br = nullptr;
        int64_t nb = 0;
        GSA_TRY(row_blocks(s, plan, itype, n, nnz, ptr, tmp, &br, &nb));
        GSA_TRY(flag.alloc(sizeof(int)));
        hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(1), 0, s, flag.as<int>());
        hipLaunchKernelGGL((k_symmetric_ptr<T, I>), dim3((unsigned)nb), dim3(ROW_WG), 0, s, n,
                           nnz, (const I*)ptr, (const I*)idx, (const T*)val, hermitian, br,
                           flag.as<int>());
        GSA_HIP(hipGetLastError());
        GSA_HIP(hipMemcpyAsync(out, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
        GSA_HIP(hipStreamSynchronize(s));
        return GSA_OK;
    });
}

int gsa_add_diag_count(void* stream, int itype, int64_t n, int64_t nnz, const void* ptr,
                       const void* idx, gsa_plan plan, void* ptr_out, int64_t* nnz_out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(nnz_out && n >= 0 && nnz >= 0 && ptr && ptr_out, GSA_ERR_INVALID,
              "gsa_add_diag_count: bad arguments");
    hipStream_t s = as_stream(stream);
    return dispatch_i(itype, [&](auto ti) -> int {
        using I = decltype(ti);
        Scratch tmp(s);
        if (n > 0) {
            const int64_t* br = nullptr;
            int64_t nb = 0;
            GSA_TRY(row_blocks(s, plan, itype, n, nnz, ptr, tmp, &br, &nb));
            hipLaunchKernelGGL(k_count_fill<I>, dim3((unsigned)std::min<int64_t>(cdiv(n, 256), 4096)),
                               dim3(256), 0, s, n, (I*)ptr_out);
            hipLaunchKernelGGL(k_diag_missing<I>, dim3((unsigned)nb), dim3(ROW_WG), 0, s, n, nnz,
                               (const I*)ptr, (const I*)idx, br, (I*)ptr_out);
            GSA_HIP(hipGetLastError());
        }
        return counts_to_ptr<I>(s, (I*)ptr_out, n, nnz_out);
    });
}

int gsa_add_diag_fill(void* stream, int vtype, int itype, int64_t n, int64_t nnz,
                      const void* ptr, const void* idx, const void* val, const void* lambda,
                      gsa_plan plan, const void* ptr_out, void* idx_out, void* val_out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(n >= 0 && nnz >= 0 && ptr && ptr_out && lambda, GSA_ERR_INVALID,
              "gsa_add_diag_fill: bad arguments");
    if (n == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) -> int {
        using T = decltype(tv);
        using I = decltype(ti);
        const T lam = host_scalar<T>(lambda, T(0));
        Scratch tmp(s);
        const int64_t* br = nullptr;
        int64_t nb = 0;
        GSA_TRY(row_blocks(s, plan, itype, n, nnz, ptr, tmp, &br, &nb));
        hipLaunchKernelGGL((k_add_diag<T, I>), dim3((unsigned)nb), dim3(ROW_WG), 0, s, n, nnz,
                           (const I*)ptr, (const I*)idx, (const T*)val, lam, br,
                           (const I*)ptr_out, (I*)idx_out, (T*)val_out);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    });
}

int gsa_add_to_dense(void* stream, int vtype, int itype, int format, int conj, int64_t m,
                     int64_t n, int64_t nnz, const void* a, const void* b, const void* val,
                     void* C, int64_t ldc) {
    GSA_TRY(ensure_init());
    GSA_CHECK(m >= 0 && n >= 0 && nnz >= 0 && C && ldc >= (m > 0 ? m : 1) &&
              (format == GSA_FMT_CSR || format == GSA_FMT_CSC || format == GSA_FMT_COO) &&
              (nnz == 0 || (a && b && val)) && (format == GSA_FMT_COO || a),
              GSA_ERR_INVALID, "gsa_add_to_dense: bad arguments");
    if (m == 0 || n == 0 || nnz == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) -> int {
        using T = decltype(tv);
        using I = decltype(ti);
        if (format == GSA_FMT_COO) {
            hipLaunchKernelGGL((k_add_dense_coo<T, I>), dim3(grid_for(nnz, 256)), dim3(256), 0, s, m,
                               n, nnz, (const I*)a, (const I*)b, (const T*)val, conj, (T*)C, ldc);
            GSA_HIP(hipGetLastError());
            return GSA_OK;
        }
        const bool cm = format == GSA_FMT_CSC;
        const int64_t nmaj = cm ? n : m, nmin = cm ? m : n;
        Scratch tmp(s);
        const int64_t* br = nullptr;
        int64_t nb = 0;
        GSA_TRY(row_blocks(s, nullptr, itype, nmaj, nnz, a, tmp, &br, &nb));
        hipLaunchKernelGGL((k_add_dense_ptr<T, I>), dim3((unsigned)nb), dim3(ROW_WG), 0, s, nmaj,
                           nmin, nnz, (const I*)a, (const I*)b, (const T*)val, cm ? 1 : 0, conj, br,
                           (T*)C, ldc);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    });
}

int gsa_diag_coo(void* stream, int vtype, int itype, int64_t n, const void* dvals,
                 const void* lambda, void* row, void* col, void* val) {
    GSA_TRY(ensure_init());
    GSA_CHECK(n >= 0 && (n == 0 || (row && col && val)) && (dvals || lambda), GSA_ERR_INVALID,
              "gsa_diag_coo: bad arguments");
    if (n == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) -> int {
        using T = decltype(tv);
        using I = decltype(ti);
        const T lam = host_scalar<T>(lambda, T(0));
        hipLaunchKernelGGL((k_diag_coo<T, I>), dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                           (const T*)dvals, lam, (I*)row, (I*)col, (T*)val);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    });
}

int gsa_spvec_dot(void* stream, int vtype, int itype, int64_t n, int64_t nnz, const void* nzind,
                  const void* nzval, const void* y, void* out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(out && n >= 0 && nnz >= 0 && (nnz == 0 || (nzind && nzval && y)), GSA_ERR_INVALID,
              "gsa_spvec_dot: bad arguments");
    memset(out, 0, acc_bytes(vtype));
    if (nnz == 0 || n == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_vai(vtype, itype, [&](auto tv, auto ti) -> int {
        using T = decltype(tv);
        using I = decltype(ti);
        using A = typename RAcc<T>::A;
        Scratch part(s);
        GSA_TRY(part.alloc(sizeof(A) * RD_PARTS));
        hipLaunchKernelGGL((k_spvec_dot<T, I>), dim3(RD_PARTS), dim3(ROW_WG), 0, s, n, nnz,
                           (const I*)nzind, (const T*)nzval, (const T*)y, part.as<A>());
        GSA_HIP(hipGetLastError());
        return finish_sum<A>(s, part.as<A>(), RD_PARTS, out);
    });
}

int gsa_norm_values(void* stream, int vtype, int64_t nnz, const void* val, double p,
                    double* out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(out && nnz >= 0, GSA_ERR_INVALID, "gsa_norm_values: bad arguments");
    *out = 0.0;
    if (nnz == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    return dispatch_va(vtype, [&](auto tv) -> int {
        using T = decltype(tv);
        Scratch part(s), res(s);
        GSA_TRY(part.alloc(sizeof(double) * 2 * RD_PARTS));
        GSA_TRY(res.alloc(sizeof(double) * 2));
        // pass 1: the largest and smallest |v| (the scale of a p-norm, and the
        // norms of order +Inf / -Inf)
        hipLaunchKernelGGL((k_abs_extrema<T>), dim3(RD_PARTS), dim3(ROW_WG), 0, s, nnz,
                           (const T*)val, part.as<double>());
        hipLaunchKernelGGL(k_extrema_parts, dim3(1), dim3(ROW_WG), 0, s, (int64_t)RD_PARTS,
                           (const double*)part.as<double>(), res.as<double>());
        GSA_HIP(hipGetLastError());
        double ext[2] = {0.0, 0.0};
        GSA_HIP(hipMemcpyAsync(ext, res.p, sizeof ext, hipMemcpyDeviceToHost, s));
        GSA_HIP(hipStreamSynchronize(s));
        if (p == INFINITY) { *out = ext[0]; return (int)GSA_OK; }
        if (p == -INFINITY) { *out = ext[1]; return (int)GSA_OK; }
        const double scale = ext[0];
        if (p != 0.0 && (scale == 0.0 || !(scale < INFINITY))) { *out = scale; return (int)GSA_OK; }
        // pass 2: sum of (|v| / scale)^p (p = 0: the count of nonzero values)
        hipLaunchKernelGGL((k_pow_sum<T>), dim3(RD_PARTS), dim3(ROW_WG), 0, s, nnz, (const T*)val,
                           p, scale, part.as<double>());
        GSA_HIP(hipGetLastError());
        double sum = 0.0;
        GSA_TRY(finish_sum<double>(s, part.as<double>(), RD_PARTS, &sum));
        if (p == 0.0) *out = sum;
        else if (p == 1.0) *out = scale * sum;
        else if (p == 2.0) *out = scale * sqrt(sum);
        else *out = scale * pow(sum, 1.0 / p);
        return (int)GSA_OK;
    });
}

}  // extern "C"
