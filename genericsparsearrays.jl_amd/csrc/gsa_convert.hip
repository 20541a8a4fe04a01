// gsa_convert.hip — CSR/CSC <-> COO and CSR <-> CSC on gfx950.
//
// CSR -> COO expands the row pointer (kernel_csr_to_coo!,
// src/conversions/conversion_kernels.jl:20-35) over fixed blocks of entries of
// a row-block plan, so long rows are split over workgroups.  COO -> CSR is the reference's
// keyed stable sort (conversions.jl:282-323): 64-bit keys (row-1)*n + (col-1)
// — the same order as the reference's Ti keys rowind*n + colind, without their
// Int32 overflow — radix-sorted stably with each entry's (column, value) pair
// carried through the passes, then a row pointer from the sorted keys (equal to
// searchsortedfirst over the sorted rows).  CSR -> CSC is a stable sort on the
// column alone: CSR storage is already row-ordered, so the result equals the
// reference's CSR -> COO -> CSC route (conversions.jl:85-86) bit for bit.
#include "gsa_scan.h"
#include "gsa_rows.h"
#include "gsa_radix_pairs.h"

namespace gsa {

struct NoPre {};

// Each workgroup's 2048 entries get their rows from the LDS row map of
// visit_owned_entries and store them coalesced: a long row is split over the
// workgroups its entries fall in.
template <typename I>
__global__ __launch_bounds__(ROW_WG) void k_expand_rows(int64_t m, int64_t nnz,
                                                        const I* __restrict__ ptr,
                                                        const int64_t* __restrict__ blk_row,
                                                        I* __restrict__ out) {
    visit_owned_entries<I>(
        m, nnz, ptr, blk_row, [](int64_t, int64_t, int64_t) {}, [](int64_t) { return NoPre{}; },
        [&](const int64_t* r, const int64_t* j, const bool* ok, const NoPre*, int64_t) {
#pragma unroll
            for (int u = 0; u < ENT_PER_LANE; ++u)
                if (ok[u]) out[j[u]] = (I)(r[u] + 1);
        });
}

// CSR -> COO in one pass: each workgroup's 2048 entries read their column and
// value with the plan's row bounds and store row, column and value together.
template <typename T, typename I>
struct ColVal {
    I c;
    T v;
};
template <typename T, typename I>
__global__ __launch_bounds__(ROW_WG) void k_csr_to_coo(int64_t m, int64_t nnz,
                                                       const I* __restrict__ ptr,
                                                       const int64_t* __restrict__ blk_row,
                                                       const I* __restrict__ colval,
                                                       const T* __restrict__ nzval,
                                                       I* __restrict__ rowind,
                                                       I* __restrict__ colind,
                                                       T* __restrict__ nzval_out) {
    visit_owned_entries<I>(
        m, nnz, ptr, blk_row, [](int64_t, int64_t, int64_t) {},
        [&](int64_t j) { return ColVal<T, I>{colval[j], nzval[j]}; },
        [&](const int64_t* r, const int64_t* j, const bool* ok, const ColVal<T, I>* p, int64_t) {
#pragma unroll
            for (int u = 0; u < ENT_PER_LANE; ++u) {
                if (j[u] >= nnz) continue;
                if (ok[u]) rowind[j[u]] = (I)(r[u] + 1);
                colind[j[u]] = p[u].c;
                nzval_out[j[u]] = p[u].v;
            }
        });
}

int expand_rows(hipStream_t s, int itype, int64_t m, int64_t nnz, const void* ptr, void* out) {
    if (m == 0 || nnz == 0) return GSA_OK;
    const int64_t nblocks = nnz / SPMV_NB + 1;
    Scratch plan(s);
    GSA_TRY(plan.alloc(sizeof(int64_t) * (nblocks + 1)));
    GSA_TRY(build_csr_plan(s, itype, m, nnz, ptr, plan.as<int64_t>(), nblocks));
    return dispatch_i(itype, [&](auto ti) {
        using I = decltype(ti);
        hipLaunchKernelGGL(k_expand_rows<I>, dim3((unsigned)nblocks), dim3(ROW_WG), 0, s, m, nnz,
                           (const I*)ptr, (const int64_t*)plan.as<int64_t>(), (I*)out);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    });
}

// keys = (major-1)*minor_dim + (minor-1), payload = position
template <typename I, typename K, typename V>
__global__ void k_make_keys(int64_t nnz, const I* __restrict__ major, const I* __restrict__ minor,
                            int64_t minor_dim, int64_t major_dim, K* __restrict__ keys,
                            V* __restrict__ pos) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t a = dclamp((int64_t)major[i] - 1, 0, major_dim - 1);
        int64_t b = minor ? dclamp((int64_t)minor[i] - 1, 0, minor_dim - 1) : 0;
        keys[i] = (K)((uint64_t)a * (uint64_t)minor_dim + (uint64_t)b);
        pos[i] = (V)i;
    }
}

// One-based row pointer from non-decreasing one-based rows (as k_ptr_from_keys).
template <typename I>
__global__ void k_ptr_from_major(int64_t nnz, const I* __restrict__ major, int64_t major_dim,
                                 I* __restrict__ ptr) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t prev = -1, cur = major_dim;
        if (i > 0) prev = dclamp((int64_t)major[i - 1] - 1, 0, major_dim - 1);
        if (i < nnz) cur = dclamp((int64_t)major[i] - 1, 0, major_dim - 1);
        for (int64_t r = prev + 1; r <= cur; ++r) ptr[r] = (I)(i + 1);
    }
}

// Order of COO storage: bit 0 of *flag when (major, minor) pairs decrease
// somewhere, bit 1 when the minor indices alone do.
template <typename I>
__global__ void k_order_flags(int64_t nnz, const I* __restrict__ major, const I* __restrict__ minor,
                              int* __restrict__ flag) {
    int f = 0;
    for (int64_t i = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        const I a0 = major[i - 1], a1 = major[i], b0 = minor[i - 1], b1 = minor[i];
        if (a1 < a0 || (a1 == a0 && b1 < b0)) f |= 1;
        if (b1 < b0) f |= 2;
    }
    // one update per workgroup, and none once another has set the same bits
    f = (__syncthreads_or(f & 1) ? 1 : 0) | (__syncthreads_or(f & 2) ? 2 : 0);
    if (threadIdx.x == 0 && (f & ~*(volatile int*)flag)) atomicOr(flag, f);
}

// One-based major pointer from the sorted keys (major = key / kdim): ptr[r]
// is one plus the first position whose major is >= r — searchsortedfirst over
// the sorted majors (conversions.jl:194-199).  Position i writes the pointers
// of the majors in (major[i-1], major[i]], so every entry of ptr[0..major_dim]
// is written once, with no atomics.
template <typename I, typename K>
__global__ void k_ptr_from_keys(int64_t nnz, const K* __restrict__ keys, uint64_t kdim,
                                int64_t major_dim, I* __restrict__ ptr) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= nnz;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t prev = -1, cur = major_dim;
        if (i > 0) prev = dclamp((int64_t)((uint64_t)keys[i - 1] / kdim), 0, major_dim - 1);
        if (i < nnz) cur = dclamp((int64_t)((uint64_t)keys[i] / kdim), 0, major_dim - 1);
        for (int64_t r = prev + 1; r <= cur; ++r) ptr[r] = (I)(i + 1);
    }
}

// Sort the entries by (major, minor) [minor may be null: by major alone] and
// write minor indices, values and the major pointer.
template <typename T, typename I>
static int sort_to_compressed(hipStream_t s, int64_t major_dim, int64_t minor_dim, int64_t nnz,
                              const I* major, const I* minor_key, const I* minor_out,
                              const T* val, I* ptr_out, I* idx_out, T* val_out,
                              const I* minor_ptr = nullptr) {
    GSA_CHECK(nnz == 0 || (major_dim > 0 && (minor_key == nullptr || minor_dim > 0)),
              GSA_ERR_INVALID, "stored entries in a matrix with a zero dimension");
    if (nnz == 0) {  // every pointer is 1 (conversions.jl:178-184, 289-295)
        GSA_TRY(zero_fill(s, ptr_out, sizeof(I) * (major_dim + 1)));
        return counts_to_ptr<I>(s, ptr_out, major_dim, nullptr);
    }
    const unsigned g = grid_for(nnz, 256);
    const bool transposition = minor_key == nullptr;  // CSR <-> CSC: key = major only
    // COO storage: one check of the (major, minor) order first.  In order (a COO
    // made from a CSR, or from a CSC for CSC): the stable sort is the identity and
    // the arrays are copied.  Ordered by the minor index only (a COO from the
    // other compressed format): the stable sort by the major index alone gives
    // the (major, minor) order, with far fewer key bits.  The call synchronises
    // once to learn which.
    const I* kminor = minor_key;
    if (!transposition) {
        Scratch flag(s);
        GSA_TRY(flag.alloc(sizeof(int)));
        GSA_TRY(zero_fill(s, flag.p, sizeof(int)));
        hipLaunchKernelGGL((k_order_flags<I>), dim3(g), dim3(256), 0, s, nnz, major, minor_key,
                           flag.as<int>());
        int flags = 3;
        GSA_HIP(hipMemcpyAsync(&flags, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
        GSA_HIP(hipStreamSynchronize(s));
        if (!(flags & 1)) {
            GSA_HIP(hipMemcpyAsync(idx_out, minor_out, sizeof(I) * nnz, hipMemcpyDeviceToDevice, s));
            GSA_HIP(hipMemcpyAsync(val_out, val, sizeof(T) * nnz, hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(k_ptr_from_major<I>, dim3(grid_for(nnz + 1, 256)), dim3(256), 0, s,
                               nnz, major, major_dim, ptr_out);
            GSA_HIP(hipGetLastError());
            return GSA_OK;
        }
        if (!(flags & 2)) kminor = nullptr;
    }
    const uint64_t maxkey = kminor
        ? (uint64_t)(major_dim - 1) * (uint64_t)minor_dim + (uint64_t)(minor_dim - 1)
        : (uint64_t)(major_dim - 1);
    const int bits = bits_for(maxkey);
    const uint64_t kdim = kminor ? (uint64_t)minor_dim : 1;
    // Entries that are not in order move far from where they were (a
    // transposition, or storage in another order): each entry's (minor, value)
    // pair rides through the radix passes with its key, so no pass reads by a
    // permutation.
    auto run = [&](auto key_t) -> int {
        using K = decltype(key_t);
        Scratch keys(s);
        GSA_TRY(keys.alloc(sizeof(K) * nnz));
        const KeyFromIdx<K, I> key0{major, kminor, major_dim, (int64_t)kdim};
        GSA_TRY((radix_sort_payload<K, T, I>(s, nnz, bits, key0, minor_out, val, keys.as<K>(),
                                             idx_out, val_out, minor_ptr, minor_dim)));
        hipLaunchKernelGGL((k_ptr_from_keys<I, K>), dim3(grid_for(nnz + 1, 256)), dim3(256), 0, s,
                           nnz, (const K*)keys.as<K>(), kdim, major_dim, ptr_out);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    };
    if (bits <= 32) return run(uint32_t{});
    return run(uint64_t{});
}

// ---- row views of COO / CSC storage (sorted-key SpMV) -------------------------
template <typename I>
__global__ void k_unsorted(int64_t nnz, const I* __restrict__ major, int* __restrict__ flag) {
    for (int64_t i = 1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz;
         i += (int64_t)gridDim.x * blockDim.x)
        if (major[i] < major[i - 1]) *flag = 1;
}


template <typename I, typename V>
__global__ void k_gather_idx(int64_t nnz, const V* __restrict__ perm, const I* __restrict__ in,
                             I* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; base < nnz;
         base += stride * GATHER_U) {
        I v[GATHER_U];
#pragma unroll
        for (int u = 0; u < GATHER_U; ++u) {
            const int64_t i = base + u * stride;
            if (i < nnz) v[u] = in[(int64_t)__builtin_nontemporal_load(perm + i)];
        }
#pragma unroll
        for (int u = 0; u < GATHER_U; ++u) {
            const int64_t i = base + u * stride;
            if (i < nnz) out[i] = v[u];
        }
    }
}

int indices_sorted(hipStream_t s, int itype, int64_t nnz, const void* idx, int* sorted) {
    *sorted = 1;
    if (nnz <= 1) return GSA_OK;
    return dispatch_i(itype, [&](auto ti) -> int {
        using I = decltype(ti);
        Scratch flag(s);
        GSA_TRY(flag.alloc(sizeof(int)));
        GSA_TRY(zero_fill(s, flag.p, sizeof(int)));
        hipLaunchKernelGGL(k_unsorted<I>, dim3(grid_for(nnz, 256)), dim3(256), 0, s, nnz,
                           (const I*)idx, flag.as<int>());
        GSA_HIP(hipGetLastError());
        int unsorted = 0;
        GSA_HIP(hipMemcpyAsync(&unsorted, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
        GSA_HIP(hipStreamSynchronize(s));
        *sorted = !unsorted;
        return GSA_OK;
    });
}

void release_row_view(RowView* v) {
    if (!v) return;
    if (v->win) gsa_plan_destroy(v->win);
    for (void* q : {v->rowptr, v->col, v->perm})
        if (q) (void)hipFree(q);
    delete v;
}

int build_row_view(hipStream_t s, int itype, int64_t m, int64_t nnz, const void* rowidx,
                   const void* colidx, bool col_is_temp, RowView& v) {
    v.itype = itype;
    v.m = m;
    v.nnz = nnz;
    return dispatch_i(itype, [&](auto ti) -> int {
        using I = decltype(ti);
        GSA_HIP(hipMalloc(&v.rowptr, sizeof(I) * (m + 1)));
        if (nnz == 0) {
            GSA_TRY(zero_fill(s, v.rowptr, sizeof(I) * (m + 1)));
            return counts_to_ptr<I>(s, (I*)v.rowptr, m, nullptr);
        }
        GSA_CHECK(m > 0, GSA_ERR_INVALID, "stored entries in a matrix with no rows");
        const unsigned g = grid_for(nnz, 256);
        Scratch flag(s);
        GSA_TRY(flag.alloc(sizeof(int)));
        GSA_TRY(zero_fill(s, flag.p, sizeof(int)));
        hipLaunchKernelGGL(k_unsorted<I>, dim3(g), dim3(256), 0, s, nnz, (const I*)rowidx,
                           flag.as<int>());
        int unsorted = 0;
        GSA_HIP(hipMemcpyAsync(&unsorted, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
        GSA_HIP(hipStreamSynchronize(s));
        if (!unsorted) {  // storage already in row order: pointer only
            hipLaunchKernelGGL(k_ptr_from_major<I>, dim3(grid_for(nnz + 1, 256)), dim3(256), 0, s,
                               nnz, (const I*)rowidx, m, (I*)v.rowptr);
            if (col_is_temp) {
                GSA_HIP(hipMalloc(&v.col, sizeof(I) * nnz));
                GSA_HIP(hipMemcpyAsync(v.col, colidx, sizeof(I) * nnz, hipMemcpyDeviceToDevice, s));
            }
            GSA_HIP(hipGetLastError());
            return GSA_OK;
        }
        // stable sort of the storage positions by row
        const int bits = bits_for((uint64_t)(m - 1));
        auto sorted = [&](auto key_t, auto pos_t, auto sorter) -> int {
            using K = decltype(key_t);
            using V = decltype(pos_t);
            Scratch keys(s);
            GSA_TRY(keys.alloc(sizeof(K) * nnz));
            GSA_HIP(hipMalloc(&v.perm, sizeof(V) * nnz));
            v.perm64 = sizeof(V) == 8;
            hipLaunchKernelGGL((k_make_keys<I, K, V>), dim3(g), dim3(256), 0, s, nnz, (const I*)rowidx,
                               (const I*)nullptr, (int64_t)1, m, keys.as<K>(), (V*)v.perm);
            GSA_TRY(sorter(keys.as<K>(), (V*)v.perm));
            GSA_HIP(hipMalloc(&v.col, sizeof(I) * nnz));
            hipLaunchKernelGGL((k_gather_idx<I, V>), dim3(g), dim3(256), 0, s, nnz,
                               (const V*)v.perm, (const I*)colidx, (I*)v.col);
            hipLaunchKernelGGL((k_ptr_from_keys<I, K>), dim3(grid_for(nnz + 1, 256)), dim3(256), 0,
                               s, nnz, (const K*)keys.as<K>(), (uint64_t)1, m, (I*)v.rowptr);
            GSA_HIP(hipGetLastError());
            GSA_HIP(hipStreamSynchronize(s));  // the keys scratch is released on return
            return GSA_OK;
        };
        if (bits <= 32 && nnz < (int64_t)UINT32_MAX)
            return sorted(uint32_t{}, uint32_t{}, [&](uint32_t* k, uint32_t* p) {
                return radix_sort_pairs_u32_u32(s, k, p, nnz, bits);
            });
        if (nnz < (int64_t)UINT32_MAX)
            return sorted(uint64_t{}, uint32_t{}, [&](uint64_t* k, uint32_t* p) {
                return radix_sort_pairs_u64_u32(s, k, p, nnz, bits);
            });
        return sorted(uint64_t{}, uint64_t{}, [&](uint64_t* k, uint64_t* p) {
            return radix_sort_pairs_u64(s, k, p, nnz, bits);
        });
    });
}

// out[i] = in[i] + delta in the output index type (host ingest and egress:
// zero-based host arrays to one-based device arrays and back, widened or
// narrowed between Int32 and Int64).  A value outside the output type's range
// raises *bad (left untouched otherwise).  in and out may be the same array.
template <typename A, typename B>
__global__ void k_index_shift(int64_t n, const A* in, B* out, int64_t delta, int* bad) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t v = (int64_t)in[i] + delta;
        if constexpr (sizeof(B) < sizeof(int64_t))
            if (v < (int64_t)INT32_MIN || v > (int64_t)INT32_MAX) *bad = 1;
        out[i] = (B)v;
    }
}

}  // namespace gsa

using namespace gsa;

extern "C" {

int gsa_index_shift(void* stream, int itype_in, int itype_out, int64_t n, const void* in,
                    void* out, int64_t delta) {
    GSA_TRY(ensure_init());
    GSA_CHECK(n >= 0 && (n == 0 || (in && out)), GSA_ERR_INVALID, "gsa_index_shift: bad arguments");
    GSA_CHECK((itype_in == GSA_I32 || itype_in == GSA_I64) &&
              (itype_out == GSA_I32 || itype_out == GSA_I64), GSA_ERR_UNSUPPORTED,
              "gsa_index_shift: bad index types %d, %d", itype_in, itype_out);
    if (n == 0) return GSA_OK;
    hipStream_t s = as_stream(stream);
    const unsigned grid = grid_for(n, 256);
    if (itype_out == GSA_I64) {
        if (itype_in == GSA_I32)
            hipLaunchKernelGGL((k_index_shift<int32_t, int64_t>), dim3(grid), dim3(256), 0, s, n,
                               (const int32_t*)in, (int64_t*)out, delta, (int*)nullptr);
        else
            hipLaunchKernelGGL((k_index_shift<int64_t, int64_t>), dim3(grid), dim3(256), 0, s, n,
                               (const int64_t*)in, (int64_t*)out, delta, (int*)nullptr);
        GSA_HIP(hipGetLastError());
        return GSA_OK;
    }
    // Int32 output: range-checked, one read-back
    Scratch flag(s);
    GSA_TRY(flag.alloc(sizeof(int)));
    GSA_TRY(zero_fill(s, flag.p, sizeof(int)));
    if (itype_in == GSA_I32)
        hipLaunchKernelGGL((k_index_shift<int32_t, int32_t>), dim3(grid), dim3(256), 0, s, n,
                           (const int32_t*)in, (int32_t*)out, delta, flag.as<int>());
    else
        hipLaunchKernelGGL((k_index_shift<int64_t, int32_t>), dim3(grid), dim3(256), 0, s, n,
                           (const int64_t*)in, (int32_t*)out, delta, flag.as<int>());
    GSA_HIP(hipGetLastError());
    int bad = 0;
    GSA_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    GSA_HIP(hipStreamSynchronize(s));
    GSA_CHECK(!bad, GSA_ERR_INVALID, "gsa_index_shift: an index does not fit Int32");
    return GSA_OK;
}

int gsa_csr_to_coo(void* stream, int vtype, int itype, int64_t m, int64_t nnz, const void* rowptr,
                   const void* colval, const void* nzval, void* rowind, void* colind,
                   void* nzval_out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(m >= 0 && nnz >= 0, GSA_ERR_INVALID, "gsa_csr_to_coo: negative size");
    hipStream_t s = as_stream(stream);
    return dispatch_vbi(vtype, itype, [&](auto tv, auto ti) {  // values moved as bits
        using T = decltype(tv);
        using I = decltype(ti);
        if (nnz == 0) return (int)GSA_OK;
        if (m == 0) {  // no row claims an entry: the arrays are copied as they are
            GSA_HIP(hipMemcpyAsync(colind, colval, sizeof(I) * nnz, hipMemcpyDeviceToDevice, s));
            GSA_HIP(hipMemcpyAsync(nzval_out, nzval, sizeof(T) * nnz, hipMemcpyDeviceToDevice, s));
            return (int)GSA_OK;
        }
        const int64_t nblocks = nnz / SPMV_NB + 1;
        Scratch plan(s);
        GSA_TRY(plan.alloc(sizeof(int64_t) * (nblocks + 1)));
        GSA_TRY(build_csr_plan(s, itype, m, nnz, rowptr, plan.as<int64_t>(), nblocks));
        hipLaunchKernelGGL((k_csr_to_coo<T, I>), dim3((unsigned)nblocks), dim3(ROW_WG), 0, s, m, nnz,
                           (const I*)rowptr, (const int64_t*)plan.as<int64_t>(), (const I*)colval,
                           (const T*)nzval, (I*)rowind, (I*)colind, (T*)nzval_out);
        GSA_HIP(hipGetLastError());
        return (int)GSA_OK;
    });
}

int gsa_coo_to_csr(void* stream, int vtype, int itype, int64_t m, int64_t n, int64_t nnz,
                   const void* rowind, const void* colind, const void* nzval, void* rowptr,
                   void* colval, void* nzval_out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(m >= 0 && n >= 0 && nnz >= 0, GSA_ERR_INVALID, "gsa_coo_to_csr: negative size");
    GSA_CHECK(m == 0 || n == 0 || (uint64_t)m <= UINT64_MAX / (uint64_t)n, GSA_ERR_INVALID,
              "gsa_coo_to_csr: m*n overflows 64-bit keys");
    hipStream_t s = as_stream(stream);
    return dispatch_vbi(vtype, itype, [&](auto tv, auto ti) {  // values moved as bits
        using T = decltype(tv);
        using I = decltype(ti);
        return sort_to_compressed<T, I>(s, m, n, nnz, (const I*)rowind, (const I*)colind,
                                        (const I*)colind, (const T*)nzval, (I*)rowptr, (I*)colval,
                                        (T*)nzval_out);
    });
}

int gsa_csr_to_csc(void* stream, int vtype, int itype, int64_t m, int64_t n, int64_t nnz,
                   const void* rowptr, const void* colval, const void* nzval, void* colptr,
                   void* rowval, void* nzval_out) {
    GSA_TRY(ensure_init());
    GSA_CHECK(m >= 0 && n >= 0 && nnz >= 0, GSA_ERR_INVALID, "gsa_csr_to_csc: negative size");
    GSA_CHECK(nnz == 0 || (m > 0 && n > 0), GSA_ERR_INVALID,
              "gsa_csr_to_csc: stored entries in a matrix with a zero dimension");
    hipStream_t s = as_stream(stream);
    return dispatch_vbi(vtype, itype, [&](auto tv, auto ti) {  // values moved as bits
        using T = decltype(tv);
        using I = decltype(ti);
        // stable sort by column only, each entry's (row, value) carried; the first
        // pass finds the rows from the row pointer tile by tile (below 2^31 rows),
        // otherwise they are expanded first
        Scratch rows(s);
        const bool from_ptr = m < ((int64_t)1 << 31);
        if (nnz > 0 && !from_ptr) {
            GSA_TRY(rows.alloc(sizeof(I) * nnz));
            GSA_TRY(expand_rows(s, itype, m, nnz, rowptr, rows.p));
        }
        return sort_to_compressed<T, I>(s, n, m, nnz, (const I*)colval, nullptr,
                                        (const I*)rows.p, (const T*)nzval, (I*)colptr,
                                        (I*)rowval, (T*)nzval_out,
                                        from_ptr ? (const I*)rowptr : nullptr);
    });
}

}  // extern "C"
