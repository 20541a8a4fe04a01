// gsa_common.h — shared host/device helpers of libgsa_hip.so (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <type_traits>
#include <cstring>
#include "../../include/gsa_hip.h"

namespace gsa {

// ------------------------------------------------------------ value types --
// Complex values in Julia's layout (re, im).  Arithmetic is Julia's
// (base/complex.jl): sums and differences by parts; the product
// (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re), every operation rounded on
// its own (the library is built without FMA contraction).
template <typename R>
struct cplx {
    R re, im;
    cplx() = default;
    __host__ __device__ constexpr cplx(R r, R i) : re(r), im(i) {}
    __host__ __device__ constexpr explicit cplx(int z) : re((R)z), im((R)0) {}
};
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) {
    return cplx<R>(a.re + b.re, a.im + b.im);
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) {
    return cplx<R>(a.re - b.re, a.im - b.im);
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator-(cplx<R> a) {
    return cplx<R>(-a.re, -a.im);
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator*(cplx<R> a, cplx<R> b) {
    const R rr = a.re * b.re, ii = a.im * b.im, ri = a.re * b.im, ir = a.im * b.re;
    return cplx<R>(rr - ii, ri + ir);
}
template <typename R>
__host__ __device__ __forceinline__ bool operator==(cplx<R> a, cplx<R> b) {
    return a.re == b.re && a.im == b.im;
}
template <typename R>
__host__ __device__ __forceinline__ bool operator!=(cplx<R> a, cplx<R> b) {
    return !(a == b);
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R>& operator+=(cplx<R>& a, cplx<R> b) {
    a = a + b;
    return a;
}

// Traits of the arithmetic types: double, float, cplx<double>, cplx<float>, and
// the integer value types, held as uint64_t / uint32_t so that sums and
// products wrap modulo 2^64 / 2^32 exactly as Julia's Int64 / Int32 do (the
// bit patterns of two's-complement and unsigned arithmetic agree).
template <typename T>
struct VT {
    static constexpr bool cx = false, f32 = false, integer = false;
};
template <>
struct VT<float> {
    static constexpr bool cx = false, f32 = true, integer = false;
};
template <typename R>
struct VT<cplx<R>> {
    static constexpr bool cx = true, f32 = false, integer = false;
};
template <>
struct VT<uint32_t> {
    static constexpr bool cx = false, f32 = false, integer = true;
};
template <>
struct VT<uint64_t> {
    static constexpr bool cx = false, f32 = false, integer = true;
};

template <typename T>
__host__ __device__ __forceinline__ T conj_v(T v) {
    if constexpr (VT<T>::cx) return T(v.re, -v.im);
    else return v;
}
template <typename T>
__host__ __device__ __forceinline__ T conj_if(T v, bool c) {
    if constexpr (VT<T>::cx) return c ? conj_v(v) : v;
    else return v;
}

// v * s for a scalar s (alpha, beta): a complex scalar whose imaginary part is
// zero multiplies the parts separately, as Julia's Complex * Real does (so 1
// leaves every value, Inf and NaN included, unchanged).
template <typename T>
__host__ __device__ __forceinline__ T mul_scalar(T v, T s) {
    if constexpr (VT<T>::cx) {
        if (s.im == 0) return T(v.re * s.re, v.im * s.re);
        return v * s;
    } else {
        return v * s;
    }
}

// ---------------------------------------------------------------- errors ----
void set_error(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define GSA_HIP(call)                                                        \
    do {                                                                     \
        hipError_t _e = (call);                                              \
        if (_e != hipSuccess) return ::gsa::hip_fail(_e, #call);             \
    } while (0)

#define GSA_CHECK(cond, code, ...)                                           \
    do {                                                                     \
        if (!(cond)) { ::gsa::set_error((code), __VA_ARGS__); return (code); } \
    } while (0)

#define GSA_TRY(expr)                                                        \
    do { int _r = (expr); if (_r != GSA_OK) return _r; } while (0)

int ensure_init();

// Stream-ordered scratch memory: device blocks cached per (device, stream); a
// block released on a stream is reused only by later work on that stream.
int ws_alloc(void** p, size_t bytes, hipStream_t s);
void ws_free(void* p, hipStream_t s);
// Free the blocks cached for stream s (its work must have drained).
void ws_release_stream(hipStream_t s);
// Bytes cached for stream s on any device.
size_t ws_cached_bytes(hipStream_t s);
// Scratch blocks held by live graphs (taken while a stream was being captured).
int64_t ws_graph_blocks();

// RAII scratch buffer; frees on the stream it was taken on.
struct Scratch {
    void* p = nullptr;
    hipStream_t s = nullptr;
    Scratch() = default;
    explicit Scratch(hipStream_t st) : s(st) {}
    int alloc(size_t bytes) { return ws_alloc(&p, bytes, s); }
    template <typename U> U* as() const { return reinterpret_cast<U*>(p); }
    ~Scratch() { if (p) ws_free(p, s); }
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Positions per thread and pass in the gather kernels (reads by a stored
// permutation): all of them are issued before the first write.
constexpr int GATHER_U = 8;

// Tiles in XCD-contiguous order: workgroup b runs on XCD b % 8, so it takes
// tile (b % 8) * ceil(ntiles / 8) + b / 8, and each XCD works through one
// contiguous eighth of the tiles (writes of neighbouring tiles to adjacent
// addresses then meet in the same L2).  Launch xcd_tile_grid(ntiles)
// workgroups; a workgroup whose tile is >= ntiles has nothing to do.
inline int64_t xcd_tile_grid(int64_t ntiles) { return 8 * ((ntiles + 7) / 8); }
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t ntiles) {
    return (b % 8) * ((ntiles + 7) / 8) + b / 8;
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// The value held by lane (lane ^ j) of the wavefront, j in {1, 2, 4, 8, 16, 32},
// by data-parallel moves in the vector ALU instead of a trip through the LDS
// crossbar (the lanes involved must be active):
//   j = 1, 2: the quad permutations [1,0,3,2] and [2,3,0,1];
//   j = 4, 8: rotations inside each row of 16 lanes (row_ror:n gives lane i the
//             value of lane (i - n) mod 16): by 12 where bit 2 of the lane is
//             clear, by 4 where it is set; by 8 for every lane;
//   j = 16, 32: the swaps of row pairs / wavefront halves (permlane16_swap,
//             permlane32_swap of the value with itself: the first result holds
//             the lower partner's copy, the second the upper's).
// The same partners as __shfl_xor(v, j, 64), so sums built on it round alike.
__device__ __forceinline__ uint32_t lane_xor32(uint32_t v, int j) {
    const unsigned lane = __lane_id();
    if (j == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    if (j == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    if (j == 4) {
        const uint32_t up = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x12C, 0xF, 0xF, false);
        const uint32_t dn = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false);
        return (lane & 4u) ? dn : up;
    }
    if (j == 8) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
    if (j == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (lane & 16u) ? (uint32_t)p[0] : (uint32_t)p[1];
    }
    const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);  // j == 32
    return (lane & 32u) ? (uint32_t)p[0] : (uint32_t)p[1];
}
// Inclusive prefix sum of an int inside each group of G lanes (G = 16, 32, 64),
// by data-parallel moves: shifts by 1, 2, 4, 8 inside each row of 16 lanes
// (lanes before the row start add nothing), then lane 15 of each even row added
// to the next row (row_bcast:15, 32- and 64-lane groups) and lane 31 to rows 2
// and 3 (row_bcast:31, 64-lane groups).  Every lane of the wavefront must be
// active.  The same sums as the __shfl_up ladder.
template <int G>
__device__ __forceinline__ int group_incl_sum(int v) {
    static_assert(G == 16 || G == 32 || G == 64, "groups of 16, 32 or 64 lanes");
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
    if constexpr (G >= 32) v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);
    if constexpr (G == 64) v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);
    return v;
}

template <typename T>
__device__ __forceinline__ T lane_xor(T v, int j) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16,
                  "lane_xor moves 32-, 64- or 128-bit values");
    if constexpr (VT<T>::cx) {
        return T(lane_xor(v.re, j), lane_xor(v.im, j));
    } else if constexpr (sizeof(T) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = lane_xor32((uint32_t)u, j);
        const uint64_t hi = lane_xor32((uint32_t)(u >> 32), j);
        return __builtin_bit_cast(T, (hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, lane_xor32(__builtin_bit_cast(uint32_t, v), j));
    }
}

// Shuffles of any value type (complex values move part by part).
template <typename T>
__device__ __forceinline__ T shfl_v(T v, int src) {
    if constexpr (VT<T>::cx) return T(shfl_v(v.re, src), shfl_v(v.im, src));
    else return __shfl(v, src, 64);
}
template <typename T>
__device__ __forceinline__ T shfl_up_v(T v, int o) {
    if constexpr (VT<T>::cx) return T(shfl_up_v(v.re, o), shfl_up_v(v.im, o));
    else return __shfl_up(v, o, 64);
}
template <typename T>
__device__ __forceinline__ T shfl_down_v(T v, int o) {
    if constexpr (VT<T>::cx) return T(shfl_down_v(v.re, o), shfl_down_v(v.im, o));
    else return __shfl_down(v, o, 64);
}

// Atomic add of any value type (complex: one atomic per part; integers wrap).
template <typename T>
__device__ __forceinline__ void atomic_add_v(T* p, T v) {
    if constexpr (VT<T>::cx) {
        auto* q = reinterpret_cast<decltype(v.re)*>(p);
        unsafeAtomicAdd(q, v.re);
        unsafeAtomicAdd(q + 1, v.im);
    } else if constexpr (std::is_same_v<T, uint64_t>) {
        atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
    } else if constexpr (std::is_same_v<T, uint32_t>) {
        atomicAdd(reinterpret_cast<unsigned int*>(p), (unsigned int)v);
    } else {
        unsafeAtomicAdd(p, v);
    }
}

// Grid size for grid-stride element loops: enough waves to fill 256 CUs.
inline unsigned grid_for(int64_t n, int block) {
    int64_t g = cdiv(n > 0 ? n : 1, block);
    if (g > 256 * 32) g = 256 * 32;
    return (unsigned)g;
}

// ------------------------------------------------------- type dispatch ----
template <typename F>
int dispatch_v(int vtype, F&& f) {
    switch (vtype) {
        case GSA_F64: return f(double{});
        case GSA_F32: return f(float{});
        default: set_error(GSA_ERR_UNSUPPORTED, "unsupported value type %d", vtype);
                 return GSA_ERR_UNSUPPORTED;
    }
}
template <typename F>
int dispatch_i(int itype, F&& f) {
    switch (itype) {
        case GSA_I64: return f(int64_t{});
        case GSA_I32: return f(int32_t{});
        default: set_error(GSA_ERR_UNSUPPORTED, "unsupported index type %d", itype);
                 return GSA_ERR_UNSUPPORTED;
    }
}
template <typename F>
int dispatch_vi(int vtype, int itype, F&& f) {
    return dispatch_v(vtype, [&](auto tv) {
        return dispatch_i(itype, [&](auto ti) { return f(tv, ti); });
    });
}
// Every arithmetic value type: Float64, Float32, ComplexF64, ComplexF32, Int64,
// Int32 (the integers as wrapping unsigned arithmetic of the same width).
template <typename F>
int dispatch_va(int vtype, F&& f) {
    switch (vtype) {
        case GSA_F64: return f(double{});
        case GSA_F32: return f(float{});
        case GSA_C128: return f(cplx<double>{});
        case GSA_C64: return f(cplx<float>{});
        case GSA_VI64: return f(uint64_t{});
        case GSA_VI32: return f(uint32_t{});
        default: set_error(GSA_ERR_UNSUPPORTED, "unsupported value type %d", vtype);
                 return GSA_ERR_UNSUPPORTED;
    }
}
template <typename F>
int dispatch_vai(int vtype, int itype, F&& f) {
    return dispatch_va(vtype, [&](auto tv) {
        return dispatch_i(itype, [&](auto ti) { return f(tv, ti); });
    });
}
// Values that are only moved (conversions): a bit pattern of the value's width.
struct alignas(16) bits128 {
    uint64_t lo, hi;
};
template <typename F>
int dispatch_vbits(int vtype, F&& f) {
    switch (vtype) {
        case GSA_F64: case GSA_C64: case GSA_VI64: return f(uint64_t{});
        case GSA_F32: case GSA_VI32: return f(uint32_t{});
        case GSA_C128: return f(bits128{});
        default: set_error(GSA_ERR_UNSUPPORTED, "unsupported value type %d", vtype);
                 return GSA_ERR_UNSUPPORTED;
    }
}
template <typename F>
int dispatch_vbi(int vtype, int itype, F&& f) {
    return dispatch_vbits(vtype, [&](auto tv) {
        return dispatch_i(itype, [&](auto ti) { return f(tv, ti); });
    });
}
inline bool vtype_real(int vtype) { return vtype == GSA_F64 || vtype == GSA_F32; }
inline size_t vtype_bytes(int vtype) {
    switch (vtype) {
        case GSA_F64: case GSA_C64: case GSA_VI64: return 8;
        case GSA_F32: case GSA_VI32: return 4;
        case GSA_C128: return 16;
        case GSA_BF16: return 2;
        default: return 0;
    }
}
// A host scalar of the value type (NULL: `def`).
template <typename T>
inline T host_scalar(const void* p, T def) {
    if (!p) return def;
    T v;
    memcpy(&v, p, sizeof(T));
    return v;
}
// beta mode: 0 = fill with zero, 1 = keep C, 2 = scale C.
template <typename T>
inline int beta_mode_of(T beta) {
    if (beta == T(0)) return 0;
    if (beta == T(1)) return 1;
    return 2;
}

// ------------------------------------------------------- plan object ----
// Row-block ownership: block b owns the rows whose first stored entry (0-based)
// lies in [b*NB, (b+1)*NB).  blk_row[b] is the first owned row, blk_row[nblocks]
// = m.  Built on the device in O(m).
constexpr int SPMV_NB = 2048;
constexpr int SPMV_CAP = SPMV_NB + 256;  // entries a workgroup stages in LDS

// Per-window staging summary: P0 (first owned entry), Ps (staged end), Pend
// (end of the last owned row) and a spill flag (last row read from HBM).
constexpr int PLAN_POS = 4;

// SpMV windows: wavefront w reduces entries [w*SPMV_WIN, (w+1)*SPMV_WIN) and
// owns the rows that start there (owner table built like the row blocks).
constexpr int SPMV_WIN = 576;
// Entries staged past a window's end, so a short row (at most SPMV_WX
// entries) that crosses it is finished, left to right, by its owner.
constexpr int SPMV_WX = 32;

// A plan is read-only once built: products keep their in-flight state (the
// SpMV's head/tail partial sums, the bins' staged products) in scratch of the
// call, so one plan may serve concurrent calls on several streams.
struct CsrPlanData {
    int itype = 0;
    int64_t m = 0, nnz = 0;
    int64_t nblocks = 0;          // workgroup windows (SPMV_NB)
    int64_t* blk_row = nullptr;   // device, nblocks+1
    int64_t* blk_pos = nullptr;   // device, PLAN_POS*nblocks
    int64_t nwin = 0;             // SpMV windows (SPMV_WIN entries each)
    int64_t* win_row = nullptr;   // device, nwin+1: first row starting in each window
    // Optional launch order of the SpMV windows (gsa_plan_set_columns): slot
    // 4*b + wave of workgroup b holds a window index or -1.
    int64_t* order = nullptr;
    int64_t nslot = 0;
    const void* order_col = nullptr;  // the column array the order was built from
    // With Int64 indices and n < 2^32: a 32-bit copy of those columns (built
    // with the order; the SpMV streams 4 bytes less per stored entry)
    uint32_t* col32 = nullptr;
    const void* ptr = nullptr;    // the row pointer the plan was built on
};

// Row ownership (and, if blk_pos is given, the staging summary) of windows of
// `nb` entries with staging capacity `cap`.
int build_csr_plan(hipStream_t s, int itype, int64_t m, int64_t nnz, const void* ptr,
                   int64_t* blk_row, int64_t nblocks, int64_t* blk_pos = nullptr,
                   int nb = SPMV_NB, int cap = SPMV_CAP);

}  // namespace gsa

namespace gsa {
void batch_plan_release(gsa_plan_s* plan);

// Row-ordered view of a COO or CSC matrix (the sorted-key form of its mul!):
// the entries in row order as a CSR row pointer, and, unless the storage is
// already in row order, each entry's column and storage position.
struct RowView {
    int itype = 0;
    int64_t m = 0, nnz = 0;
    void* rowptr = nullptr;    // I[m+1], one-based
    void* col = nullptr;       // I[nnz] one-based columns, or null: storage is row-ordered
    void* perm = nullptr;      // storage position of each entry (u32 or u64), or null
    bool perm64 = false;
    gsa_plan_s* win = nullptr; // SpMV windows over rowptr
    const void* src[2] = {nullptr, nullptr};  // the arrays it was built from (checked on use)
};

// Build v (m rows, nnz entries) from the one-based row and column of every
// stored entry, in storage order.  `col_is_temp`: colidx does not outlive the
// call, so a row-ordered storage still gets its own column array.
// Synchronises the stream.
int build_row_view(hipStream_t s, int itype, int64_t m, int64_t nnz, const void* rowidx,
                   const void* colidx, bool col_is_temp, RowView& v);
void release_row_view(RowView* v);
int expand_rows(hipStream_t s, int itype, int64_t m, int64_t nnz, const void* ptr, void* out);

// Row bins of a CSC or COO matrix for its scatter-shaped SpMV (y = A x):
// built from each entry's one-based row (rowidx) and, for CSC, the column
// pointer; null in *out when the matrix has too many rows for the bins.
// Synchronises.
struct BinPlan;
int build_bin_plan(hipStream_t s, int itype, int64_t m, int64_t n, int64_t nnz,
                   const void* rowidx, const void* colptr, BinPlan** out);
void release_bin_plan(BinPlan* p);
bool bin_plan_matches(const BinPlan* p, int itype, int64_t m, int64_t n, int64_t nnz,
                      const void* a, const void* b);
void bin_plan_set_sources(BinPlan* p, const void* a, const void* b);
int bin_spmv(hipStream_t s, const BinPlan& p, int vtype, const void* colidx,
             const void* nzval, const void* x, int64_t ldx, void* y, double alpha, double beta);
// 1 in *sorted when the one-based indices are non-decreasing.  Synchronises.
int indices_sorted(hipStream_t s, int itype, int64_t nnz, const void* idx, int* sorted);
}  // namespace gsa

struct gsa_plan_s {
    int kind = 0;  // 0 = CSR row blocks, 1 = batch schedule, 2 = COO row view only
    gsa::CsrPlanData csr;
    void* batch = nullptr;           // gsa::BatchPlan*
    gsa::RowView* view = nullptr;    // row view (CSC: beside the column blocks)
    gsa::BinPlan* bins = nullptr;    // row bins of a CSC or COO matrix (scatter-shaped SpMV)
    gsa::BinPlan* bins_t = nullptr;  // bins of the transposed product of a CSR or COO matrix
    // Gather views of scatter-shaped products with several right-hand sides: the
    // output rows of A*B (CSC, COO) or transpose(A)*B (CSR, COO) in row order, with
    // each entry's B row and storage position (gsa_plan_add_gather_view).
    gsa::RowView* gview = nullptr;
    gsa::RowView* gview_t = nullptr;
};

struct gsa_ctx_s {
    int kind = 0;
    hipStream_t stream = nullptr;
    int vtype = 0, itype = 0;
    int64_t m = 0, n = 0, nnz_out = 0;
    void* buf[8] = {nullptr};     // scratch owned by the context (freed with it)
    int64_t aux[8] = {0};
    const void* in[8] = {nullptr}; // caller arrays read again by the second phase
};
