// gsa_scan.h — device prefix-sum and index helpers shared by the structural ops.
#pragma once
#include "gsa_common.h"

namespace gsa {

// out[i] = add + in[0] + ... + in[i]  (out may alias in)
template <typename T>
int inclusive_scan(hipStream_t s, const T* in, T* out, int64_t n, T add);

// Zero a device buffer with a fill kernel (graph-capture safe, unlike a
// memset node).
int zero_fill(hipStream_t s, void* p, size_t bytes);

// ptr[1..m] hold per-row counts on entry; on exit ptr is a one-based pointer
// array (ptr[0] = 1, ptr[m] = total + 1).  If total != nullptr the stream is
// synchronised and the total returned.
template <typename I>
int counts_to_ptr(hipStream_t s, I* ptr, int64_t m, int64_t* total);

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the
// unsigned 32- or 64-bit keys; values are 32- or 64-bit payloads.  keys/vals are
// sorted in place (double-buffered through scratch).
int radix_sort_pairs_u64(hipStream_t s, uint64_t* keys, uint64_t* vals, int64_t n, int bits);
int radix_sort_pairs_u64_u32(hipStream_t s, uint64_t* keys, uint32_t* vals, int64_t n, int bits);
int radix_sort_pairs_u32_u32(hipStream_t s, uint32_t* keys, uint32_t* vals, int64_t n, int bits);

inline int bits_for(uint64_t maxval) {
    int b = 0;
    while (b < 64 && (maxval >> b) != 0) ++b;
    return b;
}

}  // namespace gsa
