// Stable LSD radix sort of (key, FP64 value) pairs on the device.
#pragma once
#include "common.hpp"

namespace spws {

// Tiles of the LSD method's own run and reduce kernels (spgemm_outer.hip).
constexpr int kRadixBlock = 256;
constexpr int kRadixItems = 8;
constexpr int kRadixTile = kRadixBlock * kRadixItems;  // keys per workgroup tile

// The sort: tiles of kSortTile keys, passes of at most kSortDigitMax bits
// (the key bits are split evenly over ceil(bits / kSortDigitMax) passes).
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kSortDigitMax = 9;
constexpr int kSortBucketsMax = 1 << kSortDigitMax;

// int32 histogram entries one pass of radix_sort_pairs over n keys uses at
// most (the scan of a pass covers as many).
inline size_t radix_hist_entries(int64_t n) {
  return (size_t)div_up(n > 0 ? n : 1, kSortTile) * kSortBucketsMax;
}

// Sorts keys[0, n) by bits [begin_bit, end_bit) carrying vals along, using
// keys_alt / vals_alt as the ping-pong buffers.  On return *in_alt says
// whether the sorted data ended up in the _alt buffers.  K is uint32_t or
// uint64_t.  Uses ctx->radix_hist and the scan scratch.  With vals_first,
// the values are read from there instead of vals (which then needs no
// copy of them; it is still filled when no pass runs).
template <class K>
int radix_sort_pairs(spws_ctx* ctx, K* keys, double* vals, K* keys_alt, double* vals_alt, int64_t n, int begin_bit,
                     int end_bit, bool* in_alt, const double* vals_first = nullptr);

}  // namespace spws
