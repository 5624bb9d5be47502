// Stable LSD radix sort of (key, FP64 value) pairs on the device.
#pragma once
#include "common.hpp"

namespace spws {

constexpr int kRadixBlock = 256;
constexpr int kRadixItems = 8;
constexpr int kRadixTile = kRadixBlock * kRadixItems;  // keys per workgroup tile
constexpr int kRadixDigit = 8;                          // bits per pass
constexpr int kRadixBuckets = 1 << kRadixDigit;

// Sorts keys[0, n) by bits [begin_bit, end_bit) carrying vals along, using
// keys_alt / vals_alt as the ping-pong buffers.  On return *in_alt says
// whether the sorted data ended up in the _alt buffers.  K is uint32_t or
// uint64_t.  Uses ctx->radix_hist and the scan scratch.
template <class K>
int radix_sort_pairs(spws_ctx* ctx, K* keys, double* vals, K* keys_alt, double* vals_alt, int64_t n, int begin_bit,
                     int end_bit, bool* in_alt);

}  // namespace spws
