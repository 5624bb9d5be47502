// Stable LSD radix sort of (key, FP64 value) pairs, 8 bits per pass.
//
// Per pass, three steps over tiles of kRadixTile keys:
//   1. k_radix_hist     per-tile digit histogram (LDS atomics), written
//                       digit-major: hist[d * ntiles + t].
//   2. exclusive scan   of the digit-major histogram -> global offset of
//                       (digit d, tile t).
//   3. k_radix_scatter  stable tile-local ranks by wavefront multisplit:
//                       each wave walks its contiguous quarter of the tile
//                       64 keys at a time; 8 ballots give the lanes holding
//                       the same digit, the group's rank is the running
//                       per-wave digit count (LDS) plus the lanes before it.
//                       Keys and values are then staged in LDS in digit
//                       order and written out so that consecutive threads
//                       store consecutive addresses of each digit's run.
#include "radix.hpp"

namespace spws {

template <class K>
__device__ __forceinline__ int radix_digit(K key, int shift, uint32_t mask) {
  return (int)((key >> shift) & (K)mask);
}

template <class K>
__global__ __launch_bounds__(kRadixBlock) void k_radix_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                            uint32_t mask, int32_t* __restrict__ hist, int64_t ntiles) {
  __shared__ int32_t h[kRadixBuckets];
  for (int d = threadIdx.x; d < kRadixBuckets; d += kRadixBlock) h[d] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
#pragma unroll
  for (int it = 0; it < kRadixItems; ++it) {
    const int64_t idx = base + (int64_t)it * kRadixBlock + threadIdx.x;
    if (idx < n) atomicAdd(&h[radix_digit(keys[idx], shift, mask)], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadixBuckets; d += kRadixBlock) hist[(int64_t)d * ntiles + blockIdx.x] = h[d];
}

template <class K>
__global__ __launch_bounds__(kRadixBlock) void k_radix_scatter(const K* __restrict__ kin, const double* __restrict__ vin,
                                                               K* __restrict__ kout, double* __restrict__ vout,
                                                               int64_t n, int shift, uint32_t mask,
                                                               const int32_t* __restrict__ offs, int64_t ntiles) {
  constexpr int kWaves = kRadixBlock / kWave;
  constexpr int kSub = kRadixTile / kWaves;  // keys per wave
  static_assert(kRadixBlock == kRadixBuckets, "one thread per digit in the prefix step");
  __shared__ K sk[kRadixTile];
  __shared__ double sv[kRadixTile];
  __shared__ int32_t wcnt[kWaves][kRadixBuckets];
  __shared__ int32_t dstart[kRadixBuckets];
  __shared__ int32_t scan_sh[kRadixBlock / kWave + 1];

  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  for (int d = threadIdx.x; d < kWaves * kRadixBuckets; d += kRadixBlock) (&wcnt[0][0])[d] = 0;
  __syncthreads();

  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
  const int64_t wbase = base + (int64_t)w * kSub;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  K key[kRadixItems];
  double val[kRadixItems];
  int32_t rank[kRadixItems];
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const int64_t idx = wbase + (int64_t)r * kWave + lane;
    const bool valid = idx < n;
    key[r] = valid ? kin[idx] : (K)0;
    val[r] = valid ? vin[idx] : 0.0;
    const int d = radix_digit(key[r], shift, mask);
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixDigit; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t m = __ballot(bit);
      match &= bit ? m : ~m;
    }
    rank[r] = 0;
    if (valid) {
      const int before = __popcll(match & lt);
      const int32_t old = wcnt[w][d];
      rank[r] = old + before;
      if (before == 0) wcnt[w][d] = old + __popcll(match);
    }
  }
  __syncthreads();
  {
    // thread d: exclusive prefix of digit d over the waves, tile total
    const int d = threadIdx.x;
    int32_t run = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
      const int32_t c = wcnt[q][d];
      wcnt[q][d] = run;
      run += c;
    }
    int32_t total;
    const int32_t st = block_excl_scan<int32_t, kRadixBlock>(run, scan_sh, &total);
    dstart[d] = st;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRadixItems; ++r) {
    const int64_t idx = wbase + (int64_t)r * kWave + lane;
    if (idx < n) {
      const int d = radix_digit(key[r], shift, mask);
      const int pos = dstart[d] + wcnt[w][d] + rank[r];
      sk[pos] = key[r];
      sv[pos] = val[r];
    }
  }
  __syncthreads();
  const int64_t remain = n - base;
  const int tile_n = remain < kRadixTile ? (int)remain : kRadixTile;
  for (int i = threadIdx.x; i < tile_n; i += kRadixBlock) {
    const K k = sk[i];
    const int d = radix_digit(k, shift, mask);
    const int64_t g = (int64_t)offs[(int64_t)d * ntiles + blockIdx.x] + (i - dstart[d]);
    kout[g] = k;
    vout[g] = sv[i];
  }
}

template <class K>
int radix_sort_pairs(spws_ctx* ctx, K* keys, double* vals, K* keys_alt, double* vals_alt, int64_t n, int begin_bit,
                     int end_bit, bool* in_alt) {
  *in_alt = false;
  if (n <= 1 || end_bit <= begin_bit) return SPWS_OK;
  SPWS_REQUIRE(n < ((int64_t)1 << 31), SPWS_ERR_UNSUPPORTED, "radix sort of %lld keys exceeds int32 offsets",
               (long long)n);
  const int64_t ntiles = div_up(n, kRadixTile);
  const int64_t hn = ntiles * kRadixBuckets;
  int rc = ctx->radix_hist.ensure((size_t)(2 * hn + 1) * sizeof(int32_t));
  if (rc) return rc;
  int32_t* hist = ctx->radix_hist.as<int32_t>();
  int32_t* offs = hist + hn;
  K* kin = keys;
  double* vin = vals;
  K* kout = keys_alt;
  double* vout = vals_alt;
  for (int shift = begin_bit; shift < end_bit; shift += kRadixDigit) {
    const int nb = end_bit - shift < kRadixDigit ? end_bit - shift : kRadixDigit;
    const uint32_t mask = (1u << nb) - 1u;
    SPWS_LAUNCH(ctx, "radix_hist", (k_radix_hist<K>), dim3((unsigned)ntiles), dim3(kRadixBlock), 0, kin, n, shift,
                mask, hist, ntiles);
    rc = exclusive_scan<int32_t, int32_t>(ctx, hist, offs, hn);
    if (rc) return rc;
    SPWS_LAUNCH(ctx, "radix_scatter", (k_radix_scatter<K>), dim3((unsigned)ntiles), dim3(kRadixBlock), 0, kin, vin,
                kout, vout, n, shift, mask, offs, ntiles);
    K* tk = kin;
    kin = kout;
    kout = tk;
    double* tv = vin;
    vin = vout;
    vout = tv;
    *in_alt = !*in_alt;
  }
  return SPWS_OK;
}

template int radix_sort_pairs<uint32_t>(spws_ctx*, uint32_t*, double*, uint32_t*, double*, int64_t, int, int, bool*);
template int radix_sort_pairs<uint64_t>(spws_ctx*, uint64_t*, double*, uint64_t*, double*, int64_t, int, int, bool*);

}  // namespace spws
