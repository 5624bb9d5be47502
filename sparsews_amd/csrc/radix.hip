// Stable LSD radix sort of (key, FP64 value) pairs, up to 9 bits per pass.
//
// The key bits are split evenly over ceil(bits / 9) passes (34 bits: 9, 9,
// 8, 8).  Per pass, three steps over tiles of kSortTile keys:
//   1. k_radix_hist     per-tile digit histogram (LDS atomics), written
//                       digit-major: hist[d * ntiles + t].
//   2. exclusive scan   of the digit-major histogram -> global offset of
//                       (digit d, tile t).
//   3. k_radix_scatter  stable tile-local ranks by wavefront multisplit:
//                       each wave walks its contiguous quarter of the tile
//                       64 keys at a time; one ballot per digit bit gives
//                       the lanes holding the same digit, the group's rank
//                       is the running per-wave digit count (LDS) plus the
//                       lanes before it.  Keys, then values, are staged in
//                       LDS in digit order (one buffer, two rounds) and
//                       written out so that consecutive threads store
//                       consecutive addresses of each digit's run.
#include "radix.hpp"

namespace spws {

template <class K>
__device__ __forceinline__ int radix_digit(K key, int shift, uint32_t mask) {
  return (int)((key >> shift) & (K)mask);
}

template <class K>
__global__ __launch_bounds__(kSortBlock) void k_radix_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                           uint32_t mask, int32_t* __restrict__ hist, int64_t ntiles) {
  __shared__ int32_t h[kSortBucketsMax];
  const int nb = (int)mask + 1;
  for (int d = threadIdx.x; d < nb; d += kSortBlock) h[d] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int64_t idx = base + (int64_t)it * kSortBlock + threadIdx.x;
    if (idx < n) atomicAdd(&h[radix_digit(keys[idx], shift, mask)], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < nb; d += kSortBlock) hist[(int64_t)d * ntiles + blockIdx.x] = h[d];
}

template <class K>
__global__ __launch_bounds__(kSortBlock, 4) void k_radix_scatter(const K* __restrict__ kin, const double* __restrict__ vin,
                                                              K* __restrict__ kout, double* __restrict__ vout,
                                                              int64_t n, int shift, uint32_t mask, int nbits,
                                                              const int32_t* __restrict__ offs, int64_t ntiles) {
  constexpr int kWaves = kSortBlock / kWave;
  constexpr int kSub = kSortTile / kWaves;  // keys per wave
  static_assert(kSortBucketsMax == 2 * kSortBlock, "two digits per thread in the prefix step");
  static_assert(sizeof(K) <= sizeof(double), "one staging buffer for keys and values");
  __shared__ double stage[kSortTile];  // keys, then values, in digit order
  // per (wavefront, digit): counts, then their prefixes over the wavefronts
  // (at most kSortTile = 4096, so 16 bits: 38.9 KB of LDS in all, four
  // workgroups per CU)
  __shared__ uint16_t wcnt[kWaves][kSortBucketsMax];
  static_assert(kSortTile < 65536, "per-wavefront digit prefixes fit 16 bits");
  __shared__ int32_t dstart[kSortBucketsMax];
  __shared__ int32_t scan_sh[kSortBlock / kWave + 1];
  K* sk = reinterpret_cast<K*>(stage);

  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  const int nb = (int)mask + 1;
  for (int d = threadIdx.x; d < kWaves * kSortBucketsMax; d += kSortBlock) (&wcnt[0][0])[d] = 0;
  __syncthreads();

  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  const int64_t wbase = base + (int64_t)w * kSub;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  K key[kSortItems];
  double val[kSortItems];
  int32_t rank[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t idx = wbase + (int64_t)r * kWave + lane;
    const bool valid = idx < n;
    key[r] = valid ? kin[idx] : (K)0;
    val[r] = valid ? vin[idx] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t idx = wbase + (int64_t)r * kWave + lane;
    const bool valid = idx < n;
    const int d = radix_digit(key[r], shift, mask);
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kSortDigitMax; ++b) {
      if (b < nbits) {  // uniform
        const bool bit = (d >> b) & 1;
        const uint64_t m = __ballot(bit);
        match &= bit ? m : ~m;
      }
    }
    rank[r] = 0;
    if (valid) {
      const int before = __popcll(match & lt);
      const int32_t old = wcnt[w][d];
      rank[r] = old + before;
      if (before == 0) wcnt[w][d] = (uint16_t)(old + __popcll(match));
    }
  }
  __syncthreads();
  {
    // thread t: digits 2t and 2t + 1: exclusive prefixes over the waves,
    // then the tile-wide exclusive prefix of the digit totals
    int32_t tot[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int d = 2 * threadIdx.x + e;
      int32_t run = 0;
      if (d < nb) {
#pragma unroll
        for (int q = 0; q < kWaves; ++q) {
          const int32_t c = wcnt[q][d];
          wcnt[q][d] = (uint16_t)run;
          run += c;
        }
      }
      tot[e] = run;
    }
    int32_t all;
    const int32_t st = block_excl_scan<int32_t, kSortBlock>(tot[0] + tot[1], scan_sh, &all);
    if (2 * (int)threadIdx.x < nb) dstart[2 * threadIdx.x] = st;
    if (2 * (int)threadIdx.x + 1 < nb) dstart[2 * threadIdx.x + 1] = st + tot[0];
  }
  __syncthreads();
  const int64_t remain = n - base;
  const int tile_n = remain < kSortTile ? (int)remain : kSortTile;
  // round 1: keys in digit order, written out; their destinations kept
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t idx = wbase + (int64_t)r * kWave + lane;
    if (idx < n) {
      const int d = radix_digit(key[r], shift, mask);
      sk[dstart[d] + wcnt[w][d] + rank[r]] = key[r];
    }
  }
  __syncthreads();
  int32_t dst[kSortItems];
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int i = it * kSortBlock + threadIdx.x;
    dst[it] = 0;
    if (i < tile_n) {
      const K k = sk[i];
      const int d = radix_digit(k, shift, mask);
      dst[it] = offs[(int64_t)d * ntiles + blockIdx.x] + (i - dstart[d]);
      kout[dst[it]] = k;
    }
  }
  __syncthreads();
  // round 2: the values through the same buffer
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t idx = wbase + (int64_t)r * kWave + lane;
    if (idx < n) {
      const int d = radix_digit(key[r], shift, mask);
      stage[dstart[d] + wcnt[w][d] + rank[r]] = val[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int i = it * kSortBlock + threadIdx.x;
    if (i < tile_n) vout[dst[it]] = stage[i];
  }
}

template <class K>
int radix_sort_pairs(spws_ctx* ctx, K* keys, double* vals, K* keys_alt, double* vals_alt, int64_t n, int begin_bit,
                     int end_bit, bool* in_alt, const double* vals_first) {
  *in_alt = false;
  if (n <= 0) return SPWS_OK;
  if (n <= 1 || end_bit <= begin_bit) {
    if (vals_first) SPWS_HIP(hipMemcpyAsync(vals, vals_first, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    return SPWS_OK;
  }
  SPWS_REQUIRE(n < ((int64_t)1 << 31), SPWS_ERR_UNSUPPORTED, "radix sort of %lld keys exceeds int32 offsets",
               (long long)n);
  const int64_t ntiles = div_up(n, kSortTile);
  int rc = ctx->radix_hist.ensure((2 * radix_hist_entries(n) + 1) * sizeof(int32_t));
  if (rc) return rc;
  int32_t* hist = ctx->radix_hist.as<int32_t>();
  int32_t* offs = hist + radix_hist_entries(n);
  K* kin = keys;
  const double* vin = vals_first ? vals_first : vals;
  K* kout = keys_alt;
  double* vout = vals_alt;
  const int passes = (int)div_up(end_bit - begin_bit, kSortDigitMax);
  int shift = begin_bit;
  for (int p = 0; p < passes; ++p) {
    // the remaining bits split evenly over the remaining passes
    const int nbits = (int)div_up(end_bit - shift, passes - p);
    const uint32_t mask = (1u << nbits) - 1u;
    const int64_t hn = ntiles << nbits;
    SPWS_LAUNCH(ctx, "radix_hist", (k_radix_hist<K>), dim3((unsigned)ntiles), dim3(kSortBlock), 0, kin, n, shift,
                mask, hist, ntiles);
    rc = exclusive_scan<int32_t, int32_t>(ctx, hist, offs, hn);
    if (rc) return rc;
    SPWS_LAUNCH(ctx, "radix_scatter", (k_radix_scatter<K>), dim3((unsigned)ntiles), dim3(kSortBlock), 0, kin, vin,
                kout, vout, n, shift, mask, nbits, offs, ntiles);
    // the next pass reads what this one wrote and writes the other buffer
    // (never vals_first)
    K* tk = kin;
    kin = kout;
    kout = tk;
    double* written = vout;
    vout = (written == vals_alt) ? vals : vals_alt;
    vin = written;
    *in_alt = !*in_alt;
    shift += nbits;
  }
  return SPWS_OK;
}

template int radix_sort_pairs<uint32_t>(spws_ctx*, uint32_t*, double*, uint32_t*, double*, int64_t, int, int, bool*,
                                        const double*);
template int radix_sort_pairs<uint64_t>(spws_ctx*, uint64_t*, double*, uint64_t*, double*, int64_t, int, int, bool*,
                                        const double*);

}  // namespace spws
