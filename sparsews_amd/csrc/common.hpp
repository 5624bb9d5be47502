// Shared device helpers and the per-context state of libspws.
#pragma once
#include <climits>
#include <limits>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "spws/spws.h"

namespace spws {

constexpr int kWave = 64;

// ---------------------------------------------------------------- errors

void set_error(const char* fmt, ...);
int fail_hip(hipError_t e, const char* what, const char* file, int line);

#define SPWS_HIP(call)                                                  \
  do {                                                                  \
    hipError_t e__ = (call);                                            \
    if (e__ != hipSuccess) return ::spws::fail_hip(e__, #call, __FILE__, __LINE__); \
  } while (0)

#define SPWS_CHECK_LAUNCH() SPWS_HIP(hipGetLastError())

#define SPWS_REQUIRE(cond, code, ...)                                   \
  do {                                                                  \
    if (!(cond)) {                                                      \
      ::spws::set_error(__VA_ARGS__);                                   \
      return (code);                                                    \
    }                                                                   \
  } while (0)

// ---------------------------------------------------------------- buffers

// A device buffer that only grows.  Growing frees the old allocation, which
// synchronises the device; steady-state calls of the same size never allocate.
// A fixed buffer (borrow()) views memory owned elsewhere: ensure() never
// allocates and fails when the view is too small, release() only forgets it.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  bool fixed = false;
  int ensure(size_t want);
  void release();
  void borrow(void* p, size_t n) {
    ptr = p;
    bytes = n;
    fixed = true;
  }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
};

struct PinnedBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t want);
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
};

// Per-kernel HIP-event timing on the context's stream (off by default).
struct KernelProfiler {
  bool on = false;
  std::string only;  // when set, only launches of this name are timed
  std::vector<hipEvent_t> pool;
  struct Pending {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::map<std::string, std::pair<double, int64_t>> totals;  // name -> (ms, launches)
};

}  // namespace spws

struct spws_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // named scratch and output buffers (see the users for their layouts)
  spws::DevBuf scan_tmp;
  spws::DevBuf row_ub, row_off, row_nnz, bin_rows, bin_meta;
  bool rw_meta_clean = false;  // bin_meta is zero (left so by the last completed product)
  bool rw_fixed_last = false;  // the last row-wise product took the fixed-slot path
  // short-row product: result counters and per-wavefront totals
  spws::DevBuf short_meta;
  spws::DevBuf slot_crd, slot_val, wide_tab, dense_val, dense_flag;
  spws::DevBuf out_pos, out_crd, out_val, out_aux;
  spws::DevBuf keys[2], vals[2], chunk_a, chunk_b;
  spws::DevBuf tr_keys, tr_vals, tr_counts, tr_out[3];  // CSR transpose: sort pairs, column counts, outputs
  spws::DevBuf stage[12];  // device copies of host inputs (TACO entry points)
  spws::DevBuf radix_hist;
  spws::DevBuf tw[10];     // TTM by roots and windows: counts, prefixes, bucketed nonzeros
  spws::DevBuf obuf[8];    // outer-product per-column metadata
  spws::DevBuf bk[11];     // outer-product bucket path: pairs, window bounds, bucket tables (two sets)
  spws::PinnedBuf host_meta;
  // row-wise SpGEMM bin boundaries (validated against the compiled table sizes)
  spws_rowwise_tuning rw{64, 256, 1024, 4096, 16, -1};
  // outer-product chunk limits (0 = defaults: 2^30 products, key-width rows)
  int64_t outer_max_products = 0;
  int32_t outer_max_rows = 0;
  // outer-product method (0 auto, 1 LSD radix sort, 2 row/window buckets) and
  // bucket window width log2 (0 = chosen from the average row product count)
  int32_t outer_method = 0;
  int32_t outer_window_log2 = 0;
  // the automatic method keeps the sums in order (3 rather than 4): set by
  // the taco entry points, whose results stand in for the reference's
  bool outer_auto_ordered = false;
  // TTM method (0 auto, 1 sorting, 2 roots and column windows)
  int32_t ttm_method = 0;
  // second stream (default priority) and its events, created on first use:
  // the outer product's per-chunk reduce overlaps the next chunk's expand
  hipStream_t aux_stream = nullptr;
  hipEvent_t aux_ready = nullptr;
  hipEvent_t aux_done[2] = {nullptr, nullptr};
  // statistics of the last call (host-readable)
  spws_stats stats{};
  spws::KernelProfiler prof;
};

namespace spws {

int ctx_activate(spws_ctx* ctx);
// The context's auxiliary stream (created with its events on first use).
int ctx_aux_stream(spws_ctx* ctx, hipStream_t* out);

// Kernel timing hooks: prof_begin records a start event (or returns null when
// profiling is off or the launch is not the one being timed), prof_end
// records the stop event; prof_collect turns the
// events recorded so far into totals (call it after the stream is idle).
hipEvent_t prof_begin(spws_ctx* ctx, const char* name);
void prof_end(spws_ctx* ctx, const char* name, hipEvent_t start);
void prof_collect(spws_ctx* ctx);

#define SPWS_LAUNCH(ctx, name, kern, grid, block, shmem, ...)                 \
  do {                                                                        \
    hipEvent_t e0__ = ::spws::prof_begin(ctx, name);                          \
    hipLaunchKernelGGL(kern, grid, block, shmem, (ctx)->stream, __VA_ARGS__); \
    SPWS_CHECK_LAUNCH();                                                      \
    ::spws::prof_end(ctx, name, e0__);                                        \
  } while (0)

// ---------------------------------------------------------------- scan

// Exclusive prefix sum of n values into out[0..n]; out[n] receives the total.
// Tin/Tout in {int32_t, int64_t}.  Uses ctx->scan_tmp.
template <class Tin, class Tout>
int exclusive_scan(spws_ctx* ctx, const Tin* in, Tout* out, int64_t n);
// Bytes of ctx->scan_tmp that exclusive_scan of n values of elem_bytes uses.
size_t scan_scratch_bytes(int64_t n, size_t elem_bytes);

// Resolves the no-allocation entry point's scratch address on `device`
// (current), once per device; spws_ctx_create calls it for its device.
int noalloc_scratch_init(int device);

// ---------------------------------------------------------------- device helpers

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// v moved by a DPP pattern (CTRL) within the rows of ROW_MASK; lanes with no
// source lane take `none` (32- or 64-bit T, moved as one or two dwords).
template <int CTRL, int ROW_MASK, class T>
__device__ __forceinline__ T dpp_move(T v, T none) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, none),
                                                             __builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xF, false));
  } else {
    const uint64_t x = __builtin_bit_cast(uint64_t, v), o = __builtin_bit_cast(uint64_t, none);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)o, (int)(uint32_t)x, CTRL, ROW_MASK, 0xF,
                                                              false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(o >> 32), (int)(uint32_t)(x >> 32),
                                                              CTRL, ROW_MASK, 0xF, false);
    return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
  }
}

// Inclusive wavefront scan with an associative op and its identity, all in
// DPP moves (no LDS): row shifts by 1, 2, 4 and 8 give each row of 16 its
// own prefixes, then the ends of rows 0 and 2 are broadcast into rows 1
// and 3, and the end of rows 0-1 into rows 2-3.  Every lane of the
// wavefront must be active.
template <class T, class Op>
__device__ __forceinline__ T wave_incl_scan_op(T v, T id, Op op) {
  v = op(v, dpp_move<0x111, 0xF>(v, id));  // row_shr:1
  v = op(v, dpp_move<0x112, 0xF>(v, id));  // row_shr:2
  v = op(v, dpp_move<0x114, 0xF>(v, id));  // row_shr:4
  v = op(v, dpp_move<0x118, 0xF>(v, id));  // row_shr:8
  v = op(v, dpp_move<0x142, 0xA>(v, id));  // row_bcast:15
  v = op(v, dpp_move<0x143, 0xC>(v, id));  // row_bcast:31
  return v;
}

// v of lane n of this lane's row of 16 (DPP row_newbcast: no LDS); n must
// fold to a constant where it is used (unrolled loops do).
template <class T>
__device__ __forceinline__ T row_lane(T v, int n) {
  switch (n & 15) {
    case 0: return dpp_move<0x150, 0xF>(v, v);
    case 1: return dpp_move<0x151, 0xF>(v, v);
    case 2: return dpp_move<0x152, 0xF>(v, v);
    case 3: return dpp_move<0x153, 0xF>(v, v);
    case 4: return dpp_move<0x154, 0xF>(v, v);
    case 5: return dpp_move<0x155, 0xF>(v, v);
    case 6: return dpp_move<0x156, 0xF>(v, v);
    case 7: return dpp_move<0x157, 0xF>(v, v);
    case 8: return dpp_move<0x158, 0xF>(v, v);
    case 9: return dpp_move<0x159, 0xF>(v, v);
    case 10: return dpp_move<0x15A, 0xF>(v, v);
    case 11: return dpp_move<0x15B, 0xF>(v, v);
    case 12: return dpp_move<0x15C, 0xF>(v, v);
    case 13: return dpp_move<0x15D, 0xF>(v, v);
    case 14: return dpp_move<0x15E, 0xF>(v, v);
    case 15: return dpp_move<0x15F, 0xF>(v, v);
  }
  return v;
}

// The value of lane 63, in every lane.
template <class T>
__device__ __forceinline__ T wave_last(T v) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), kWave - 1));
  } else {
    const uint64_t x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, kWave - 1);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), kWave - 1);
    return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
  }
}

// Inclusive wavefront max-scan of unsigned values (every lane active).
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v) {
  return wave_incl_scan_op(v, 0u, [](uint32_t x, uint32_t y) { return max(x, y); });
}

// Integer wavefront scans and reductions (every lane active).
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  return wave_incl_scan_op(v, (T)0, [](T a, T b) { return a + b; });
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
  return wave_last(wave_incl_scan_op(v, std::numeric_limits<T>::lowest(), [](T a, T b) { return max(a, b); }));
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
  return wave_last(wave_incl_scan(v));
}

// v of lane (lane ^ D) without LDS: DPP quad permutes, half-row and row
// mirrors and the row rotate by 8, and gfx950's 16- and 32-lane swaps
// (each swap returns both halves exchanged; the lane keeps the half that
// came from its partner).
template <int D>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
  if constexpr (D == 1) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
  } else if constexpr (D == 2) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
  } else if constexpr (D == 3) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x1B, 0xF, 0xF, true);  // quad_perm [3,2,1,0]
  } else if constexpr (D == 4) {
    return lane_xor<3>(lane_xor<7>(v));
  } else if constexpr (D == 7) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, true);  // row_half_mirror
  } else if constexpr (D == 8) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, true);  // row_ror:8
  } else if constexpr (D == 15) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, true);  // row_mirror
  } else if constexpr (D == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane_id() & 16) ? p[0] : p[1];
  } else if constexpr (D == 31) {
    return lane_xor<16>(lane_xor<15>(v));
  } else if constexpr (D == 32) {
    const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane_id() & 32) ? p[0] : p[1];
  } else {
    static_assert(D == 63, "lane_xor: unsupported distance");
    return lane_xor<32>(lane_xor<31>(v));
  }
}

// One comparator column of the flip bitonic network on R independent
// 64-key sequences (register r of every lane): the lower lane of each pair
// keeps the minimum.
template <int D, int R>
__device__ __forceinline__ void wave_sort64_step(uint32_t (&k)[R], bool low) {
  uint32_t p[R];
#pragma unroll
  for (int r = 0; r < R; ++r) p[r] = lane_xor<D>(k[r]);
  // the median of (key, partner's key, 0 or ~0): the minimum in the lower
  // lane, the maximum in the upper one
  const uint32_t bound = low ? 0u : ~0u;
#pragma unroll
  for (int r = 0; r < R; ++r) k[r] = max(min(k[r], p[r]), min(max(k[r], p[r]), bound));
}

// Sorts R independent sequences of 64 unsigned keys (register r of every
// lane is sequence r) ascending; only the first n_pow2 (a power of two,
// 2..64) entries of each take part, the rest must already hold 0xFFFFFFFF.
// Every exchange is a DPP move or a lane swap, none goes through LDS.
template <int R>
__device__ __forceinline__ void wave_sort64(uint32_t (&k)[R], int n_pow2) {
  const int lane = lane_id();
  const bool b0 = !(lane & 1), b1 = !(lane & 2), b2 = !(lane & 4), b3 = !(lane & 8), b4 = !(lane & 16);
  wave_sort64_step<1>(k, b0);
  if (n_pow2 >= 4) {
    wave_sort64_step<3>(k, b1);
    wave_sort64_step<1>(k, b0);
  }
  if (n_pow2 >= 8) {
    wave_sort64_step<7>(k, b2);
    wave_sort64_step<2>(k, b1);
    wave_sort64_step<1>(k, b0);
  }
  if (n_pow2 >= 16) {
    wave_sort64_step<15>(k, b3);
    wave_sort64_step<4>(k, b2);
    wave_sort64_step<2>(k, b1);
    wave_sort64_step<1>(k, b0);
  }
  if (n_pow2 >= 32) {
    wave_sort64_step<31>(k, b4);
    wave_sort64_step<8>(k, b3);
    wave_sort64_step<4>(k, b2);
    wave_sort64_step<2>(k, b1);
    wave_sort64_step<1>(k, b0);
  }
  if (n_pow2 >= 64) {
    wave_sort64_step<63>(k, !(lane & 32));
    wave_sort64_step<16>(k, b4);
    wave_sort64_step<8>(k, b3);
    wave_sort64_step<4>(k, b2);
    wave_sort64_step<2>(k, b1);
    wave_sort64_step<1>(k, b0);
  }
}

// Wavefront bitonic sort, ascending ("flip" network: stage k first compares
// e with e ^ (k - 1), then e with e ^ j for j = k/4 .. 1, so padding with
// INT_MAX sorts any n).  R keys per lane, element e = r * 64 + lane; the
// first n_pow2 elements (a power of two, at least 2) take part.  Every
// distance is a compile-time constant: its lane part is a DPP move or lane
// swap (lane_xor), its register part a renaming, and each comparator one
// median of (key, partner, INT_MIN or INT_MAX).
template <int D, int R>
__device__ __forceinline__ void wave_bitonic_step(int32_t (&v)[R]) {
  constexpr int DL = D & (kWave - 1), DR = D >> 6;
  // the lower element of each pair (bit M of e clear) keeps the minimum
  constexpr int M = (D & (D + 1)) == 0 ? (D + 1) >> 1 : D;  // flip distance 2^k - 1, or a power of two
  const int lane = lane_id();
  int32_t p[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int32_t src = v[r ^ DR];
    if constexpr (DL != 0) p[r] = (int32_t)lane_xor<DL>((uint32_t)src);
    else p[r] = src;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool low = M < kWave ? (lane & M) == 0 : ((r << 6) & M) == 0;
    const int32_t bound = low ? INT_MIN : INT_MAX;
    v[r] = max(min(v[r], p[r]), min(max(v[r], p[r]), bound));
  }
}

template <int J, int R>
__device__ __forceinline__ void wave_bitonic_merge(int32_t (&v)[R]) {
  if constexpr (J >= 1) {
    wave_bitonic_step<J, R>(v);
    wave_bitonic_merge<J / 2, R>(v);
  }
}

template <int KK, int R>
__device__ __forceinline__ void wave_bitonic_stages(int32_t (&v)[R], int n_pow2) {
  if constexpr (KK <= kWave * R) {
    if (n_pow2 >= KK) {  // uniform
      wave_bitonic_step<KK - 1, R>(v);
      wave_bitonic_merge<KK / 4, R>(v);
      wave_bitonic_stages<KK * 2, R>(v, n_pow2);
    }
  }
}

template <int R>
__device__ __forceinline__ void wave_bitonic(int32_t (&v)[R], int n_pow2) {
  static_assert(R == 1 || R == 2 || R == 4 || R == 8, "R keys per lane: a power of two up to 8");
  wave_bitonic_stages<2, R>(v, n_pow2);
}

// Exclusive scan over a block of BLOCK threads; returns this thread's prefix
// and writes the block total to *total.  `sh` holds BLOCK/64 + 1 entries.
template <class T, int BLOCK>
__device__ __forceinline__ T block_excl_scan(T v, T* sh, T* total) {
  constexpr int NW = BLOCK / kWave;
  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  T incl = wave_incl_scan(v);
  if (lane == kWave - 1) sh[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int i = 0; i < NW; ++i) {
      T t = sh[i];
      sh[i] = run;
      run += t;
    }
    sh[NW] = run;
  }
  __syncthreads();
  T res = sh[w] + incl - v;
  *total = sh[NW];
  __syncthreads();
  return res;
}

// Fibonacci hash of a column index into a table of 2^(32 - shift) slots
// (the high bits of the product, so strided column patterns spread out).
__device__ __forceinline__ uint32_t hash_col(uint32_t k, uint32_t shift) {
  return (k * 2654435761u) >> shift;
}

__host__ __device__ __forceinline__ int log2_pow2(uint32_t v) {
  int r = 0;
  while ((1u << r) < v) ++r;
  return r;
}

__host__ __device__ __forceinline__ int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

__host__ __device__ __forceinline__ uint32_t next_pow2_u32(uint32_t v) {
  v--;
  v |= v >> 1;
  v |= v >> 2;
  v |= v >> 4;
  v |= v >> 8;
  v |= v >> 16;
  return v + 1;
}

}  // namespace spws
