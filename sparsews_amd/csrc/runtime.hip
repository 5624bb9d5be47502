// Context, error reporting, device buffers and the device prefix sum.
#include <cstdarg>
#include <cstdlib>

#include "common.hpp"

namespace spws {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int fail_hip(hipError_t e, const char* what, const char* file, int line) {
  set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return SPWS_ERR_OOM;
  return SPWS_ERR_DEVICE;
}

int DevBuf::ensure(size_t want) {
  if (want <= bytes && ptr) return SPWS_OK;
  if (fixed) {
    set_error("a fixed scratch region of %zu bytes cannot hold %zu", bytes, want);
    return SPWS_ERR_UNSUPPORTED;
  }
  release();
  size_t sz = want < 256 ? 256 : want;
  // grow geometrically so a slowly growing workload does not reallocate each call
  sz = sz + sz / 4;
  hipError_t e = hipMalloc(&ptr, sz);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ptr = nullptr;
    bytes = 0;
    set_error("hipMalloc(%zu bytes) failed: %s", sz, hipGetErrorString(e));
    return SPWS_ERR_OOM;
  }
  bytes = sz;
  return SPWS_OK;
}

void DevBuf::release() {
  if (ptr && !fixed) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
  fixed = false;
}

int PinnedBuf::ensure(size_t want) {
  if (want <= bytes && ptr) return SPWS_OK;
  release();
  size_t sz = want < 4096 ? 4096 : want;
  hipError_t e = hipHostMalloc(&ptr, sz, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ptr = nullptr;
    bytes = 0;
    set_error("hipHostMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
    return SPWS_ERR_OOM;
  }
  bytes = sz;
  return SPWS_OK;
}

void PinnedBuf::release() {
  if (ptr) (void)hipHostFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

int ctx_activate(spws_ctx* ctx) {
  SPWS_REQUIRE(ctx != nullptr, SPWS_ERR_INVALID, "null context");
  SPWS_HIP(hipSetDevice(ctx->device));
  return SPWS_OK;
}

int ctx_aux_stream(spws_ctx* ctx, hipStream_t* out) {
  if (!ctx->aux_stream) {
    SPWS_HIP(hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
    SPWS_HIP(hipEventCreateWithFlags(&ctx->aux_ready, hipEventDisableTiming));
    for (hipEvent_t& e : ctx->aux_done) SPWS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  *out = ctx->aux_stream;
  return SPWS_OK;
}

static hipEvent_t prof_event(spws_ctx* ctx) {
  if (!ctx->prof.pool.empty()) {
    hipEvent_t e = ctx->prof.pool.back();
    ctx->prof.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

hipEvent_t prof_begin(spws_ctx* ctx, const char* name) {
  if (!ctx->prof.on) return nullptr;
  if (!ctx->prof.only.empty() && ctx->prof.only != name) return nullptr;
  hipEvent_t e = prof_event(ctx);
  if (e) (void)hipEventRecord(e, ctx->stream);
  return e;
}

void prof_end(spws_ctx* ctx, const char* name, hipEvent_t start) {
  if (!ctx->prof.on || !start) return;
  hipEvent_t e = prof_event(ctx);
  if (!e) return;
  (void)hipEventRecord(e, ctx->stream);
  ctx->prof.pending.push_back({name, start, e});
}

void prof_collect(spws_ctx* ctx) {
  if (ctx->prof.pending.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->prof.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto& t = ctx->prof.totals[p.name];
      t.first += ms;
      t.second += 1;
    } else {
      (void)hipGetLastError();
    }
    ctx->prof.pool.push_back(p.a);
    ctx->prof.pool.push_back(p.b);
  }
  ctx->prof.pending.clear();
}

// ---------------------------------------------------------------- scan
//
// Reduce-then-scan: kernel 1 reduces tiles of kScanTile values; then, up to
// kScanFuseTiles tiles, one kernel in which each workgroup sums the earlier
// tiles' sums itself and scans its tile (coalesced: each wavefront takes a
// quarter of the tile 64 values at a time, DPP wavefront scans with a
// running carry); beyond that, a single-workgroup scan of the tile sums
// and a rescan of each tile with its offset.  Inputs of at most four tiles
// take a single-workgroup path.

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;
constexpr int64_t kScanFuseTiles = 8192;  // earlier-tile sums each workgroup reads at most

template <class Tin, class Tout>
__global__ __launch_bounds__(kScanBlock) void k_scan_tile_reduce(const Tin* __restrict__ in, int64_t n,
                                                                 Tout* __restrict__ tile_sums) {
  __shared__ Tout sh[kScanBlock / kWave + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  Tout s = 0;
#pragma unroll
  for (int it = 0; it < kScanItems; ++it) {
    int64_t idx = base + (int64_t)it * kScanBlock + threadIdx.x;
    if (idx < n) s += (Tout)in[idx];
  }
  s = wave_sum(s);
  if (lane_id() == 0) sh[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    Tout t = 0;
    for (int w = 0; w < kScanBlock / kWave; ++w) t += sh[w];
    tile_sums[blockIdx.x] = t;
  }
}

// Single-workgroup exclusive scan of n values (any n); out[n] = total.
template <class Tin, class Tout>
__global__ __launch_bounds__(kScanBlock) void k_scan_single(const Tin* __restrict__ in, int64_t n,
                                                            Tout* __restrict__ out) {
  __shared__ Tout sh[kScanBlock / kWave + 1];
  Tout carry = 0;
  for (int64_t base = 0; base < n; base += kScanTile) {
    // each thread owns kScanItems consecutive values of this tile
    Tout v[kScanItems];
    Tout local = 0;
    const int64_t mine = base + (int64_t)threadIdx.x * kScanItems;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
      int64_t idx = mine + it;
      v[it] = idx < n ? (Tout)in[idx] : (Tout)0;
      local += v[it];
    }
    Tout total;
    Tout pre = block_excl_scan<Tout, kScanBlock>(local, sh, &total);
    Tout run = carry + pre;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
      int64_t idx = mine + it;
      if (idx < n) out[idx] = run;
      run += v[it];
    }
    carry += total;
  }
  if (threadIdx.x == 0) out[n] = carry;
}

template <class Tin, class Tout>
__global__ __launch_bounds__(kScanBlock) void k_scan_tile_apply(const Tin* __restrict__ in, int64_t n,
                                                                const Tout* __restrict__ tile_off,
                                                                Tout* __restrict__ out) {
  __shared__ Tout sh[kScanBlock / kWave + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  Tout v[kScanItems];
  Tout local = 0;
  const int64_t mine = base + (int64_t)threadIdx.x * kScanItems;
#pragma unroll
  for (int it = 0; it < kScanItems; ++it) {
    int64_t idx = mine + it;
    v[it] = idx < n ? (Tout)in[idx] : (Tout)0;
    local += v[it];
  }
  Tout total;
  Tout pre = block_excl_scan<Tout, kScanBlock>(local, sh, &total);
  Tout run = tile_off[blockIdx.x] + pre;
#pragma unroll
  for (int it = 0; it < kScanItems; ++it) {
    int64_t idx = mine + it;
    if (idx < n) out[idx] = run;
    run += v[it];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = tile_off[gridDim.x];
}

// Tile scan with the tile's offset summed in place from the tile sums:
// out[i] = sum of in[0, i); the last workgroup also writes out[n].
template <class Tin, class Tout>
__global__ __launch_bounds__(kScanBlock) void k_scan_tile_fused(const Tin* __restrict__ in, int64_t n,
                                                                const Tout* __restrict__ tile_sums,
                                                                Tout* __restrict__ out) {
  constexpr int kWaves = kScanBlock / kWave;
  constexpr int kRounds = kScanTile / kScanBlock;  // rounds of 64 values per wavefront
  __shared__ Tout sh[2 * kWaves];
  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  // this tile's offset: the sums of the tiles before it
  Tout off = 0;
  for (int64_t t = threadIdx.x; t < (int64_t)blockIdx.x; t += kScanBlock) off += tile_sums[t];
  // this wavefront's quarter of the tile
  const int64_t wbase = (int64_t)blockIdx.x * kScanTile + (int64_t)w * (kRounds * kWave);
  Tout v[kRounds];
  Tout mine = 0;
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int64_t idx = wbase + r * kWave + lane;
    v[r] = idx < n ? (Tout)in[idx] : (Tout)0;
    mine += v[r];
  }
  off = wave_sum(off);
  mine = wave_sum(mine);
  if (lane == 0) {
    sh[w] = off;
    sh[kWaves + w] = mine;
  }
  __syncthreads();
  Tout carry = 0, total = 0;
#pragma unroll
  for (int q = 0; q < kWaves; ++q) {
    carry += sh[q];  // the tile offset, summed over the threads
    if (q < w) carry += sh[kWaves + q];
    total += sh[kWaves + q];
  }
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int64_t idx = wbase + r * kWave + lane;
    const Tout incl = wave_incl_scan(v[r]) + carry;
    if (idx < n) out[idx] = incl - v[r];
    carry = wave_last(incl);
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    Tout base = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) base += sh[q];
    out[n] = base + total;
  }
}

size_t scan_scratch_bytes(int64_t n, size_t elem_bytes) {
  if (n <= 4 * (int64_t)kScanTile) return 0;
  return (size_t)(div_up(n, kScanTile) + 1) * 2 * elem_bytes;
}

template <class Tin, class Tout>
int exclusive_scan(spws_ctx* ctx, const Tin* in, Tout* out, int64_t n) {
  if (n <= 4 * (int64_t)kScanTile) {
    SPWS_LAUNCH(ctx, "scan_single", (k_scan_single<Tin, Tout>), dim3(1), dim3(kScanBlock), 0, in, n, out);
    return SPWS_OK;
  }
  const int64_t tiles = div_up(n, kScanTile);
  int rc = ctx->scan_tmp.ensure((size_t)(tiles + 1) * 2 * sizeof(Tout));
  if (rc) return rc;
  Tout* sums = ctx->scan_tmp.as<Tout>();
  Tout* offs = sums + (tiles + 1);
  SPWS_LAUNCH(ctx, "scan_tile_reduce", (k_scan_tile_reduce<Tin, Tout>), dim3((unsigned)tiles), dim3(kScanBlock),
              0, in, n, sums);
  if (tiles <= kScanFuseTiles) {
    SPWS_LAUNCH(ctx, "scan_tile_apply", (k_scan_tile_fused<Tin, Tout>), dim3((unsigned)tiles), dim3(kScanBlock), 0,
                in, n, (const Tout*)sums, out);
    return SPWS_OK;
  }
  SPWS_LAUNCH(ctx, "scan_single", (k_scan_single<Tout, Tout>), dim3(1), dim3(kScanBlock), 0, sums, tiles, offs);
  SPWS_LAUNCH(ctx, "scan_tile_apply", (k_scan_tile_apply<Tin, Tout>), dim3((unsigned)tiles), dim3(kScanBlock), 0,
              in, n, offs, out);
  return SPWS_OK;
}

__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_copy1(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

template int exclusive_scan<int32_t, int32_t>(spws_ctx*, const int32_t*, int32_t*, int64_t);
template int exclusive_scan<int32_t, int64_t>(spws_ctx*, const int32_t*, int64_t*, int64_t);
template int exclusive_scan<int64_t, int64_t>(spws_ctx*, const int64_t*, int64_t*, int64_t);

}  // namespace spws

using namespace spws;

extern "C" {

const char* spws_last_error(void) { return g_last_error.c_str(); }

int spws_ctx_create(int device, void* stream, spws_ctx** out) {
  SPWS_REQUIRE(out != nullptr, SPWS_ERR_INVALID, "null output pointer");
  int ndev = 0;
  SPWS_HIP(hipGetDeviceCount(&ndev));
  SPWS_REQUIRE(device >= 0 && device < ndev, SPWS_ERR_INVALID, "device %d out of range (%d devices)",
               device, ndev);
  SPWS_HIP(hipSetDevice(device));
  spws_ctx* ctx = new spws_ctx();
  ctx->device = device;
  if (stream) {
    ctx->stream = static_cast<hipStream_t>(stream);
  } else {
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete ctx;
      return fail_hip(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    ctx->own_stream = true;
  }
  int rc = ctx->host_meta.ensure(4096);
  if (!rc) rc = noalloc_scratch_init(device);
  if (rc) {
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return rc;
  }
  *out = ctx;
  return SPWS_OK;
}

int spws_ctx_destroy(spws_ctx* ctx) {
  if (!ctx) return SPWS_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  DevBuf* bufs[] = {&ctx->scan_tmp, &ctx->row_ub,   &ctx->row_off,    &ctx->row_nnz,   &ctx->bin_rows,
                    &ctx->bin_meta, &ctx->slot_crd, &ctx->slot_val,   &ctx->wide_tab,  &ctx->dense_val,
                    &ctx->dense_flag, &ctx->out_pos, &ctx->out_crd,   &ctx->out_val,   &ctx->out_aux,
                    &ctx->keys[0],  &ctx->keys[1],  &ctx->vals[0],    &ctx->vals[1],   &ctx->chunk_a,
                    &ctx->chunk_b,  &ctx->tr_keys,  &ctx->tr_vals,    &ctx->tr_counts, &ctx->tr_out[0],
                    &ctx->tr_out[1], &ctx->tr_out[2], &ctx->short_meta};
  if (ctx->aux_stream) (void)hipStreamSynchronize(ctx->aux_stream);
  for (DevBuf* b : bufs) b->release();
  for (DevBuf& b : ctx->stage) b.release();
  for (DevBuf& b : ctx->obuf) b.release();
  for (DevBuf& b : ctx->bk) b.release();
  for (DevBuf& b : ctx->tw) b.release();
  if (ctx->aux_ready) (void)hipEventDestroy(ctx->aux_ready);
  for (hipEvent_t e : ctx->aux_done)
    if (e) (void)hipEventDestroy(e);
  if (ctx->aux_stream) (void)hipStreamDestroy(ctx->aux_stream);
  ctx->radix_hist.release();
  ctx->host_meta.release();
  prof_collect(ctx);
  for (hipEvent_t e : ctx->prof.pool) (void)hipEventDestroy(e);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SPWS_OK;
}

int spws_ctx_set_stream(spws_ctx* ctx, void* stream) {
  SPWS_REQUIRE(ctx != nullptr, SPWS_ERR_INVALID, "null context");
  if (ctx->own_stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    ctx->own_stream = false;
  }
  if (stream) {
    ctx->stream = static_cast<hipStream_t>(stream);
  } else {
    SPWS_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->own_stream = true;
  }
  return SPWS_OK;
}

void* spws_ctx_get_stream(spws_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int spws_ctx_get_stats(spws_ctx* ctx, spws_stats* out) {
  SPWS_REQUIRE(ctx && out, SPWS_ERR_INVALID, "null argument");
  *out = ctx->stats;
  return SPWS_OK;
}

int spws_dev_alloc(spws_ctx* ctx, size_t bytes, void** out) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_REQUIRE(out != nullptr, SPWS_ERR_INVALID, "null output pointer");
  hipError_t e = hipMalloc(out, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return SPWS_ERR_OOM;
  }
  return SPWS_OK;
}

int spws_dev_free(spws_ctx* ctx, void* p) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  if (p) SPWS_HIP(hipFree(p));
  return SPWS_OK;
}

int spws_memcpy_h2d(spws_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  if (bytes == 0) return SPWS_OK;
  SPWS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  SPWS_HIP(hipStreamSynchronize(ctx->stream));
  return SPWS_OK;
}

int spws_memcpy_d2h(spws_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  if (bytes == 0) return SPWS_OK;
  SPWS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SPWS_HIP(hipStreamSynchronize(ctx->stream));
  return SPWS_OK;
}

int spws_sync(spws_ctx* ctx) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_HIP(hipStreamSynchronize(ctx->stream));
  return SPWS_OK;
}

void spws_host_free(void* p) { free(p); }

int spws_device_copy(spws_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_REQUIRE((dst && src) || bytes == 0, SPWS_ERR_INVALID, "null pointer");
  SPWS_REQUIRE(((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0, SPWS_ERR_INVALID,
               "device_copy needs 16-byte aligned pointers");
  const int64_t n16 = (int64_t)(bytes / 16);
  const int64_t tail = (int64_t)(bytes % 16);
  if (n16 > 0) {
    int64_t g = div_up(n16, 256);
    if (g > 8192) g = 8192;
    SPWS_LAUNCH(ctx, "copy16", k_copy16, dim3((unsigned)g), dim3(256), 0, (const uint4*)src, (uint4*)dst, n16);
  }
  if (tail > 0) {
    SPWS_LAUNCH(ctx, "copy1", k_copy1, dim3(1), dim3(256), 0, (const uint8_t*)src + n16 * 16,
                (uint8_t*)dst + n16 * 16, tail);
  }
  return SPWS_OK;
}

int spws_ctx_set_profiling(spws_ctx* ctx, int on) {
  SPWS_REQUIRE(ctx != nullptr, SPWS_ERR_INVALID, "null context");
  prof_collect(ctx);
  ctx->prof.on = on != 0;
  return SPWS_OK;
}

int spws_ctx_set_profiling_only(spws_ctx* ctx, const char* name) {
  SPWS_REQUIRE(ctx != nullptr, SPWS_ERR_INVALID, "null context");
  prof_collect(ctx);
  ctx->prof.only = name ? name : "";
  return SPWS_OK;
}

int spws_ctx_reset_kernel_times(spws_ctx* ctx) {
  SPWS_REQUIRE(ctx != nullptr, SPWS_ERR_INVALID, "null context");
  prof_collect(ctx);
  ctx->prof.totals.clear();
  return SPWS_OK;
}

int spws_ctx_kernel_times(spws_ctx* ctx, int index, char* name, size_t name_len, double* total_ms,
                          int64_t* launches) {
  SPWS_REQUIRE(ctx != nullptr, SPWS_ERR_INVALID, "null context");
  prof_collect(ctx);
  SPWS_REQUIRE(index >= 0 && (size_t)index < ctx->prof.totals.size(), SPWS_ERR_INVALID,
               "kernel index %d out of range", index);
  auto it = ctx->prof.totals.begin();
  std::advance(it, index);
  if (name && name_len) {
    strncpy(name, it->first.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (total_ms) *total_ms = it->second.first;
  if (launches) *launches = it->second.second;
  return SPWS_OK;
}

}  // extern "C"
