// The workspace API of utils/wspace.h on the device: Sort (sort a
// coordinate workspace by its coordinates) and Merge (merge a sorted
// workspace into a sorted COO, adding the values of equal coordinates).
//
// Records are the reference's wspace struct with an FP64 value:
// {int32 crd[2]; double val;} (16 bytes).  Coordinates are packed into a
// uint64 key whose order is the signed lexicographic order of (crd0, crd1)
// (sign bits flipped), so an empty slot (crd0 = -1, init_wspace) sorts before
// every real coordinate, as in the reference's esc_cmp order.
#include <algorithm>
#include <vector>

#include "radix.hpp"

namespace spws {
namespace {

struct WsRecord {
  int32_t crd[2];
  double val;
};

__device__ __forceinline__ uint64_t ws_key(int32_t c0, int32_t c1) {
  return ((uint64_t)((uint32_t)c0 ^ 0x80000000u) << 32) | (uint64_t)((uint32_t)c1 ^ 0x80000000u);
}

__global__ __launch_bounds__(256) void k_ws_pack(const WsRecord* __restrict__ rec, int64_t n,
                                                 uint64_t* __restrict__ keys, double* __restrict__ vals) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const WsRecord r = rec[i];
    keys[i] = ws_key(r.crd[0], r.crd[1]);
    vals[i] = r.val;
  }
}

__global__ __launch_bounds__(256) void k_ws_pack_coo(const int32_t* __restrict__ c0, const int32_t* __restrict__ c1,
                                                     const double* __restrict__ v, int64_t n,
                                                     uint64_t* __restrict__ keys, double* __restrict__ vals) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    keys[i] = ws_key(c0[i], c1[i]);
    vals[i] = v[i];
  }
}

// sorted keys -> records, ascending or reversed
__global__ __launch_bounds__(256) void k_ws_unpack(const uint64_t* __restrict__ keys, const double* __restrict__ vals,
                                                   int64_t n, int rev, WsRecord* __restrict__ rec) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t src = rev ? n - 1 - i : i;
    const uint64_t k = keys[src];
    WsRecord r;
    r.crd[0] = (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
    r.crd[1] = (int32_t)((uint32_t)k ^ 0x80000000u);
    r.val = vals[src];
    rec[i] = r;
  }
}

// heads of runs of equal keys: flag[i] = 1 for the first of each run
__global__ __launch_bounds__(256) void k_ws_heads(const uint64_t* __restrict__ keys, int64_t n,
                                                  int32_t* __restrict__ flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// each head sums its run in sorted order and writes one COO entry
__global__ __launch_bounds__(256) void k_ws_reduce(const uint64_t* __restrict__ keys, const double* __restrict__ vals,
                                                   int64_t n, const int32_t* __restrict__ flag,
                                                   const int32_t* __restrict__ off, int32_t* __restrict__ c0,
                                                   int32_t* __restrict__ c1, double* __restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (!flag[i]) continue;
    const uint64_t k = keys[i];
    double s = vals[i];
    for (int64_t q = i + 1; q < n && keys[q] == k; ++q) s += vals[q];
    const int32_t u = off[i];
    c0[u] = (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
    c1[u] = (int32_t)((uint32_t)k ^ 0x80000000u);
    v[u] = s;
  }
}

int ws_buffers(spws_ctx* ctx, int64_t n) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  int rc;
  if ((rc = ctx->keys[0].ensure(nn * 8)) || (rc = ctx->keys[1].ensure(nn * 8)) || (rc = ctx->vals[0].ensure(nn * 8)) ||
      (rc = ctx->vals[1].ensure(nn * 8)) || (rc = ctx->stage[8].ensure(nn * sizeof(WsRecord))))
    return rc;
  return SPWS_OK;
}

unsigned grid_of(int64_t n) { return (unsigned)std::min<int64_t>(div_up(n > 0 ? n : 1, 256), 16384); }

}  // namespace
}  // namespace spws

using namespace spws;

extern "C" {

int spws_wspace_sort(spws_ctx* ctx, void* records, int64_t n, int rev, int64_t* nvalid) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_REQUIRE(n >= 0 && (records || n == 0), SPWS_ERR_INVALID, "null records");
  SPWS_REQUIRE(n < ((int64_t)1 << 31), SPWS_ERR_UNSUPPORTED, "workspace too large");
  hipStream_t s = ctx->stream;
  int64_t valid = n;
  if (n > 0) {
    if ((rc = ws_buffers(ctx, n))) return rc;
    WsRecord* d = ctx->stage[8].as<WsRecord>();
    SPWS_HIP(hipMemcpyAsync(d, records, (size_t)n * sizeof(WsRecord), hipMemcpyHostToDevice, s));
    uint64_t* k = ctx->keys[0].as<uint64_t>();
    double* v = ctx->vals[0].as<double>();
    SPWS_LAUNCH(ctx, "ws_pack", k_ws_pack, dim3(grid_of(n)), dim3(256), 0, d, n, k, v);
    bool alt = false;
    if ((rc = radix_sort_pairs<uint64_t>(ctx, k, v, ctx->keys[1].as<uint64_t>(), ctx->vals[1].as<double>(), n, 0, 64,
                                         &alt)))
      return rc;
    if (alt) {
      k = ctx->keys[1].as<uint64_t>();
      v = ctx->vals[1].as<double>();
    }
    SPWS_LAUNCH(ctx, "ws_unpack", k_ws_unpack, dim3(grid_of(n)), dim3(256), 0, k, v, n, rev, d);
    SPWS_HIP(hipMemcpyAsync(records, d, (size_t)n * sizeof(WsRecord), hipMemcpyDeviceToHost, s));
    SPWS_HIP(hipStreamSynchronize(s));
    if (rev) {
      // descending: empty slots (crd0 == -1) are last; count the others
      const WsRecord* r = static_cast<const WsRecord*>(records);
      valid = n;
      while (valid > 0 && r[valid - 1].crd[0] == -1) --valid;
    }
  }
  if (nvalid) *nvalid = valid;
  return SPWS_OK;
}

int spws_wspace_merge(spws_ctx* ctx, int32_t* coo_crd0, int32_t* coo_crd1, double* coo_vals, int64_t coo_size,
                      const void* records, int64_t n_acc, int64_t* out_size) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_REQUIRE(coo_size >= 0 && n_acc >= 0 && out_size, SPWS_ERR_INVALID, "bad sizes");
  SPWS_REQUIRE((coo_size == 0 || (coo_crd0 && coo_crd1 && coo_vals)) && (n_acc == 0 || records), SPWS_ERR_INVALID,
               "null array");
  const int64_t n = coo_size + n_acc;
  SPWS_REQUIRE(n < ((int64_t)1 << 31), SPWS_ERR_UNSUPPORTED, "workspace too large");
  hipStream_t s = ctx->stream;
  if (n == 0) {
    *out_size = 0;
    return SPWS_OK;
  }
  if ((rc = ws_buffers(ctx, n)) || (rc = ctx->stage[9].ensure((size_t)n * 4 * 2)) ||
      (rc = ctx->stage[10].ensure((size_t)n * 8)) || (rc = ctx->stage[11].ensure(((size_t)n + 1) * 4 * 2)))
    return rc;
  uint64_t* k = ctx->keys[0].as<uint64_t>();
  double* v = ctx->vals[0].as<double>();
  int32_t* c = ctx->stage[9].as<int32_t>();
  double* cv = ctx->stage[10].as<double>();
  // the COO first, then the workspace: equal coordinates sum COO + workspace
  if (coo_size > 0) {
    SPWS_HIP(hipMemcpyAsync(c, coo_crd0, (size_t)coo_size * 4, hipMemcpyHostToDevice, s));
    SPWS_HIP(hipMemcpyAsync(c + n, coo_crd1, (size_t)coo_size * 4, hipMemcpyHostToDevice, s));
    SPWS_HIP(hipMemcpyAsync(cv, coo_vals, (size_t)coo_size * 8, hipMemcpyHostToDevice, s));
    SPWS_LAUNCH(ctx, "ws_pack_coo", k_ws_pack_coo, dim3(grid_of(coo_size)), dim3(256), 0, c, c + n, cv, coo_size, k,
                v);
  }
  if (n_acc > 0) {
    WsRecord* d = ctx->stage[8].as<WsRecord>();
    SPWS_HIP(hipMemcpyAsync(d, records, (size_t)n_acc * sizeof(WsRecord), hipMemcpyHostToDevice, s));
    SPWS_LAUNCH(ctx, "ws_pack", k_ws_pack, dim3(grid_of(n_acc)), dim3(256), 0, d, n_acc, k + coo_size, v + coo_size);
  }
  bool alt = false;
  if ((rc = radix_sort_pairs<uint64_t>(ctx, k, v, ctx->keys[1].as<uint64_t>(), ctx->vals[1].as<double>(), n, 0, 64,
                                       &alt)))
    return rc;
  if (alt) {
    k = ctx->keys[1].as<uint64_t>();
    v = ctx->vals[1].as<double>();
  }
  int32_t* flag = ctx->stage[11].as<int32_t>();
  int32_t* off = flag + n + 1;
  SPWS_LAUNCH(ctx, "ws_heads", k_ws_heads, dim3(grid_of(n)), dim3(256), 0, k, n, flag);
  if ((rc = exclusive_scan<int32_t, int32_t>(ctx, flag, off, n))) return rc;
  SPWS_LAUNCH(ctx, "ws_reduce", k_ws_reduce, dim3(grid_of(n)), dim3(256), 0, k, v, n, flag, off, c, c + n, cv);
  int32_t* hm = ctx->host_meta.as<int32_t>();
  SPWS_HIP(hipMemcpyAsync(hm, off + n, 4, hipMemcpyDeviceToHost, s));
  SPWS_HIP(hipStreamSynchronize(s));
  const int64_t u = *hm;
  SPWS_HIP(hipMemcpyAsync(coo_crd0, c, (size_t)u * 4, hipMemcpyDeviceToHost, s));
  SPWS_HIP(hipMemcpyAsync(coo_crd1, c + n, (size_t)u * 4, hipMemcpyDeviceToHost, s));
  SPWS_HIP(hipMemcpyAsync(coo_vals, cv, (size_t)u * 8, hipMemcpyDeviceToHost, s));
  SPWS_HIP(hipStreamSynchronize(s));
  *out_size = u;
  return SPWS_OK;
}

}  // extern "C"
