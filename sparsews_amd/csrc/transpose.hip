// CSR transpose on the device, the second half of the reference's non-fused
// transposed product (benchmark/CSC_CSR_T/nonfuse/CSC_CSR_T_hash.h:352-400,
// transpose(A, B) with A = B^T): there the entries of each column are
// counted, the counts prefixed into the output row pointers, and the input
// rows scattered in row order, so every output row lists its columns (the
// input rows) in increasing order.
//
// Here each entry becomes a key (column << 32 | row) carrying its value; a
// stable LSD radix sort on the column bits alone keeps the row order of
// equal columns (the input is row-major).  The column counts come from the
// sorted keys: the last entry of a column's run adds its end position, the
// first subtracts its start, so a column takes at most two uncontended
// atomics instead of one per entry; their prefix sum gives the output row
// pointers.  Pure data movement: the output is bit-identical to the
// reference's.
#include "radix.hpp"

namespace spws {
namespace {

__global__ __launch_bounds__(256) void k_tr_zero(int32_t n, int32_t* __restrict__ counts) {
  for (int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) counts[c] = 0;
}

// one wavefront per input row: keys and values
__global__ __launch_bounds__(256) void k_tr_keys(int32_t m, const int32_t* __restrict__ pos,
                                                 const int32_t* __restrict__ crd, const double* __restrict__ vals,
                                                 uint64_t* __restrict__ keys, double* __restrict__ kv) {
  const int lane = lane_id();
  const int32_t waves = gridDim.x * (blockDim.x / kWave);
  for (int32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < m; i += waves) {
    const int32_t b = pos[i], e = pos[i + 1];
    for (int32_t p = b + lane; p < e; p += kWave) {
      keys[p] = ((uint64_t)(uint32_t)crd[p] << 32) | (uint32_t)i;
      kv[p] = vals[p];
    }
  }
}

// sorted keys -> the transpose's column indices (the input rows), values and
// per-column counts (run end + 1 minus run start, added by the run's first
// and last entries)
__global__ __launch_bounds__(256) void k_tr_out(int64_t nnz, const uint64_t* __restrict__ keys,
                                                const double* __restrict__ kv, int32_t* __restrict__ counts,
                                                int32_t* __restrict__ out_crd, double* __restrict__ out_vals) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t key = keys[p];
    const uint32_t c = (uint32_t)(key >> 32);
    const bool first = p == 0 || (uint32_t)(keys[p - 1] >> 32) != c;
    const bool last = p == nnz - 1 || (uint32_t)(keys[p + 1] >> 32) != c;
    if (first || last) atomicAdd(&counts[c], (last ? (int32_t)(p + 1) : 0) - (first ? (int32_t)p : 0));
    out_crd[p] = (int32_t)(uint32_t)key;
    out_vals[p] = kv[p];
  }
}

}  // namespace
}  // namespace spws

using namespace spws;

extern "C" int spws_csr_transpose_dev(spws_ctx* ctx, int32_t m, int32_t n, int64_t nnz, const int32_t* pos,
                                      const int32_t* crd, const double* vals, int32_t* out_pos, int32_t* out_crd,
                                      double* out_vals) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_REQUIRE(m >= 0 && n >= 0 && nnz >= 0, SPWS_ERR_INVALID, "negative dimension or nnz");
  SPWS_REQUIRE(nnz <= INT32_MAX, SPWS_ERR_UNSUPPORTED, "nnz = %lld exceeds the int32 index range", (long long)nnz);
  SPWS_REQUIRE(pos && out_pos && (nnz == 0 || (crd && vals && out_crd && out_vals)), SPWS_ERR_INVALID,
               "null pointer argument");
  const size_t nn = (size_t)(nnz > 0 ? nnz : 1);
  if ((rc = ctx->tr_keys.ensure(2 * nn * 8)) || (rc = ctx->tr_vals.ensure(2 * nn * 8)) ||
      (rc = ctx->tr_counts.ensure((size_t)(n > 0 ? n : 1) * 4)))
    return rc;
  uint64_t* keys = ctx->tr_keys.as<uint64_t>();
  double* kv = ctx->tr_vals.as<double>();
  int32_t* counts = ctx->tr_counts.as<int32_t>();
  if (n > 0)
    SPWS_LAUNCH(ctx, "tr_zero", k_tr_zero, dim3((unsigned)std::min<int64_t>(div_up(n, 256), 4096)), dim3(256), 0, n,
                counts);
  if (m > 0 && nnz > 0) {
    SPWS_LAUNCH(ctx, "tr_keys", k_tr_keys, dim3((unsigned)std::min<int64_t>(div_up(m, 4), 65536)), dim3(256), 0, m,
                pos, crd, vals, keys, kv);
    // column bits: enough for n - 1
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < (int64_t)n) ++bits;
    bool in_alt = false;
    if ((rc = radix_sort_pairs<uint64_t>(ctx, keys, kv, keys + nn, kv + nn, nnz, 32, 32 + bits, &in_alt))) return rc;
    SPWS_LAUNCH(ctx, "tr_out", k_tr_out, dim3((unsigned)std::min<int64_t>(div_up(nnz, 256), 16384)), dim3(256), 0,
                nnz, in_alt ? keys + nn : keys, in_alt ? kv + nn : kv, counts, out_crd, out_vals);
  }
  return exclusive_scan<int32_t, int32_t>(ctx, counts, out_pos, n);
}
