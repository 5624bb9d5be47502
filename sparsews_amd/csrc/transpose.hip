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
#include <mutex>

#include "radix.hpp"

namespace spws {
namespace {

// bits that hold every column index below n
int column_bits(int32_t n) {
  int bits = 1;
  while (bits < 31 && ((int64_t)1 << bits) < (int64_t)n) ++bits;
  return bits;
}

__global__ __launch_bounds__(256) void k_tr_zero(int32_t n, int32_t* __restrict__ counts) {
  for (int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) counts[c] = 0;
}

// one wavefront per input row: keys and values (row pointers int32, or
// int64 as the outer product returns them; nnz stays below 2^31 either way)
template <typename PosT>
__global__ __launch_bounds__(256) void k_tr_keys(int32_t m, const PosT* __restrict__ pos,
                                                 const int32_t* __restrict__ crd, const double* __restrict__ vals,
                                                 uint64_t* __restrict__ keys, double* __restrict__ kv) {
  const int lane = lane_id();
  const int32_t waves = gridDim.x * (blockDim.x / kWave);
  for (int32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < m; i += waves) {
    const int32_t b = (int32_t)pos[i], e = (int32_t)pos[i + 1];
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

// The transpose on a context: its workspaces are the context's buffers
// (tr_keys, tr_vals, tr_counts, and the radix and scan scratch).
template <typename PosT>
static int transpose_run(spws_ctx* ctx, int32_t m, int32_t n, int64_t nnz, const PosT* pos, const int32_t* crd,
                         const double* vals, int32_t* out_pos, int32_t* out_crd, double* out_vals) {
  int rc;
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
    SPWS_LAUNCH(ctx, "tr_keys", k_tr_keys<PosT>, dim3((unsigned)std::min<int64_t>(div_up(m, 4), 65536)), dim3(256), 0, m,
                pos, crd, vals, keys, kv);
    bool in_alt = false;
    if ((rc = radix_sort_pairs<uint64_t>(ctx, keys, kv, keys + nn, kv + nn, nnz, 32, 32 + column_bits(n), &in_alt)))
      return rc;
    SPWS_LAUNCH(ctx, "tr_out", k_tr_out, dim3((unsigned)std::min<int64_t>(div_up(nnz, 256), 16384)), dim3(256), 0,
                nnz, in_alt ? keys + nn : keys, in_alt ? kv + nn : kv, counts, out_crd, out_vals);
  }
  return exclusive_scan<int32_t, int32_t>(ctx, counts, out_pos, n);
}

static int check_args(int32_t m, int32_t n, int64_t nnz, const void* pos, const int32_t* crd, const double* vals,
                      const int32_t* out_pos, const int32_t* out_crd, const double* out_vals) {
  SPWS_REQUIRE(m >= 0 && n >= 0 && nnz >= 0, SPWS_ERR_INVALID, "negative dimension or nnz");
  SPWS_REQUIRE(nnz <= INT32_MAX, SPWS_ERR_UNSUPPORTED, "nnz = %lld exceeds the int32 index range", (long long)nnz);
  SPWS_REQUIRE(pos && out_pos && (nnz == 0 || (crd && vals && out_crd && out_vals)), SPWS_ERR_INVALID,
               "null pointer argument");
  return SPWS_OK;
}

// ------------------------------------------------ no-allocation entry point
//
// spws_csr_transpose_dev_noalloc runs transpose_run on a context that lives
// on the host stack and whose workspaces are fixed views into `scratch`, a
// file-scope device array: nothing is allocated on the host or the device.
// Each GPU holds its own copy of the array (at its own address) for the life
// of the process; its address on a device comes from hipGetSymbolAddress,
// taken once per device with that device current and cached here.  The
// array's name is only a host-side handle: the kernels get the looked-up
// address.  Calls that overlap on one GPU share the array and race.

constexpr size_t kNoAllocScratchBytes = (size_t)48 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

// The scratch array's address on `device`, which must be current.
static int scratch_address(int device, unsigned char** out) {
  SPWS_REQUIRE(device >= 0 && device < kMaxDevices, SPWS_ERR_UNSUPPORTED, "device %d beyond the scratch table",
               device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    SPWS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
  return SPWS_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of each workspace the transpose of nnz entries into n rows uses
// (keys, values, column counts, radix histograms, scan partials).
struct TransposeScratch {
  size_t keys, vals, counts, hist, scan, total;
  TransposeScratch(int64_t nnz, int32_t n) {
    const size_t nn = (size_t)(nnz > 0 ? nnz : 1);
    keys = align256(2 * nn * 8);
    vals = align256(2 * nn * 8);
    counts = align256((size_t)(n > 0 ? n : 1) * 4);
    const size_t hn = radix_hist_entries((int64_t)nn);
    hist = align256((2 * hn + 1) * 4);
    scan = align256(std::max(scan_scratch_bytes((int64_t)hn, 4), scan_scratch_bytes((int64_t)n, 4)) + 4);
    total = keys + vals + counts + hist + scan;
  }
};

}  // namespace

int noalloc_scratch_init(int device) {
  unsigned char* p;
  return scratch_address(device, &p);
}

}  // namespace spws

using namespace spws;

extern "C" int spws_csr_transpose_dev(spws_ctx* ctx, int32_t m, int32_t n, int64_t nnz, const int32_t* pos,
                                      const int32_t* crd, const double* vals, int32_t* out_pos, int32_t* out_crd,
                                      double* out_vals) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  if ((rc = check_args(m, n, nnz, pos, crd, vals, out_pos, out_crd, out_vals))) return rc;
  return transpose_run(ctx, m, n, nnz, pos, crd, vals, out_pos, out_crd, out_vals);
}

extern "C" int spws_csr64_transpose_dev(spws_ctx* ctx, int32_t m, int32_t n, int64_t nnz, const int64_t* pos,
                                        const int32_t* crd, const double* vals, int32_t* out_pos, int32_t* out_crd,
                                        double* out_vals) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  if ((rc = check_args(m, n, nnz, pos, crd, vals, out_pos, out_crd, out_vals))) return rc;
  return transpose_run(ctx, m, n, nnz, pos, crd, vals, out_pos, out_crd, out_vals);
}

extern "C" int spws_csr_transpose_dev_noalloc(void* stream, int32_t m, int32_t n, int64_t nnz, const int32_t* pos,
                                              const int32_t* crd, const double* vals, int32_t* out_pos,
                                              int32_t* out_crd, double* out_vals) {
  int rc = check_args(m, n, nnz, pos, crd, vals, out_pos, out_crd, out_vals);
  if (rc) return rc;
  const TransposeScratch need(nnz, n);
  SPWS_REQUIRE(need.total <= kNoAllocScratchBytes, SPWS_ERR_UNSUPPORTED,
               "the no-allocation transpose of %lld entries into %d rows needs %zu bytes of scratch; the fixed "
               "scratch holds %zu (about 33 bytes per entry plus 4 per row)",
               (long long)nnz, n, need.total, kNoAllocScratchBytes);
  int device = 0;
  SPWS_HIP(hipGetDevice(&device));
  unsigned char* base;
  if ((rc = scratch_address(device, &base))) return rc;
  spws_ctx local;  // host stack: stream and fixed workspace views only
  local.device = device;
  local.stream = static_cast<hipStream_t>(stream);
  size_t off = 0;
  local.tr_keys.borrow(base + off, need.keys);
  off += need.keys;
  local.tr_vals.borrow(base + off, need.vals);
  off += need.vals;
  local.tr_counts.borrow(base + off, need.counts);
  off += need.counts;
  local.radix_hist.borrow(base + off, need.hist);
  off += need.hist;
  local.scan_tmp.borrow(base + off, need.scan);
  return transpose_run(&local, m, n, nnz, pos, crd, vals, out_pos, out_crd, out_vals);
}
