// Batched row-wise SpGEMM: many products of uneven sizes in one hipGraph
// launch, and the request loop (plain and continuously batched) on top.
//
// A batch stacks its requests into one block-diagonal product: request b's
// A rows occupy rows [row0_b, row0_b + m_b) of the batch, its B rows occupy
// [brow0_b, brow0_b + k_b), and its output columns are shifted by col0_b.
// Row-wise SpGEMM of blockdiag(A_b) * blockdiag(B_b) is blockdiag(A_b * B_b),
// so the row binning of the single-product path load-balances across all
// requests at once, and the arithmetic of every entry is unchanged.
//
// The pipeline is captured once per engine into a hipGraph whose kernels read
// every size from a device-side descriptor and request table; a launch copies
// the (pinned) table and descriptor in, so the same graph executes any batch
// that fits the engine's capacities:
//   H2D request table + descriptor
//   k_batch_gather     operands -> engine pools, indices rebased
//   k_meta_clear, k_rowstat, device-sized scan, k_bin_offsets, k_binscatter
//   numeric kernels of the wave / wave2 / block / block2 / wide paths
//   (persistent grids; each reads its row count from the bin metadata; the
//   block bins use the LDS bitmap kernel over each row's own column range
//   when the engine bounds the request width by max_cols)
//   device-sized scan -> C_pos
//   k_batch_req_meta   each request's nnz (and whether its output buffer holds it)
//   k_batch_compact    rows into their request's output buffer when it has
//                      one, into the engine's pools otherwise (columns
//                      shifted back per row)
//   k_batch_total, D2H of the batch nnz and of each request's nnz
// No memset is recorded into the graph: clears are kernels.  The dense path
// is not used in batches: the stacked column space is the sum of the
// requests' widths, so long rows take the HBM hash path instead.
//
// Admission trusts the caller's upper bounds (products, max_row_products).
// The graph recomputes them on the device (k_rowstat) and k_batch_guard
// compares them with the engine's pools before any kernel writes into them:
// when a bound was under-reported, the batch's row count is set to zero, so
// every later kernel has nothing to do, and the batch reports an error.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <thread>
#include <vector>

#include "rowwise_kernels.hpp"

struct spws_batch_req {  // device-visible per-request descriptor
  const int32_t* A_pos;
  const int32_t* A_crd;
  const double* A_val;
  const int32_t* B_pos;
  const int32_t* B_crd;
  const double* B_val;
  int32_t m, k, n;
  int32_t row0, brow0, col0;
  int64_t annz0, bnnz0, annz, bnnz;
  // optional output (device): row pointers [m + 1], columns and values
  // [out_cap]; null out_pos = the result stays in the engine's pools
  int32_t* out_pos;
  int32_t* out_crd;
  double* out_val;
  int64_t out_cap;
};

struct spws_batch_desc {
  int32_t nreq, rows, brows, cols;
  int64_t annz, bnnz;
  int32_t overflow, pad;  // set on the device when an upper bound was under-reported
};

enum BatchPool : int {
  BP_REQ = 0, BP_DESC, BP_APOS, BP_ACRD, BP_AVAL, BP_BPOS, BP_BCRD, BP_BVAL, BP_COLOFF, BP_ROWUB, BP_ROWOFF,
  BP_ROWNNZ, BP_BINROWS, BP_META, BP_SCRD, BP_SVAL, BP_CPOS, BP_CCRD, BP_CVAL, BP_WKEY, BP_WVAL, BP_SCAN, BP_TOTAL,
  BP_REQNNZ, BP_COUNT
};

struct spws_batch {
  spws_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  spws_batch_caps caps{};
  spws::DevBuf pool[BP_COUNT];
  spws_batch_req* h_req = nullptr;
  spws_batch_desc* h_desc = nullptr;
  int64_t* h_total = nullptr;
  int64_t* h_req_nnz = nullptr;  // per request: output nnz (-1: out_cap too small, -2: overflow)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;
  bool in_flight = false;
  spws_batch_desc cur{};
  int64_t cur_ub = 0;
  int64_t wide_slots = 0;
  int32_t wide_groups = 0;
  spws::PathParams pp{};
};

namespace spws {
namespace {

constexpr int kDScanBlocks = 256;
constexpr int kDScanThreads = 256;
constexpr int kDScanItems = 16;
constexpr int kDScanTile = kDScanThreads * kDScanItems;

__device__ __forceinline__ void dscan_range(int64_t n, int64_t* beg, int64_t* end) {
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  *beg = (int64_t)blockIdx.x * chunk;
  *end = *beg + chunk < n ? *beg + chunk : n;
  if (*beg > n) *beg = n;
}

// Exclusive scan whose length lives in device memory: each of a fixed number
// of blocks owns a contiguous range (reduce, scan of the partials, rescan).
template <class Tin, class Tout>
__global__ __launch_bounds__(kDScanThreads) void k_dscan_reduce(const Tin* __restrict__ in,
                                                                const int32_t* __restrict__ n_dev,
                                                                Tout* __restrict__ part) {
  __shared__ Tout sh[kDScanThreads / kWave + 1];
  int64_t beg, end;
  dscan_range(*n_dev, &beg, &end);
  Tout s = 0;
  for (int64_t i = beg + threadIdx.x; i < end; i += kDScanThreads) s += (Tout)in[i];
  s = wave_sum(s);
  if (lane_id() == 0) sh[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    Tout t = 0;
    for (int w = 0; w < kDScanThreads / kWave; ++w) t += sh[w];
    part[blockIdx.x] = t;
  }
}

template <class Tout>
__global__ __launch_bounds__(kDScanThreads) void k_dscan_top(Tout* __restrict__ part) {
  __shared__ Tout sh[kDScanThreads / kWave + 1];
  static_assert(kDScanBlocks == kDScanThreads, "one partial per thread");
  const Tout v = part[threadIdx.x];
  Tout total;
  const Tout pre = block_excl_scan<Tout, kDScanThreads>(v, sh, &total);
  part[threadIdx.x] = pre;
  if (threadIdx.x == 0) part[kDScanBlocks] = total;
}

template <class Tin, class Tout>
__global__ __launch_bounds__(kDScanThreads) void k_dscan_apply(const Tin* __restrict__ in,
                                                               const int32_t* __restrict__ n_dev,
                                                               const Tout* __restrict__ part, Tout* __restrict__ out) {
  __shared__ Tout sh[kDScanThreads / kWave + 1];
  const int64_t n = *n_dev;
  int64_t beg, end;
  dscan_range(n, &beg, &end);
  Tout carry = part[blockIdx.x];
  for (int64_t base = beg; base < end; base += kDScanTile) {
    Tout v[kDScanItems];
    Tout local = 0;
    const int64_t mine = base + (int64_t)threadIdx.x * kDScanItems;
#pragma unroll
    for (int it = 0; it < kDScanItems; ++it) {
      const int64_t idx = mine + it;
      v[it] = idx < end ? (Tout)in[idx] : (Tout)0;
      local += v[it];
    }
    Tout total;
    const Tout pre = block_excl_scan<Tout, kDScanThreads>(local, sh, &total);
    Tout run = carry + pre;
#pragma unroll
    for (int it = 0; it < kDScanItems; ++it) {
      const int64_t idx = mine + it;
      if (idx < end) out[idx] = run;
      run += v[it];
    }
    carry += total;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = part[kDScanBlocks];
}

// dst_crd[p] = crd[p] + shift, dst_val[p] = val[p] for p = t0, t0 + st, ...:
// kGatherUnroll strides loaded before any is stored, so a large request's
// copy is bound by bandwidth rather than by one load latency per stride
constexpr int kGatherUnroll = 8;

__device__ __forceinline__ void copy_rebased(const int32_t* __restrict__ crd, const double* __restrict__ val,
                                             int64_t n, int32_t shift, int32_t* __restrict__ dst_crd,
                                             double* __restrict__ dst_val, int64_t t0, int64_t st) {
  for (int64_t p0 = t0; p0 < n; p0 += st * kGatherUnroll) {
    int32_t c[kGatherUnroll];
    double v[kGatherUnroll];
#pragma unroll
    for (int u = 0; u < kGatherUnroll; ++u) {
      const int64_t p = p0 + u * st;
      if (p < n) {
        c[u] = crd[p];
        v[u] = val[p];
      }
    }
#pragma unroll
    for (int u = 0; u < kGatherUnroll; ++u) {
      const int64_t p = p0 + u * st;
      if (p < n) {
        dst_crd[p] = c[u] + shift;
        dst_val[p] = v[u];
      }
    }
  }
}

// operands of every request into the pools, indices rebased; per row, the
// request's first column (row_coloff[row]), width (row_coloff[max_rows +
// row]) and index (row_coloff[2 * max_rows + row])
__global__ __launch_bounds__(256) void k_batch_gather(const spws_batch_req* __restrict__ reqs,
                                                      const spws_batch_desc* __restrict__ desc,
                                                      int32_t* __restrict__ A_pos, int32_t* __restrict__ A_crd,
                                                      double* __restrict__ A_val, int32_t* __restrict__ B_pos,
                                                      int32_t* __restrict__ B_crd, double* __restrict__ B_val,
                                                      int32_t* __restrict__ row_coloff, int32_t max_rows) {
  const int32_t r = blockIdx.y;
  if (r >= desc->nreq) return;
  const spws_batch_req q = reqs[r];
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = t0; i <= q.m; i += st) {
    A_pos[q.row0 + i] = (int32_t)(q.annz0 + q.A_pos[i]);
    if (i < q.m) {
      row_coloff[q.row0 + i] = q.col0;
      row_coloff[max_rows + q.row0 + i] = q.n;  // the row's request width
      row_coloff[2 * max_rows + q.row0 + i] = r;
    }
  }
  for (int64_t i = t0; i <= q.k; i += st) B_pos[q.brow0 + i] = (int32_t)(q.bnnz0 + q.B_pos[i]);
  copy_rebased(q.A_crd, q.A_val, q.annz, q.brow0, A_crd + q.annz0, A_val + q.annz0, t0, st);
  copy_rebased(q.B_crd, q.B_val, q.bnnz, q.col0, B_crd + q.bnnz0, B_val + q.bnnz0, t0, st);
}

__global__ void k_batch_total(const int32_t* __restrict__ C_pos, const spws_batch_desc* __restrict__ desc,
                              int64_t* __restrict__ total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *total = desc->overflow ? -2 : C_pos[desc->rows];
}

// The true upper bounds (k_rowstat) against the pools: the sum of all row
// bounds must fit the slot and output pools, and no row may exceed the
// per-row capacity the hash-table pool was sized for.  On overflow the
// batch's row count becomes zero (every later kernel reads it), the bins
// are emptied and the overflow flag is set.
__global__ void k_batch_guard(spws_batch_desc* __restrict__ desc, int32_t* __restrict__ bmeta,
                              const unsigned long long* __restrict__ maxub, int64_t cap_products,
                              int64_t cap_row_products) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long mx = 0;
  for (int p = 0; p < P_EMPTY; ++p) mx = maxub[p] > mx ? maxub[p] : mx;
  if (maxub[8] > (unsigned long long)cap_products || mx > (unsigned long long)cap_row_products) {
    desc->rows = 0;
    desc->overflow = 1;
    for (int p = 0; p < 8; ++p) bmeta[kMetaCount + p] = 0;
  }
}

// Per request, after the output row pointers are known: req_nnz[r] = its
// nnz, -1 when it has an output buffer smaller than that (its rows then stay
// in the pools), -2 when the batch overflowed.  A request with no rows gets
// its single row pointer here.
__global__ __launch_bounds__(256) void k_batch_req_meta(const spws_batch_req* __restrict__ reqs,
                                                        const spws_batch_desc* __restrict__ desc,
                                                        const int32_t* __restrict__ C_pos,
                                                        int64_t* __restrict__ req_nnz) {
  for (int32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < desc->nreq; r += gridDim.x * blockDim.x) {
    if (desc->overflow) {
      req_nnz[r] = -2;
      continue;
    }
    const spws_batch_req& q = reqs[r];
    const int64_t nnz = (int64_t)C_pos[q.row0 + q.m] - C_pos[q.row0];
    const bool fits = nnz <= q.out_cap;
    req_nnz[r] = (q.out_pos && !fits) ? -1 : nnz;
    if (q.out_pos && fits && q.m == 0) q.out_pos[0] = 0;
  }
}

// Slot regions -> contiguous rows, a wavefront per row (k_compact's loop),
// with each row written straight into its request's output buffer when it
// has one that holds the request (row pointers rebased to the request's
// first entry; the request's last row also writes the closing pointer), and
// into the engine's output pools otherwise.
__global__ __launch_bounds__(256) void k_batch_compact(const spws_batch_desc* __restrict__ desc,
                                                       const int64_t* __restrict__ row_off,
                                                       const int32_t* __restrict__ C_pos,
                                                       const int32_t* __restrict__ slot_crd,
                                                       const double* __restrict__ slot_val,
                                                       const int32_t* __restrict__ row_coloff,
                                                       const int32_t* __restrict__ row_req,
                                                       const spws_batch_req* __restrict__ reqs,
                                                       const int64_t* __restrict__ req_nnz,
                                                       int32_t* __restrict__ C_crd, double* __restrict__ C_val) {
  const int32_t m = desc->rows;
  const int lane = lane_id();
  const int32_t waves = gridDim.x * (blockDim.x / kWave);
  for (int32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < m; i += waves) {
    const int32_t dst = C_pos[i];
    const int32_t n = C_pos[i + 1] - dst;
    const int64_t src = row_off[i];
    const int32_t shift = row_coloff[i];
    const int32_t r = row_req[i];
    int32_t* oc = C_crd + dst;
    double* ov = C_val + dst;
    const spws_batch_req& q = reqs[r];
    if (q.out_pos && req_nnz[r] >= 0) {
      const int32_t start = C_pos[q.row0];
      oc = q.out_crd + (dst - start);
      ov = q.out_val + (dst - start);
      if (lane == 0) {
        q.out_pos[i - q.row0] = dst - start;
        if (i == q.row0 + q.m - 1) q.out_pos[q.m] = dst + n - start;
      }
    }
    for (int32_t e0 = 0; e0 < n; e0 += kWave * kCompactUnroll) {
      int32_t c[kCompactUnroll];
      double v[kCompactUnroll];
#pragma unroll
      for (int u = 0; u < kCompactUnroll; ++u) {
        const int32_t e = e0 + u * kWave + lane;
        if (e < n) {
          c[u] = slot_crd[src + e];
          v[u] = slot_val[src + e];
        }
      }
#pragma unroll
      for (int u = 0; u < kCompactUnroll; ++u) {
        const int32_t e = e0 + u * kWave + lane;
        if (e < n) {
          oc[e] = c[u] - shift;
          ov[e] = v[u];
        }
      }
    }
  }
}

}  // namespace

template <class Tin, class Tout>
static void dscan(hipStream_t s, const Tin* in, const int32_t* n_dev, Tout* part, Tout* out) {
  hipLaunchKernelGGL((k_dscan_reduce<Tin, Tout>), dim3(kDScanBlocks), dim3(kDScanThreads), 0, s, in, n_dev, part);
  hipLaunchKernelGGL((k_dscan_top<Tout>), dim3(1), dim3(kDScanThreads), 0, s, part);
  hipLaunchKernelGGL((k_dscan_apply<Tin, Tout>), dim3(kDScanBlocks), dim3(kDScanThreads), 0, s, in, n_dev, part, out);
}

// Enqueue the whole batch pipeline on b->stream (captured into the graph).
static void enqueue_batch(spws_batch* b) {
  hipStream_t s = b->stream;
  const spws_batch_caps& c = b->caps;
  auto P = [&](int i) { return b->pool[i].ptr; };
  spws_batch_req* d_req = (spws_batch_req*)P(BP_REQ);
  spws_batch_desc* d_desc = (spws_batch_desc*)P(BP_DESC);
  (void)hipMemcpyAsync(d_req, b->h_req, sizeof(spws_batch_req) * (size_t)c.max_requests, hipMemcpyHostToDevice, s);
  (void)hipMemcpyAsync(d_desc, b->h_desc, sizeof(spws_batch_desc), hipMemcpyHostToDevice, s);
  int32_t* Ap = (int32_t*)P(BP_APOS);
  int32_t* Ai = (int32_t*)P(BP_ACRD);
  double* Ax = (double*)P(BP_AVAL);
  int32_t* Bp = (int32_t*)P(BP_BPOS);
  int32_t* Bi = (int32_t*)P(BP_BCRD);
  double* Bx = (double*)P(BP_BVAL);
  int32_t* coloff = (int32_t*)P(BP_COLOFF);
  hipLaunchKernelGGL(k_batch_gather, dim3(64, c.max_requests), dim3(256), 0, s, d_req, d_desc, Ap, Ai, Ax, Bp, Bi, Bx,
                     coloff, (int32_t)c.max_rows);
  const int32_t* row_ncols = coloff + c.max_rows;
  const int32_t* m_dev = &d_desc->rows;
  int64_t* row_ub = (int64_t*)P(BP_ROWUB);
  int64_t* row_off = (int64_t*)P(BP_ROWOFF);
  int32_t* row_nnz = (int32_t*)P(BP_ROWNNZ);
  int32_t* bin_rows = (int32_t*)P(BP_BINROWS);
  int32_t* bmeta = (int32_t*)P(BP_META);
  unsigned long long* maxub = (unsigned long long*)(bmeta + 32);
  int32_t* scrd = (int32_t*)P(BP_SCRD);
  double* sval = (double*)P(BP_SVAL);
  int32_t* Cp = (int32_t*)P(BP_CPOS);
  int32_t* Ci = (int32_t*)P(BP_CCRD);
  double* Cx = (double*)P(BP_CVAL);
  const unsigned grows = (unsigned)std::min<int64_t>(div_up(c.max_rows, 256), 4096);
  hipLaunchKernelGGL(k_meta_clear, dim3(1), dim3(64), 0, s, bmeta, maxub);
  const unsigned gstat = (unsigned)std::min<int64_t>(div_up(c.max_rows, kRsRows), 16384);
  hipLaunchKernelGGL(k_rowstat, dim3(gstat), dim3(kRsBlock), 0, s, 0, m_dev, 1, Ap, Ai, Bp, row_ub, bmeta, maxub,
                     b->pp);
  hipLaunchKernelGGL(k_batch_guard, dim3(1), dim3(64), 0, s, d_desc, bmeta, maxub, c.max_products,
                     c.max_row_products);
  dscan<int64_t, int64_t>(s, row_ub, m_dev, (int64_t*)P(BP_SCAN), row_off);
  hipLaunchKernelGGL(k_bin_offsets, dim3(1), dim3(64), 0, s, bmeta);
  hipLaunchKernelGGL(k_binscatter, dim3(grows), dim3(256), 0, s, 0, m_dev, 1, row_ub, bmeta, bin_rows, row_nnz, b->pp);
  const unsigned gwave = (unsigned)std::min<int64_t>(c.max_rows, 8192);
  const unsigned gblk = (unsigned)std::min<int64_t>(c.max_rows, 1024);
  // the block bins' row-per-workgroup kernel: a wider grid keeps more rows in flight
  const unsigned gbin = (unsigned)std::min<int64_t>(c.max_rows, 8192);
  hipLaunchKernelGGL((k_row_wave<128, 1, true, kWaveRowsK>), dim3(gwave), dim3(64), 0, s, bin_rows, bmeta, (int)P_WAVE,
                     Ap, Ai, Ax, Bp, Bi, Bx, row_off, scrd, sval, row_nnz);
  hipLaunchKernelGGL((k_row_wave<512, 4, true>), dim3(gwave), dim3(64), 0, s, bin_rows, bmeta, (int)P_WAVE2, Ap, Ai,
                     Ax, Bp, Bi, Bx, row_off, scrd, sval, row_nnz);
  if (c.max_cols > 0 && c.max_cols <= kBitmapCols) {
    // every request is at most max_cols wide: LDS bitmap over each row's own column range
    const int64_t v1 = std::min<int64_t>(b->pp.block_ub, c.max_cols), v2 = std::min<int64_t>(b->pp.block2_ub, c.max_cols);
    hipLaunchKernelGGL((k_row_bitmap_rows<kBitmapRowsK, kBitmapItems1>), dim3(gbin), dim3(kBitmapBlock),
                       bitmap_lds_bytes(c.max_cols, v1, true, kBitmapBlock), s,
                       bin_rows, bmeta, (int)P_BLOCK, c.max_cols, Ap, Ai, Ax, Bp, Bi, Bx, row_off, scrd, sval, row_nnz,
                       coloff, row_ncols);
    hipLaunchKernelGGL((k_row_bitmap_rows<kBitmapRowsK, kBitmapItems2>), dim3(gbin), dim3(kBitmapBlock),
                       bitmap_lds_bytes(c.max_cols, v2, true, kBitmapBlock), s,
                       bin_rows, bmeta, (int)P_BLOCK2, c.max_cols, Ap, Ai, Ax, Bp, Bi, Bx, row_off, scrd, sval,
                       row_nnz, coloff, row_ncols);
  } else {
    hipLaunchKernelGGL((k_row_block<2048, 1024, 256, true>), dim3(gblk), dim3(256), 0, s, bin_rows, bmeta,
                       (int)P_BLOCK, Ap, Ai, Ax, Bp, Bi, Bx, row_off, scrd, sval, row_nnz);
    hipLaunchKernelGGL((k_row_block<8192, 4096, 512, true>), dim3(gblk), dim3(512), 0, s, bin_rows, bmeta,
                       (int)P_BLOCK2, Ap, Ai, Ax, Bp, Bi, Bx, row_off, scrd, sval, row_nnz);
  }
  if (c.max_cols > 0 && c.max_cols <= kBitmapCols) {
    // products staged in the output pools, which the compaction fills later;
    // rows of at most acc_cap entries accumulate in LDS
    const size_t bm_bytes = bitmap_lds_bytes(c.max_cols, 0, false, kBitmapLongBlock);
    const int32_t acc_cap = (int32_t)std::min<int64_t>(c.max_cols, (int64_t)((kLdsBudget - bm_bytes) / 8));
    hipLaunchKernelGGL(k_row_bitmap_staged<kBitmapLongBlock>, dim3(gblk), dim3(kBitmapLongBlock),
                       bitmap_lds_bytes(c.max_cols, acc_cap, true, kBitmapLongBlock), s, bin_rows, bmeta,
                       (int)P_WIDE, c.max_cols, Ap, Ai, Ax, Bp, Bi, Bx, row_off, scrd, sval, row_nnz, coloff,
                       row_ncols, Ci, Cx, acc_cap);
  } else {
    hipLaunchKernelGGL((k_row_wide<true>), dim3(b->wide_groups), dim3(kWideBlock), 0, s, bin_rows, bmeta,
                       (int)P_WIDE, Ap, Ai, Ax, Bp, Bi, Bx, row_ub, row_off, scrd, sval, row_nnz,
                       (int32_t*)P(BP_WKEY), (double*)P(BP_WVAL), b->wide_slots);
  }
  dscan<int32_t, int32_t>(s, row_nnz, m_dev, (int32_t*)P(BP_SCAN), Cp);
  int64_t* req_nnz = (int64_t*)P(BP_REQNNZ);
  hipLaunchKernelGGL(k_batch_req_meta, dim3((unsigned)div_up(c.max_requests, 256)), dim3(256), 0, s, d_req, d_desc,
                     Cp, req_nnz);
  hipLaunchKernelGGL(k_batch_compact, dim3((unsigned)std::min<int64_t>(div_up(c.max_rows, 4), 16384)), dim3(256), 0,
                     s, d_desc, row_off, Cp, scrd, sval, coloff, coloff + 2 * (size_t)c.max_rows, d_req, req_nnz, Ci,
                     Cx);
  hipLaunchKernelGGL(k_batch_total, dim3(1), dim3(64), 0, s, Cp, d_desc, (int64_t*)P(BP_TOTAL));
  (void)hipMemcpyAsync(b->h_total, P(BP_TOTAL), 8, hipMemcpyDeviceToHost, s);
  (void)hipMemcpyAsync(b->h_req_nnz, req_nnz, 8 * (size_t)c.max_requests, hipMemcpyDeviceToHost, s);
}

static void batch_release(spws_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->done) (void)hipEventDestroy(b->done);
  for (DevBuf& p : b->pool) p.release();
  if (b->h_req) (void)hipHostFree(b->h_req);
  if (b->h_desc) (void)hipHostFree(b->h_desc);
  if (b->h_total) (void)hipHostFree(b->h_total);
  if (b->h_req_nnz) (void)hipHostFree(b->h_req_nnz);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

static int batch_build(spws_batch* b) {
  const spws_batch_caps& c = b->caps;
  const size_t rows = (size_t)c.max_rows, brows = (size_t)c.max_brows;
  const size_t annz = (size_t)c.max_annz, bnnz = (size_t)c.max_bnnz, ub = (size_t)c.max_products;
  b->wide_slots = std::max<int64_t>(64, next_pow2_u32((uint32_t)(2 * c.max_row_products)));
  b->wide_groups = 64;
  const size_t sizes[BP_COUNT] = {
      sizeof(spws_batch_req) * (size_t)c.max_requests, sizeof(spws_batch_desc), (rows + 1) * 4, annz * 4, annz * 8,
      (brows + 1) * 4, bnnz * 4, bnnz * 8, rows * 12, rows * 8, (rows + 1) * 8, rows * 4, rows * 4, 64 * 8, ub * 4,
      ub * 8, (rows + 1) * 4, ub * 4, ub * 8, (size_t)b->wide_groups * b->wide_slots * 4,
      (size_t)b->wide_groups * b->wide_slots * 8, (kDScanBlocks + 1) * 8, 8, (size_t)c.max_requests * 8};
  int rc;
  for (int i = 0; i < BP_COUNT; ++i)
    if ((rc = b->pool[i].ensure(sizes[i] > 0 ? sizes[i] : 8))) return rc;
  SPWS_HIP(hipHostMalloc((void**)&b->h_req, sizeof(spws_batch_req) * (size_t)c.max_requests, hipHostMallocDefault));
  SPWS_HIP(hipHostMalloc((void**)&b->h_desc, sizeof(spws_batch_desc), hipHostMallocDefault));
  SPWS_HIP(hipHostMalloc((void**)&b->h_total, 8, hipHostMallocDefault));
  SPWS_HIP(hipHostMalloc((void**)&b->h_req_nnz, 8 * (size_t)c.max_requests, hipHostMallocDefault));
  memset(b->h_req, 0, sizeof(spws_batch_req) * (size_t)c.max_requests);
  memset(b->h_desc, 0, sizeof(spws_batch_desc));
  SPWS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  SPWS_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
  // SPWS_BATCH_EAGER=1: no graph; every launch enqueues the same kernels
  // on the engine's stream (for kernel tracers that do not follow graphs)
  const char* eager = getenv("SPWS_BATCH_EAGER");
  if (eager && eager[0] == '1') return SPWS_OK;
  SPWS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  enqueue_batch(b);
  hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
  if (e != hipSuccess) return fail_hip(e, "hipStreamEndCapture (batch pipeline)", __FILE__, __LINE__);
  SPWS_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return SPWS_OK;
}

}  // namespace spws

using namespace spws;

extern "C" {

int spws_batch_create(spws_ctx* ctx, const spws_batch_caps* caps, spws_batch** out) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_REQUIRE(caps && out, SPWS_ERR_INVALID, "null argument");
  SPWS_REQUIRE(caps->max_requests >= 1 && caps->max_requests <= 65535 && caps->max_rows >= 1 &&
                   caps->max_brows >= 1 && caps->max_annz >= 1 && caps->max_bnnz >= 1 && caps->max_products >= 1 &&
                   caps->max_row_products >= 1,
               SPWS_ERR_INVALID, "batch capacities must be positive (max_requests <= 65535)");
  SPWS_REQUIRE(caps->max_products <= INT32_MAX && caps->max_annz <= INT32_MAX && caps->max_bnnz <= INT32_MAX &&
                   caps->max_row_products <= (1 << 28),
               SPWS_ERR_UNSUPPORTED, "batch capacities exceed the int32 index range");
  SPWS_REQUIRE(caps->max_cols >= 0 && caps->max_cols <= kBitmapCols, SPWS_ERR_INVALID,
               "max_cols must be 0 (any width) or at most %d", kBitmapCols);
  spws_batch* b = new spws_batch();
  b->ctx = ctx;
  b->caps = *caps;
  const spws_rowwise_tuning& t = ctx->rw;
  // dense_ratio 0: the dense path is never chosen for the stacked product
  b->pp = PathParams{t.wave_ub, t.wave2_ub, t.block_ub, t.block2_ub, 0, -1};
  rc = batch_build(b);
  if (rc) {
    batch_release(b);
    delete b;
    return rc;
  }
  *out = b;
  return SPWS_OK;
}

int spws_batch_destroy(spws_batch* b) {
  if (!b) return SPWS_OK;
  (void)hipSetDevice(b->ctx->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_release(b);
  delete b;
  return SPWS_OK;
}

int spws_batch_reset(spws_batch* b) {
  SPWS_REQUIRE(b != nullptr, SPWS_ERR_INVALID, "null batch");
  SPWS_REQUIRE(!b->in_flight, SPWS_ERR_INVALID, "batch is still running");
  b->cur = spws_batch_desc{};
  b->cur_ub = 0;
  return SPWS_OK;
}

int spws_batch_add(spws_batch* b, int32_t m, int32_t k, int32_t n, const int32_t* A_pos, const int32_t* A_crd,
                   const double* A_vals, int64_t annz, const int32_t* B_pos, const int32_t* B_crd,
                   const double* B_vals, int64_t bnnz, int64_t products, int64_t max_row_products, int32_t* slot) {
  SPWS_REQUIRE(b != nullptr, SPWS_ERR_INVALID, "null batch");
  SPWS_REQUIRE(!b->in_flight, SPWS_ERR_INVALID, "batch is still running");
  SPWS_REQUIRE(m >= 0 && k >= 0 && n >= 0 && annz >= 0 && bnnz >= 0 && products >= 0, SPWS_ERR_INVALID,
               "negative size");
  const spws_batch_caps& c = b->caps;
  spws_batch_desc& d = b->cur;
  // memory-based admission: every pool must hold the request (products is the
  // upper bound of its output entries)
  if (max_row_products > c.max_row_products) {
    set_error("a row of %lld products exceeds the engine's row capacity", (long long)max_row_products);
    return SPWS_ERR_UNSUPPORTED;
  }
  if (c.max_cols > 0 && n > c.max_cols) {
    set_error("a request of %d columns exceeds the engine's column capacity %d", n, c.max_cols);
    return SPWS_ERR_UNSUPPORTED;
  }
  if (d.nreq + 1 > c.max_requests || (int64_t)d.rows + m > c.max_rows || (int64_t)d.brows + k > c.max_brows ||
      d.annz + annz > c.max_annz || d.bnnz + bnnz > c.max_bnnz || b->cur_ub + products > c.max_products ||
      (int64_t)d.cols + n > INT32_MAX) {
    set_error("the batch is full");
    return SPWS_ERR_OOM;
  }
  spws_batch_req& q = b->h_req[d.nreq];
  q.A_pos = A_pos;
  q.A_crd = A_crd;
  q.A_val = A_vals;
  q.B_pos = B_pos;
  q.B_crd = B_crd;
  q.B_val = B_vals;
  q.m = m;
  q.k = k;
  q.n = n;
  q.row0 = d.rows;
  q.brow0 = d.brows;
  q.col0 = d.cols;
  q.annz0 = d.annz;
  q.bnnz0 = d.bnnz;
  q.annz = annz;
  q.bnnz = bnnz;
  q.out_pos = nullptr;
  q.out_crd = nullptr;
  q.out_val = nullptr;
  q.out_cap = 0;
  if (slot) *slot = d.nreq;
  d.nreq += 1;
  d.rows += m;
  d.brows += k;
  d.cols += n;
  d.annz += annz;
  d.bnnz += bnnz;
  b->cur_ub += products;
  return SPWS_OK;
}

int spws_batch_set_output(spws_batch* b, int32_t slot, int32_t* C_pos, int32_t* C_crd, double* C_vals,
                          int64_t capacity) {
  SPWS_REQUIRE(b != nullptr, SPWS_ERR_INVALID, "null batch");
  SPWS_REQUIRE(!b->in_flight, SPWS_ERR_INVALID, "batch is still running");
  SPWS_REQUIRE(slot >= 0 && slot < b->cur.nreq, SPWS_ERR_INVALID, "slot %d out of range", slot);
  SPWS_REQUIRE(capacity >= 0, SPWS_ERR_INVALID, "negative output capacity");
  SPWS_REQUIRE(C_pos == nullptr || capacity == 0 || (C_crd && C_vals), SPWS_ERR_INVALID,
               "an output buffer with capacity needs columns and values");
  spws_batch_req& q = b->h_req[slot];
  q.out_pos = C_pos;
  q.out_crd = C_crd;
  q.out_val = C_vals;
  q.out_cap = C_pos ? capacity : 0;
  return SPWS_OK;
}

int spws_batch_launch(spws_batch* b) {
  SPWS_REQUIRE(b != nullptr, SPWS_ERR_INVALID, "null batch");
  SPWS_REQUIRE(!b->in_flight, SPWS_ERR_INVALID, "batch is already running");
  int rc = ctx_activate(b->ctx);
  if (rc) return rc;
  *b->h_desc = b->cur;
  *b->h_total = -1;
  if (b->exec) {
    SPWS_HIP(hipGraphLaunch(b->exec, b->stream));
  } else {
    enqueue_batch(b);
    SPWS_CHECK_LAUNCH();
  }
  SPWS_HIP(hipEventRecord(b->done, b->stream));
  b->in_flight = true;
  return SPWS_OK;
}

int spws_batch_poll(spws_batch* b, int* done) {
  SPWS_REQUIRE(b && done, SPWS_ERR_INVALID, "null argument");
  if (!b->in_flight) {
    *done = 1;
    return SPWS_OK;
  }
  hipError_t e = hipEventQuery(b->done);
  if (e == hipErrorNotReady) {
    *done = 0;
    return SPWS_OK;
  }
  if (e != hipSuccess) return fail_hip(e, "hipEventQuery (batch)", __FILE__, __LINE__);
  b->in_flight = false;
  *done = 1;
  return SPWS_OK;
}

int spws_batch_wait(spws_batch* b) {
  SPWS_REQUIRE(b != nullptr, SPWS_ERR_INVALID, "null batch");
  if (!b->in_flight) return SPWS_OK;
  SPWS_HIP(hipEventSynchronize(b->done));
  b->in_flight = false;
  return SPWS_OK;
}

static int batch_overflow_error() {
  set_error("an upper bound given at admission (products or max_row_products) was below the true one; "
            "the batch was not computed");
  return SPWS_ERR_INVALID;
}

int spws_batch_nnz(spws_batch* b, int64_t* nnz) {
  SPWS_REQUIRE(b && nnz, SPWS_ERR_INVALID, "null argument");
  SPWS_REQUIRE(!b->in_flight, SPWS_ERR_INVALID, "batch is still running");
  if (*b->h_total == -2) return batch_overflow_error();
  *nnz = *b->h_total;
  return SPWS_OK;
}

int spws_batch_result_nnz(spws_batch* b, int32_t slot, int64_t* nnz) {
  SPWS_REQUIRE(b && nnz, SPWS_ERR_INVALID, "null argument");
  SPWS_REQUIRE(!b->in_flight, SPWS_ERR_INVALID, "batch is still running");
  SPWS_REQUIRE(slot >= 0 && slot < b->cur.nreq, SPWS_ERR_INVALID, "slot %d out of range", slot);
  if (*b->h_total == -2) return batch_overflow_error();
  *nnz = b->h_req_nnz[slot];
  return SPWS_OK;
}

int spws_batch_result(spws_batch* b, int32_t slot, int64_t* nnz, int32_t* pos_host, int32_t** crd, double** vals) {
  SPWS_REQUIRE(b != nullptr, SPWS_ERR_INVALID, "null batch");
  SPWS_REQUIRE(!b->in_flight, SPWS_ERR_INVALID, "batch is still running");
  SPWS_REQUIRE(slot >= 0 && slot < b->cur.nreq, SPWS_ERR_INVALID, "slot %d out of range", slot);
  if (*b->h_total == -2) return batch_overflow_error();
  int rc = ctx_activate(b->ctx);
  if (rc) return rc;
  const spws_batch_req& q = b->h_req[slot];
  const int32_t* Cp = b->pool[BP_CPOS].as<int32_t>();
  std::vector<int32_t> tmp((size_t)q.m + 1);
  SPWS_HIP(hipMemcpyAsync(tmp.data(), Cp + q.row0, ((size_t)q.m + 1) * 4, hipMemcpyDeviceToHost, b->stream));
  SPWS_HIP(hipStreamSynchronize(b->stream));
  const int32_t base = tmp[0];
  if (pos_host)
    for (int32_t i = 0; i <= q.m; ++i) pos_host[i] = tmp[i] - base;
  if (nnz) *nnz = tmp[q.m] - base;
  // a request with an output buffer that holds it was written there
  const bool delivered = q.out_pos && b->h_req_nnz[slot] >= 0;
  if (crd) *crd = delivered ? q.out_crd : b->pool[BP_CCRD].as<int32_t>() + base;
  if (vals) *vals = delivered ? q.out_val : b->pool[BP_CVAL].as<double>() + base;
  return SPWS_OK;
}

void* spws_batch_stream(spws_batch* b) { return b ? (void*)b->stream : nullptr; }

}  // extern "C"
