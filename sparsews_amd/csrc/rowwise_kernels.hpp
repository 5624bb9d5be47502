// Row-wise SpGEMM device code shared by the single-product driver
// (spgemm_rowwise.hip) and the batched engine (batch.hip).  See
// spgemm_rowwise.hip for the pipeline and the reference it follows.
#pragma once
#include <climits>

#include "common.hpp"

namespace spws {
namespace {

enum RowPath : int { P_WAVE = 0, P_WAVE2 = 1, P_BLOCK = 2, P_BLOCK2 = 3, P_DENSE = 4, P_WIDE = 5, P_EMPTY = 6 };
constexpr int kPaths = 7;

// compiled workspace capacities (max ub per path; tables hold 2x slots)
constexpr int kWaveUbMax = 64;
constexpr int kWave2UbMax = 256;
constexpr int kBlockUbMax = 1024;
constexpr int kBlock2UbMax = 4096;
constexpr int kWideLdsSort = 4096;  // wide rows with nnz <= this sort in LDS

struct PathParams {
  int32_t wave_ub, wave2_ub, block_ub, block2_ub, dense_ratio, force;
};

__host__ __device__ __forceinline__ int classify_row(int64_t ub, int32_t ncols, const PathParams& p) {
  if (ub == 0) return P_EMPTY;
  if (p.force >= 0) return p.force;
  if (ub <= p.wave_ub) return P_WAVE;
  if (ub <= p.wave2_ub) return P_WAVE2;
  if (ub <= p.block_ub) return P_BLOCK;
  if (ub <= p.block2_ub) return P_BLOCK2;
  if (ub * (int64_t)p.dense_ratio >= (int64_t)ncols) return P_DENSE;
  return P_WIDE;
}

// Bin metadata on the device (int32 bmeta): [0, 8) rows per path, [8, 16)
// offset of each path's rows in bin_rows, [16, 24) scatter cursors; maxub
// (uint64): [0, 8) largest row upper bound per path, [8] sum of all upper
// bounds.
constexpr int kMetaCount = 0, kMetaOff = 8, kMetaCursor = 16, kMetaInts = 24;

__device__ __forceinline__ int32_t dev_or(const int32_t* p, int32_t v) { return p ? *p : v; }

// Tiles of kRsRows consecutive rows, one workgroup of kRsBlock threads per
// tile (grid-stride).  The tile's A entries are walked as one flat range,
// kRsUnroll strides of kRsBlock in flight per thread, so rows of any length
// cost a few memory latencies: each entry's len(B_j) is added to its row's
// LDS accumulator (the row found by a binary search over the tile's row
// pointers in LDS; a wavefront whose entries all lie in one row adds once).
// Each row's path is counted in an LDS histogram.
constexpr int kRsBlock = 256;
constexpr int kRsRows = 64;
constexpr int kRsRowsLog2 = 6;
constexpr int kRsUnroll = 8;
static_assert(kRsRows == 1 << kRsRowsLog2 && kRsRows <= kRsBlock, "tile rows");

__global__ __launch_bounds__(kRsBlock) void k_rowstat(int32_t m_host, const int32_t* __restrict__ m_dev,
                                                      int32_t ncols, const int32_t* __restrict__ Ap,
                                                      const int32_t* __restrict__ Ai, const int32_t* __restrict__ Bp,
                                                      int64_t* __restrict__ row_ub, int32_t* __restrict__ bmeta,
                                                      unsigned long long* __restrict__ maxub, PathParams pp) {
  __shared__ int cnt[8];
  __shared__ unsigned long long mx[8];
  __shared__ int32_t sAp[kRsRows + 1];
  __shared__ unsigned long long sub[kRsRows];
  const int t = threadIdx.x;
  if (t < 8) {
    cnt[t] = 0;
    mx[t] = 0;
  }
  const int32_t m = dev_or(m_dev, m_host);
  unsigned long long tot = 0;
  for (int32_t r0 = blockIdx.x * kRsRows; r0 < m; r0 += gridDim.x * kRsRows) {
    const int32_t nr = min(kRsRows, m - r0);
    __syncthreads();  // the previous tile's readers are done
    if (t <= nr) sAp[t] = Ap[r0 + t];  // nr + 1 row pointers
    if (t < kRsRows) sub[t] = 0;
    __syncthreads();
    const int32_t a0 = sAp[0], a1 = sAp[nr];
    // int64 positions: a0 + kRsBlock * kRsUnroll may pass INT32_MAX
    for (int64_t p0 = a0; p0 < a1; p0 += kRsBlock * kRsUnroll) {
      int32_t j[kRsUnroll], len[kRsUnroll];
#pragma unroll
      for (int u = 0; u < kRsUnroll; ++u) {
        const int64_t p = p0 + u * kRsBlock + t;
        j[u] = p < a1 ? Ai[p] : -1;
      }
#pragma unroll
      for (int u = 0; u < kRsUnroll; ++u) len[u] = j[u] >= 0 ? Bp[j[u] + 1] - Bp[j[u]] : 0;
#pragma unroll
      for (int u = 0; u < kRsUnroll; ++u) {
        if (p0 + u * kRsBlock >= a1) break;  // workgroup-uniform: no entries left
        // the tile row holding entry p: the last q with sAp[q] <= p (p past
        // the range maps to the last row and adds 0)
        const int32_t p = (int32_t)min(p0 + u * kRsBlock + t, (int64_t)a1 - 1);
        int lo = 0, hi = nr - 1;
#pragma unroll
        for (int it = 0; it < kRsRowsLog2; ++it) {
          const int mid = (lo + hi + 1) >> 1;
          if (sAp[mid] <= p) lo = mid; else hi = mid - 1;
        }
        if (__all(lo == __shfl(lo, 0, kWave))) {
          const int32_t wsum = wave_sum(len[u]);
          if (lane_id() == 0 && wsum) atomicAdd(&sub[lo], (unsigned long long)wsum);
        } else if (len[u]) {
          atomicAdd(&sub[lo], (unsigned long long)len[u]);
        }
      }
    }
    __syncthreads();
    if (t < nr) {
      const unsigned long long ub = sub[t];
      row_ub[r0 + t] = (int64_t)ub;
      const int path = classify_row((int64_t)ub, ncols, pp);
      atomicAdd(&cnt[path], 1);
      atomicMax(&mx[path], ub);
      tot += ub;
    }
  }
  tot = wave_sum(tot);
  if (lane_id() == 0 && tot) atomicAdd(&maxub[8], tot);
  __syncthreads();
  if (t < kPaths && cnt[t]) {
    atomicAdd(&bmeta[kMetaCount + t], cnt[t]);
    atomicMax(&maxub[t], mx[t]);
  }
}

// offsets of the paths' row lists (exclusive prefix of the counts)
__global__ void k_bin_offsets(int32_t* __restrict__ bmeta) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    int32_t run = 0;
    for (int p = 0; p < 8; ++p) {
      bmeta[kMetaOff + p] = run;
      if (p < P_EMPTY) run += bmeta[kMetaCount + p];
    }
  }
}

// ints of the metadata block: bmeta [0, kMetaInts), then 9 uint64 maxub
// from int 32 on
constexpr int kMetaBlockInts = 32 + 2 * 9;

// zero the bin metadata (a kernel, so the sequence can be graph-captured)
__global__ void k_meta_clear(int32_t* __restrict__ bmeta, unsigned long long* __restrict__ maxub) {
  const int t = threadIdx.x;
  if (t < kMetaInts) bmeta[t] = 0;
  if (t < 9) maxub[t] = 0;  // [8]: sum of the upper bounds
}

// Row ids grouped by path, a tile of kRsBlock rows per workgroup: ranks
// within each wavefront from ballots, wavefront offsets within the tile in
// LDS, and one global cursor atomic per path and tile (issued by separate
// threads, so a tile waits for one atomic latency).
__global__ __launch_bounds__(kRsBlock) void k_binscatter(int32_t m_host, const int32_t* __restrict__ m_dev,
                                                         int32_t ncols, const int64_t* __restrict__ row_ub,
                                                         int32_t* __restrict__ bmeta, int32_t* __restrict__ bin_rows,
                                                         int32_t* __restrict__ row_nnz, PathParams pp) {
  constexpr int NW = kRsBlock / kWave;
  __shared__ int32_t woff[NW][8];
  __shared__ int32_t gbase[8];
  const int32_t m = dev_or(m_dev, m_host);
  const int t = threadIdx.x, lane = lane_id(), w = t / kWave;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int32_t base = blockIdx.x * kRsBlock; base < m; base += gridDim.x * kRsBlock) {
    const int32_t i = base + t;
    const int path = i < m ? classify_row(row_ub[i], ncols, pp) : -1;
    if (i < m && path == P_EMPTY) row_nnz[i] = 0;
    int rank = 0;
#pragma unroll
    for (int p = 0; p < P_EMPTY; ++p) {
      const uint64_t mask = __ballot(path == p);
      if (path == p) rank = __popcll(mask & lt);
      if (lane == 0) woff[w][p] = __popcll(mask);
    }
    __syncthreads();
    if (t < P_EMPTY) {
      int32_t run = 0;
      for (int v = 0; v < NW; ++v) {
        const int32_t c = woff[v][t];
        woff[v][t] = run;
        run += c;
      }
      gbase[t] = run ? bmeta[kMetaOff + t] + atomicAdd(&bmeta[kMetaCursor + t], run) : 0;
    }
    __syncthreads();
    if (path >= 0 && path < P_EMPTY) bin_rows[gbase[path] + woff[w][path] + rank] = i;
    __syncthreads();  // woff and gbase are reused by the next tile
  }
}

// ------------------------------------------------------------------ product expansion
//
// The products of row i are enumerated 64 A-entries at a time: each lane
// loads one A entry (j, a) and len(B_j); an inclusive wave scan gives the
// product offsets, and each lane then finds, for its product index q, the
// owning A entry by a binary search over the lanes' offsets (six
// ds_bpermute rounds).  Rows with fewer than two chunks of A entries per
// wavefront: every wavefront scans the same chunk and takes its own strides
// of the chunk's products, so rows with few A entries but long B rows still
// use all wavefronts.  Longer rows: each wavefront takes whole chunks.  Loop
// bounds are wave-uniform, so the shuffles see every lane.

template <bool kValues, class F>
__device__ __forceinline__ void expand_products(int32_t a_beg, int32_t a_end, int32_t widx, int32_t nwaves,
                                                const int32_t* __restrict__ Ai, const double* __restrict__ Ax,
                                                const int32_t* __restrict__ Bp, const int32_t* __restrict__ Bi,
                                                const double* __restrict__ Bx, F&& insert) {
  // kU strides of products are located and loaded before any is inserted,
  // so their B loads are in flight together
  constexpr int kU = 4;
  const int lane = lane_id();
  // rows of at least two chunks per wavefront: each wavefront takes whole
  // chunks of its own (chunk widx, widx + nwaves, ...), so the chunks' load
  // chains run side by side instead of one after another
  const bool own_chunks = a_end - a_beg >= 2 * kWave * nwaves;
  const int32_t c_first = own_chunks ? a_beg + widx * kWave : a_beg;
  const int32_t c_step = own_chunks ? nwaves * kWave : kWave;
  if (own_chunks) {
    widx = 0;
    nwaves = 1;
  }
  for (int32_t c = c_first; c < a_end; c += c_step) {
    const int32_t p = c + lane;
    int32_t bstart = 0, blen = 0;
    double a = 0.0;
    if (p < a_end) {
      const int32_t j = Ai[p];
      bstart = Bp[j];
      blen = Bp[j + 1] - bstart;
      if (kValues) a = Ax[p];
    }
    const int32_t incl = wave_incl_scan(blen);
    const int32_t excl = incl - blen;
    const int32_t total = __shfl(incl, kWave - 1, kWave);
    for (int32_t q0 = widx * kWave * kU; q0 < total; q0 += nwaves * kWave * kU) {
      // strides past the row's end are skipped (wave-uniform)
      const int nu = min(kU, (total - q0 + kWave - 1) / kWave);
      int32_t kb[kU];
      double oa[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (u >= nu) break;
        // product index clamped into the row (lanes past the end are masked at insert)
        const int32_t q = min(q0 + u * kWave + lane, total - 1);
        // owner: the last lane t with excl_t <= q
        int lo = 0, hi = kWave - 1;
#pragma unroll
        for (int it = 0; it < 6; ++it) {
          const int mid = (lo + hi + 1) >> 1;
          const int32_t em = __shfl(excl, mid, kWave);
          if (em <= q) lo = mid; else hi = mid - 1;
        }
        kb[u] = __shfl(bstart, lo, kWave) + (q - __shfl(excl, lo, kWave));
        oa[u] = kValues ? __shfl(a, lo, kWave) : 0.0;
      }
      int32_t col[kU];
      double bv[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (u >= nu) break;
        col[u] = Bi[kb[u]];
        bv[u] = kValues ? Bx[kb[u]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (u >= nu) break;
        if (q0 + u * kWave + lane < total) insert(col[u], kValues ? oa[u] * bv[u] : 0.0);
      }
    }
  }
}

// The products of K rows gathered into registers: product g of row k (in
// the order of expand_products) belongs to thread g % BLOCK as item
// g / BLOCK.  A row's A entries are taken 64 at a time (kMultiChunk false:
// the caller guarantees at most 64); with K > 1 each level of the first
// chunk (A entries, B row bounds, B entries) is loaded for all K rows
// together, so their load chains overlap.  Rows with take[k] false are
// skipped (workgroup-uniform), and the caller guarantees that a taken row
// has at most BLOCK * ITEMS products.  Unused items hold col = -1.  The
// chunk loop costs registers: the block-bin kernel, whose occupancy
// depends on them, takes single-chunk rows only.
template <int BLOCK, int K, int ITEMS, bool kMultiChunk>
__device__ __forceinline__ void gather_rows_products(const bool (&take)[K], const int32_t (&a_beg)[K],
                                                     const int32_t (&a_end)[K], const int32_t* __restrict__ Ai,
                                                     const double* __restrict__ Ax, const int32_t* __restrict__ Bp,
                                                     const int32_t* __restrict__ Bi, const double* __restrict__ Bx,
                                                     int32_t (&col)[K][ITEMS], double (&val)[K][ITEMS]) {
  const int lane = lane_id();
  const int32_t wbase = (int32_t)threadIdx.x - lane;
  int32_t j[K], bstart[K], blen[K];
  double a[K];
  // first chunk of every row: A entries, then B row bounds
#pragma unroll
  for (int k = 0; k < K; ++k) {
    j[k] = -1;
    a[k] = 0.0;
    if (take[k] && a_beg[k] + lane < a_end[k]) {
      j[k] = Ai[a_beg[k] + lane];
      a[k] = Ax[a_beg[k] + lane];
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    bstart[k] = 0;
    blen[k] = 0;
    if (j[k] >= 0) {
      bstart[k] = Bp[j[k]];
      blen[k] = Bp[j[k] + 1] - bstart[k];
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int u = 0; u < ITEMS; ++u) {
      col[k][u] = -1;
      val[k][u] = 0.0;
    }
    if (!take[k]) continue;
    int32_t base = 0;  // products of the row's earlier chunks
    const int32_t c_end = kMultiChunk ? a_end[k] : min(a_end[k], a_beg[k] + kWave);
    for (int32_t c = a_beg[k]; c < c_end; c += kWave) {
      if (c != a_beg[k]) {  // later chunks: their A entries and B row bounds
        const int32_t p = c + lane;
        j[k] = -1;
        a[k] = 0.0;
        bstart[k] = blen[k] = 0;
        if (p < a_end[k]) {
          j[k] = Ai[p];
          a[k] = Ax[p];
          bstart[k] = Bp[j[k]];
          blen[k] = Bp[j[k] + 1] - bstart[k];
        }
      }
      const int32_t incl = wave_incl_scan(blen[k]);
      const int32_t excl = incl - blen[k];
      const int32_t total = __shfl(incl, kWave - 1, kWave);
      int32_t kb[ITEMS];
      double oa[ITEMS];
#pragma unroll
      for (int u = 0; u < ITEMS; ++u) {
        kb[u] = -1;
        oa[u] = 0.0;
        // this wave's item-u products are [g0, g0 + 64); skipped
        // wave-uniformly when none of them lies in this chunk
        const int32_t g0 = u * BLOCK + wbase;
        if (g0 + kWave <= base || g0 >= base + total) continue;
        const int32_t qr = g0 + lane - base;
        const int32_t q = min(max(qr, 0), total - 1);
        int lo = 0, hi = kWave - 1;
#pragma unroll
        for (int it = 0; it < 6; ++it) {
          const int mid = (lo + hi + 1) >> 1;
          const int32_t em = __shfl(excl, mid, kWave);
          if (em <= q) lo = mid; else hi = mid - 1;
        }
        const int32_t kk = __shfl(bstart[k], lo, kWave) + (q - __shfl(excl, lo, kWave));
        const double ao = __shfl(a[k], lo, kWave);
        if (qr >= 0 && qr < total) {
          kb[u] = kk;
          oa[u] = ao;
        }
      }
#pragma unroll
      for (int u = 0; u < ITEMS; ++u) {
        if (kb[u] >= 0) {
          col[k][u] = Bi[kb[u]];
          val[k][u] = oa[u] * Bx[kb[u]];
        }
      }
      base += total;
    }
  }
}

// ------------------------------------------------------------------ LDS hash

template <bool kValues>
__device__ __forceinline__ void lds_insert(int32_t* keys, double* vals, uint32_t shift, uint32_t mask,
                                           int32_t col, double v) {
  uint32_t h = hash_col((uint32_t)col, shift);
  for (;;) {
    int32_t cur = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == -1) {
      cur = atomicCAS(&keys[h], -1, col);
      if (cur == -1) cur = col;
    }
    if (cur == col) {
      if (kValues) atomicAdd(&vals[h], v);
      return;
    }
    h = (h + 1) & mask;
  }
}

__device__ __forceinline__ double lds_lookup(const int32_t* keys, const double* vals, uint32_t shift, uint32_t mask,
                                             int32_t col) {
  uint32_t h = hash_col((uint32_t)col, shift);
  while (keys[h] != col) h = (h + 1) & mask;
  return vals[h];
}

// ------------------------------------------------------------------ bitonic sorts
//
// "Flip" bitonic network: every comparator is ascending; stage k first
// compares e with e ^ (k - 1), then e with e ^ j for j = k/4 .. 1.  Padding
// with INT_MAX (or skipping partners beyond n) therefore sorts any n.

// LDS bitonic of P (power of two) keys, all threads of the block participate.
template <int BLOCK>
__device__ __forceinline__ void block_bitonic_lds(int32_t* list, int P) {
  for (int k = 2; k <= P; k <<= 1) {
    for (int t = threadIdx.x; t < P / 2; t += BLOCK) {
      const int half = k >> 1;
      const int b = t / half, o = t % half;
      const int lo = b * k + o, hi = b * k + k - 1 - o;
      const int32_t x = list[lo], y = list[hi];
      if (x > y) {
        list[lo] = y;
        list[hi] = x;
      }
    }
    __syncthreads();
    for (int j = k >> 2; j >= 1; j >>= 1) {
      for (int t = threadIdx.x; t < P / 2; t += BLOCK) {
        const int b = t / j, o = t % j;
        const int lo = b * 2 * j + o, hi = lo + j;
        const int32_t x = list[lo], y = list[hi];
        if (x > y) {
          list[lo] = y;
          list[hi] = x;
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ wave-per-row kernel
//
// One 64-thread workgroup (one wavefront), K rows per step of a grid-stride
// loop, TS = 2 * max ub slots per row.  Every row's products (at most 64 * R
// by its bin) are gathered into registers, the first chunk of A entries of
// the K rows one load level at a time, so their load chains overlap.

// rows per step of the first wave bin: two rows per step overlap their
// load chains but measure slower than one (the per-row LDS work dominates)
constexpr int kWaveRowsK = 1;

// bin_rows == nullptr: rows 0 .. direct_m - 1 in order, each with a fixed
// region of kWave slots at r * kWave (every row has ub <= 64).
template <int TS, int R, bool kValues, int K = 1>
__global__ __launch_bounds__(64) void k_row_wave(const int32_t* __restrict__ bin_rows, const int32_t* __restrict__ bmeta,
                                                 int path,
                                                 const int32_t* __restrict__ Ap, const int32_t* __restrict__ Ai,
                                                 const double* __restrict__ Ax, const int32_t* __restrict__ Bp,
                                                 const int32_t* __restrict__ Bi, const double* __restrict__ Bx,
                                                 const int64_t* __restrict__ row_off, int32_t* __restrict__ slot_crd,
                                                 double* __restrict__ slot_val, int32_t* __restrict__ row_nnz,
                                                 int32_t direct_m = 0, const int64_t* __restrict__ direct_ub = nullptr) {
  static_assert(TS >= 2 * kWave * R / 2, "table too small");
  __shared__ int32_t keys[K][TS];
  __shared__ double vals[K][kValues ? TS : 1];
  __shared__ int32_t list[K][kWave * R];
  constexpr uint32_t kMask = TS - 1;
  const uint32_t shift = 32 - log2_pow2(TS);
  const int lane = lane_id();
  const bool direct = bin_rows == nullptr;
  const int32_t count = direct ? direct_m : bmeta[kMetaCount + path];
  const int32_t* rows = direct ? nullptr : bin_rows + bmeta[kMetaOff + path];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int32_t r0 = blockIdx.x * K; r0 < count; r0 += gridDim.x * K) {
    int32_t i[K], a_beg[K], a_end[K];
    bool take[K];
#pragma unroll
    for (int k = 0; k < K; ++k) i[k] = r0 + k < count ? (direct ? r0 + k : rows[r0 + k]) : -1;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      a_beg[k] = a_end[k] = 0;
      if (i[k] >= 0) {
        a_beg[k] = Ap[i[k]];
        a_end[k] = Ap[i[k] + 1];
      }
      // every row of the bins has ub <= 64 * R; over all rows (direct_ub
      // given), a row beyond that is skipped with a count of 0
      take[k] = i[k] >= 0 && (direct_ub == nullptr || direct_ub[i[k]] <= (int64_t)kWave * R);
#pragma unroll
      for (int s = lane; s < TS; s += kWave) {
        keys[k][s] = -1;
        if (kValues) vals[k][s] = 0.0;
      }
    }
    int32_t pc[K][R];
    double pv[K][R];
    gather_rows_products<kWave, K, R, true>(take, a_beg, a_end, Ai, Ax, Bp, Bi, Bx, pc, pv);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
      for (int u = 0; u < R; ++u)
        if (pc[k][u] >= 0) lds_insert<kValues>(keys[k], vals[k], shift, kMask, pc[k][u], pv[k][u]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (i[k] < 0) continue;
      // compaction of the occupied slots (order irrelevant: sorted next)
      int32_t nnz = 0;
#pragma unroll
      for (int s0 = 0; s0 < TS; s0 += kWave) {
        const int32_t key = keys[k][s0 + lane];
        const uint64_t occ = __ballot(key != -1);
        if (kValues && key != -1) list[k][nnz + __popcll(occ & lt)] = key;
        nnz += __popcll(occ);
      }
      if (!kValues) {
        if (lane == 0) row_nnz[i[k]] = nnz;
        continue;
      }
      __syncthreads();
      int32_t v[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int e = q * kWave + lane;
        v[q] = e < nnz ? list[k][e] : INT_MAX;
      }
      int n_pow2 = 2;
      while (n_pow2 < nnz) n_pow2 <<= 1;
      wave_bitonic<R>(v, n_pow2);
      const int64_t off = direct ? (int64_t)i[k] * kWave : row_off[i[k]];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int e = q * kWave + lane;
        if (e < nnz) {
          slot_crd[off + e] = v[q];
          slot_val[off + e] = lds_lookup(keys[k], vals[k], shift, kMask, v[q]);
        }
      }
      if (lane == 0) row_nnz[i[k]] = nnz;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ short-row compaction
//
// When every row has ub <= 64, k_row_wave runs over the rows directly (no
// bins) into fixed 64-entry slot regions, and this kernel turns the slots
// into C: each workgroup owns kCfRows consecutive rows, adds up the counts of
// all rows before them (at most kCfMaxRows rows, so a few hundred loads per
// thread), scans its own rows' counts and copies the rows, writing C_pos.
// Two launches, no upper-bound scan, no bins and no count scan.

constexpr int kCfRows = 64;  // a multiple of 4: the earlier rows' counts are read as int4
constexpr int kCfBlock = 256;
constexpr int kCfPart = 8;  // partial sums (16-byte loads in flight per thread) of the row-count prefix
constexpr int32_t kCfMaxRows = 1 << 16;

// ------------------------------------------------------------------ short-row product in two launches
//
// When every row has at most 64 products (and at most 64 A entries):
//   k_short_rows     wavefront w owns the q consecutive rows from w q
//                    (q <= kShortRows).  It gathers the products of all its
//                    rows together -- row pointers, A entries, B row
//                    pointers, B entries: four dependent loads in all, not
//                    four per row -- and stages them in LDS (column, A value
//                    * B value); then all its rows at once, row r in
//                    register r across the lanes: a bitonic sort by column
//                    carrying the values, and a scan that sums runs of equal
//                    columns.  The wavefront's rows go out packed into its
//                    slot region (q * 64 entries from w q 64), with the row
//                    counts and the wavefront's totals (entries, products).
//                    Each workgroup also writes its own totals.
//   k_short_compact  a wavefront per wavefront of the first kernel: the
//                    totals of all earlier workgroups (each workgroup sums
//                    them itself: at most kShortMaxWaves / 4 values), the
//                    prefix over its own wavefronts, then C_pos and a
//                    coalesced copy of the wavefront's packed entries into
//                    C, every load issued before the first store.
// No upper-bound pass, no bins, no host round trip in between.  A single
// launch with a look-back across wavefronts was measured slower than both
// kernels together: status words polled across the XCDs cost a coherent
// round trip each, and a same-address ticket counter serialises.
// A wavefront with a row that does not fit reports a total of -1; the
// compaction then writes nothing past the totals, out[1] reads 1, and the
// caller runs the general pipeline (which recomputes C).  The last
// compaction workgroup writes all of out with plain stores (nothing is
// cleared beforehand):
//   out[0] products, out[1] 1 if a row did not fit (else 0), out[2] nnz(C).
constexpr int kShortBlock = 256;
constexpr int kShortWaves = kShortBlock / kWave;
constexpr int kShortRows = 8;           // rows per wavefront at most
constexpr int kShortAEntries = 128;     // A entries per wavefront at most
constexpr int kShortProducts = 512;     // staged products per wavefront at most
constexpr int kShortMaxWaves = 1 << 14; // wavefronts of the first kernel at most
constexpr int kShortCompactWaves = kShortWaves;  // first-kernel wavefronts per compaction workgroup (one per wavefront)

// Sorts R independent 64-entry sequences (register r of every lane is
// sequence r, entry = lane) by key, carrying the values along; the first
// n_pow2 entries of each take part (the rest must be INT_MAX already).
// Equal keys keep their own values: a lane takes its partner's entry only
// on a strict comparison.
template <int R>
__device__ __forceinline__ void wave_sort_pairs(int32_t (&k)[R], double (&v)[R], int n_pow2) {
  const int lane = lane_id();
  auto step = [&](int d, bool low) {
    int32_t pk[R];
    double pv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      pk[r] = __shfl_xor(k[r], d, kWave);
      pv[r] = __shfl_xor(v[r], d, kWave);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool take = low ? pk[r] < k[r] : pk[r] > k[r];
      if (take) {
        k[r] = pk[r];
        v[r] = pv[r];
      }
    }
  };
  for (int kk = 2; kk <= n_pow2; kk <<= 1) {
    step(kk - 1, (lane & (kk >> 1)) == 0);
    for (int j = kk >> 2; j >= 1; j >>= 1) step(j, (lane & j) == 0);
  }
}

// kPacked: columns below 2^26 - 1, so that (column << 6 | position in the
// row) fits 32 bits below the padding key; the sort then moves one dword
// per exchange and the values are read from LDS at the sorted positions.
template <int R, bool kPacked>
[[maybe_unused]] __global__ __launch_bounds__(kShortBlock) void k_short_rows(
    int32_t m, int32_t q, int32_t nwaves, const int32_t* __restrict__ Ap, const int32_t* __restrict__ Ai,
    const double* __restrict__ Ax, const int32_t* __restrict__ Bp, const int32_t* __restrict__ Bi,
    const double* __restrict__ Bx, int32_t* __restrict__ slot_crd, double* __restrict__ slot_val,
    int32_t* __restrict__ row_nnz, int32_t* __restrict__ wave_tot, int32_t* __restrict__ wg_tot,
    int32_t* __restrict__ wg_prod) {
  __shared__ int32_t s_col[kShortWaves][kShortProducts];
  __shared__ double s_pv[kShortWaves][kShortProducts];
  __shared__ double s_av[kShortWaves][kShortAEntries];
  __shared__ uint8_t s_ai[kShortWaves][kShortProducts];
  static_assert(kShortAEntries <= 256, "A entry indices are stored in a byte");
  static_assert(kShortAEntries == 2 * kWave, "two A entries per lane");
  const int wv = threadIdx.x / kWave;
  int32_t* scol = s_col[wv];
  double* spv = s_pv[wv];
  double* sav = s_av[wv];
  uint8_t* sai = s_ai[wv];
  const int lane = lane_id();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int32_t w = blockIdx.x * kShortWaves + wv;
  __shared__ int32_t s_wt[kShortWaves], s_wp[kShortWaves];
  // the wavefront's totals: entries (-1 when a row does not fit), products
  int32_t my_tot = 0, my_prod = 0;
  auto wave_rows = [&]() {
  const int32_t r0 = (int32_t)min((int64_t)w * q, (int64_t)m);
  const int32_t nr = min(q, m - r0);  // <= q <= R

  // row pointers, A entries and B row pointers of all the rows
  const int32_t rp = (lane <= nr && nr > 0) ? Ap[r0 + lane] : 0;
  const int32_t a0 = __shfl(rp, 0, kWave), a1 = __shfl(rp, nr, kWave);
  const int32_t nA = a1 - a0;
  bool over = nA > kShortAEntries;
  int32_t bs0 = 0, bl0 = 0, bs1 = 0, bl1 = 0;
  double av0 = 0.0, av1 = 0.0;
  if (!over) {
    if (lane < nA) {
      const int32_t j = Ai[a0 + lane];
      av0 = Ax[a0 + lane];
      bs0 = Bp[j];
      bl0 = Bp[j + 1] - bs0;
    }
    if (kWave + lane < nA) {
      const int32_t j = Ai[a0 + kWave + lane];
      av1 = Ax[a0 + kWave + lane];
      bs1 = Bp[j];
      bl1 = Bp[j + 1] - bs1;
    }
  }
  // exclusive product offsets over the wavefront's A entries (two halves)
  const int32_t inc0 = wave_incl_scan(bl0);
  const int32_t tot0 = __shfl(inc0, kWave - 1, kWave);
  const int32_t inc1 = wave_incl_scan(bl1) + tot0;
  const int32_t P = __shfl(inc1, kWave - 1, kWave);
  const int32_t ex0 = inc0 - bl0, ex1 = inc1 - bl1;
  // lane r < nr: its row's A entries [lo, hi) and products [pb, pe)
  auto E = [&](int32_t x) -> int32_t {  // exclusive offset of A entry x (P at x = nA)
    const int32_t e0 = __shfl(ex0, x & (kWave - 1), kWave);
    const int32_t e1 = __shfl(ex1, x & (kWave - 1), kWave);
    return x >= nA ? P : (x < kWave ? e0 : e1);
  };
  const int32_t lo = rp - a0, hi = __shfl(rp, min(lane + 1, kWave - 1), kWave) - a0;
  const int32_t pb = E(lo), pe = E(hi);
  if (lane < nr && (hi - lo > kWave || pe - pb > kWave)) over = true;
  over = __any(over || P > kShortProducts);
  if (over) {  // the caller discards the result
    my_tot = -1;
    if (lane < nr) row_nnz[r0 + lane] = 0;
    return;
  }

  // every product staged in LDS at its A entry's offset, in two steps so
  // that the B loads are coalesced: each A entry first writes, for each of
  // its products, the product's B index (over scol) and its own index;
  // then lane l loads products l, l + 64, ... (consecutive products of a B
  // row sit in consecutive lanes), all loads of the wavefront in flight
  {
    sav[lane] = av0;
    sav[kWave + lane] = av1;
    const int32_t maxlen = wave_max(max(bl0, bl1));
    for (int32_t t = 0; t < maxlen; ++t) {
      if (t < bl0) {
        scol[ex0 + t] = bs0 + t;
        sai[ex0 + t] = (uint8_t)lane;
      }
      if (t < bl1) {
        scol[ex1 + t] = bs1 + t;
        sai[ex1 + t] = (uint8_t)(kWave + lane);
      }
    }
    __builtin_amdgcn_wave_barrier();
    constexpr int kU = kShortProducts / kWave;
    int32_t c[kU];
    double x[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int32_t p = u * kWave + lane;
      if (u * kWave < P) {  // uniform
        const int32_t idx = p < P ? scol[p] : 0;
        c[u] = p < P ? Bi[idx] : 0;
        x[u] = p < P ? Bx[idx] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int32_t p = u * kWave + lane;
      if (u * kWave < P && p < P) {
        scol[p] = c[u];
        spv[p] = sav[sai[p]] * x[u];
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  // all rows at once, row r in register r (its products across the lanes,
  // INT_MAX past the end): sort by column, then sum runs of equal columns
  const int32_t rlen = lane < nr ? pe - pb : 0;
  const int32_t maxrow = wave_max(rlen);
  int n_pow2 = 2;
  while (n_pow2 < maxrow) n_pow2 <<= 1;
  int32_t kc[R];
  double kv[R];
  if constexpr (kPacked) {
    uint32_t key[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t rb = __builtin_amdgcn_readlane(pb, r);
      const int32_t len = r < nr ? __builtin_amdgcn_readlane(rlen, r) : 0;
      key[r] = lane < len ? ((uint32_t)scol[rb + lane] << 6) | (uint32_t)lane : 0xFFFFFFFFu;
    }
    wave_sort64<R>(key, n_pow2);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t rb = __builtin_amdgcn_readlane(pb, r);
      const bool valid = key[r] != 0xFFFFFFFFu;
      kc[r] = valid ? (int32_t)(key[r] >> 6) : INT_MAX;
      kv[r] = valid ? spv[rb + (int32_t)(key[r] & 63u)] : 0.0;
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t rb = __builtin_amdgcn_readlane(pb, r);
      const int32_t len = r < nr ? __builtin_amdgcn_readlane(rlen, r) : 0;
      kc[r] = lane < len ? scol[rb + lane] : INT_MAX;
      kv[r] = lane < len ? spv[rb + lane] : 0.0;
    }
    wave_sort_pairs<R>(kc, kv, n_pow2);
  }
  // runs of equal columns are contiguous: an inclusive scan restricted to
  // equal keys leaves each run's sum at its last lane (only rows that have
  // a repeated column take it)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int32_t prev = __shfl_up(kc[r], 1, kWave);
    if (__ballot(lane > 0 && kc[r] != INT_MAX && prev == kc[r])) {
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const int32_t pk = __shfl_up(kc[r], d, kWave);
        const double ps = __shfl_up(kv[r], d, kWave);
        if (lane >= d && pk == kc[r]) kv[r] += ps;
      }
    }
  }
  // the rows packed into the wavefront's slot region: each run's last lane
  // writes (column, sum) at its rank
  const int64_t base = (int64_t)w * q * kWave;
  int32_t loc = 0, my_nnz = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int32_t nxt = __shfl_down(kc[r], 1, kWave);
    const bool last = kc[r] != INT_MAX && (lane == kWave - 1 || nxt != kc[r]);
    const uint64_t lm = __ballot(last);
    if (last) {
      const int64_t dst = base + loc + __popcll(lm & lt);
      slot_crd[dst] = kc[r];
      slot_val[dst] = kv[r];
    }
    const int32_t cnt = __popcll(lm);
    if (lane == r) my_nnz = cnt;
    loc += cnt;
  }
  if (lane < nr) row_nnz[r0 + lane] = my_nnz;
  my_tot = loc;
  my_prod = P;
  };
  // the last workgroup's spare wavefronts own no rows and add nothing
  if (w < nwaves) {
    wave_rows();
    if (lane == 0) wave_tot[w] = my_tot;
  }
  if (lane == 0) {
    s_wt[wv] = my_tot;
    s_wp[wv] = my_prod;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t t = 0, pr = 0;
    bool bad = false;
#pragma unroll
    for (int i = 0; i < kShortWaves; ++i) {
      bad |= s_wt[i] < 0;
      t += s_wt[i];
      pr += s_wp[i];
    }
    wg_tot[blockIdx.x] = bad ? -1 : t;
    wg_prod[blockIdx.x] = pr;
  }
}

[[maybe_unused]] __global__ __launch_bounds__(kShortBlock) void k_short_compact(
    int32_t m, int32_t q, int32_t nwaves, const int32_t* __restrict__ row_nnz, const int32_t* __restrict__ wave_tot,
    const int32_t* __restrict__ wg_tot, const int32_t* __restrict__ wg_prod, const int32_t* __restrict__ slot_crd,
    const double* __restrict__ slot_val,
    int32_t* __restrict__ C_pos, int32_t* __restrict__ C_crd, double* __restrict__ C_val,
    unsigned long long* __restrict__ out) {
  __shared__ int64_t s_part[2][kShortWaves];
  __shared__ int32_t s_bad[kShortWaves + 1];  // per wavefront, then the workgroup's verdict
  __shared__ int32_t s_off[kShortCompactWaves + 1];
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  const int32_t w0 = blockIdx.x * kShortCompactWaves;
  const int32_t w1 = min(w0 + kShortCompactWaves, nwaves);
  // entries of every earlier workgroup of the first kernel (one per
  // workgroup here too), and the products of all of them for the last
  // workgroup's statistics; a total of -1 marks a workgroup with a row that
  // did not fit
  static_assert(kShortCompactWaves == kShortWaves, "one compaction workgroup per first-kernel workgroup");
  int64_t sn = 0, sp = 0;
  bool bad = false;
  for (int32_t t = threadIdx.x; t < (int32_t)blockIdx.x; t += kShortBlock) {
    const int32_t x = wg_tot[t];
    sn += x;
    bad |= x < 0;
  }
  if (w1 == nwaves)
    for (int32_t t = threadIdx.x; t < (int32_t)gridDim.x; t += kShortBlock) sp += wg_prod[t];
  sn = wave_sum(sn);
  sp = wave_sum(sp);
  const bool wbad = __any(bad);
  if (lane == 0) {
    s_part[0][wv] = sn;
    s_part[1][wv] = sp;
    s_bad[wv] = wbad;
  }
  __syncthreads();
  if (wv == 0) {
    int64_t n = 0, pr = 0;
    bool any_bad = false;
#pragma unroll
    for (int i = 0; i < kShortWaves; ++i) {
      n += s_part[0][i];
      pr += s_part[1][i];
      any_bad |= s_bad[i] != 0;
    }
    // this workgroup's wavefronts (one per lane): exclusive offsets, and the
    // end at index w1 - w0 (lanes past w1 add nothing)
    const int32_t t = w0 + lane;
    const bool mine = lane < kShortCompactWaves && t < w1;
    const int32_t tot = mine ? wave_tot[t] : 0;
    any_bad |= __any(mine && tot < 0);
    const int32_t incl = wave_incl_scan(tot);
    if (lane <= kShortCompactWaves) s_off[lane] = (int32_t)n + incl - tot;
    if (lane == 0) s_bad[kShortWaves] = any_bad;
    if (w1 == nwaves && lane == kWave - 1) {
      if (!any_bad) C_pos[m] = (int32_t)n + incl;
      out[0] = (unsigned long long)pr;
      out[1] = any_bad ? 1ull : 0ull;
      out[2] = (unsigned long long)((int32_t)n + incl);
    }
  }
  __syncthreads();
  if (s_bad[kShortWaves]) return;  // a row did not fit: the caller recomputes C
  // one wavefront per first-kernel wavefront: row pointers, then the copy
  // (all of its packed entries loaded before any is stored)
  const int32_t t = w0 + wv;
  if (t >= w1) return;
  const int32_t r0 = (int32_t)min((int64_t)t * q, (int64_t)m);
  const int32_t nr = min(q, m - r0);
  const int32_t off = s_off[wv];
  const int32_t cnt = lane < nr ? row_nnz[r0 + lane] : 0;
  const int32_t incl = wave_incl_scan(cnt);
  if (lane < nr) C_pos[r0 + lane] = off + incl - cnt;
  const int32_t tot = s_off[wv + 1] - off;
  const int64_t base = (int64_t)t * q * kWave;
  constexpr int kU = kShortRows;  // at most kShortRows * 64 entries per wavefront
  int32_t c[kU];
  double v[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int32_t e = u * kWave + lane;
    c[u] = e < tot ? slot_crd[base + e] : 0;
    v[u] = e < tot ? slot_val[base + e] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int32_t e = u * kWave + lane;
    if (e < tot) {
      C_crd[off + e] = c[u];
      C_val[off + e] = v[u];
    }
  }
}

template <bool kValues>
__global__ __launch_bounds__(kCfBlock) void k_compact_fixed(int32_t m, const int32_t* __restrict__ row_nnz,
                                                            const int32_t* __restrict__ slot_crd,
                                                            const double* __restrict__ slot_val,
                                                            int32_t* __restrict__ C_pos, int32_t* __restrict__ C_crd,
                                                            double* __restrict__ C_val, int32_t* __restrict__ meta_zero) {
  __shared__ int32_t sh[kCfBlock / kWave + 1];
  // the product's metadata block, read by the host before this launch, is
  // left zero for the next product
  if (meta_zero && blockIdx.x == 0 && threadIdx.x < kMetaBlockInts) meta_zero[threadIdx.x] = 0;
  __shared__ int32_t pre[kCfRows], cnt[kCfRows];
  __shared__ int32_t s_base, s_total;
  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  const int32_t r0 = blockIdx.x * kCfRows;
  // counts of all earlier rows, four per 16-byte load (r0 is a multiple of
  // kCfRows, so r0 / 4 whole quads); kCfPart independent partial sums: the
  // loads are in flight together (one round per 8,192 earlier rows)
  const int4* __restrict__ nnz4 = reinterpret_cast<const int4*>(row_nnz);
  const int32_t q0 = r0 / 4;
  int32_t part[kCfPart];
#pragma unroll
  for (int u = 0; u < kCfPart; ++u) part[u] = 0;
  for (int32_t i0 = threadIdx.x; i0 < q0; i0 += kCfPart * kCfBlock) {
#pragma unroll
    for (int u = 0; u < kCfPart; ++u) {
      const int32_t i = i0 + u * kCfBlock;
      if (i < q0) {
        const int4 c4 = nnz4[i];
        part[u] += (c4.x + c4.y) + (c4.z + c4.w);
      }
    }
  }
  int32_t before = 0;
#pragma unroll
  for (int u = 0; u < kCfPart; ++u) before += part[u];
  int32_t tot;
  (void)block_excl_scan<int32_t, kCfBlock>(before, sh, &tot);
  if (w == 0) {
    const int32_t c = (lane < kCfRows && r0 + lane < m) ? row_nnz[r0 + lane] : 0;
    const int32_t incl = wave_incl_scan(c);
    pre[lane] = incl - c;
    cnt[lane] = c;
    if (lane == kWave - 1) s_total = incl;
    if (lane == 0) s_base = tot;
  }
  __syncthreads();
  const int32_t base = s_base;
  // every row of this wavefront is loaded before any is stored
  constexpr int kPer = kCfRows / (kCfBlock / kWave);
  int32_t cc[kPer];
  double vv[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int q = w + u * (kCfBlock / kWave);
    const int32_t r = r0 + q;
    if (kValues && r < m && lane < cnt[q]) {
      cc[u] = slot_crd[(int64_t)r * kWave + lane];
      vv[u] = slot_val[(int64_t)r * kWave + lane];
    }
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int q = w + u * (kCfBlock / kWave);
    const int32_t r = r0 + q;
    if (r >= m) break;
    const int32_t at = base + pre[q];
    if (lane == 0) C_pos[r] = at;
    if (kValues && lane < cnt[q]) {
      C_crd[at + lane] = cc[u];
      C_val[at + lane] = vv[u];
    }
  }
  if (threadIdx.x == 0 && r0 + kCfRows >= m) C_pos[m] = base + s_total;
}

// ------------------------------------------------------------------ workgroup-per-row kernel

template <int TS, int MAXN, int BLOCK, bool kValues>
__global__ __launch_bounds__(BLOCK) void k_row_block(const int32_t* __restrict__ bin_rows,
                                                     const int32_t* __restrict__ bmeta, int path,
                                                     const int32_t* __restrict__ Ap, const int32_t* __restrict__ Ai,
                                                     const double* __restrict__ Ax, const int32_t* __restrict__ Bp,
                                                     const int32_t* __restrict__ Bi, const double* __restrict__ Bx,
                                                     const int64_t* __restrict__ row_off,
                                                     int32_t* __restrict__ slot_crd, double* __restrict__ slot_val,
                                                     int32_t* __restrict__ row_nnz) {
  __shared__ int32_t keys[TS];
  __shared__ double vals[kValues ? TS : 1];
  __shared__ int32_t list[kValues ? MAXN : 1];
  __shared__ int32_t counter;
  constexpr uint32_t kMask = TS - 1;
  const uint32_t shift = 32 - log2_pow2(TS);
  const int wave = threadIdx.x / kWave;
  constexpr int kWaves = BLOCK / kWave;
  const int32_t count = bmeta[kMetaCount + path];
  const int32_t* rows = bin_rows + bmeta[kMetaOff + path];
  for (int32_t r = blockIdx.x; r < count; r += gridDim.x) {
    const int32_t i = rows[r];
    for (int s = threadIdx.x; s < TS; s += BLOCK) {
      keys[s] = -1;
      if (kValues) vals[s] = 0.0;
    }
    if (threadIdx.x == 0) counter = 0;
    __syncthreads();
    expand_products<kValues>(Ap[i], Ap[i + 1], wave, BLOCK / kWave, Ai, Ax, Bp, Bi, Bx,
                             [&](int32_t col, double v) { lds_insert<kValues>(keys, vals, shift, kMask, col, v); });
    __syncthreads();
    for (int s = threadIdx.x; s < TS; s += BLOCK) {
      const int32_t key = keys[s];
      if (key != -1) {
        const int pos = atomicAdd(&counter, 1);
        if (kValues) list[pos] = key;
      }
    }
    __syncthreads();
    const int32_t nnz = counter;
    if (!kValues) {
      if (threadIdx.x == 0) row_nnz[i] = nnz;
      __syncthreads();
      continue;
    }
    int P = 2;
    while (P < nnz) P <<= 1;
    for (int e = nnz + threadIdx.x; e < P; e += BLOCK) list[e] = INT_MAX;
    __syncthreads();
    block_bitonic_lds<BLOCK>(list, P);
    const int64_t off = row_off[i];
    for (int e = threadIdx.x; e < nnz; e += BLOCK) {
      const int32_t col = list[e];
      slot_crd[off + e] = col;
      slot_val[off + e] = lds_lookup(keys, vals, shift, kMask, col);
    }
    if (threadIdx.x == 0) row_nnz[i] = nnz;
    __syncthreads();
    (void)kWaves;
  }
}

// ------------------------------------------------------------------ bitmap kernel
//
// Rows whose output has at most kBitmapCols columns: one workgroup per row.
// Pass 1 over the row's products sets each column's bit in an LDS bitmap of
// all columns; a prefix of the word popcounts gives every column its rank in
// the sorted row; pass 2 adds each product at its column's rank.  The row is
// then written in column order with no hash probing and no sort.  Values
// accumulate in LDS (ds_add_f64); a long row (kGlobal) whose entries do not
// fit there accumulates straight into its slot region in HBM (global FP64
// atomics), which holds at least nnz entries.  In a stacked batch a
// row's columns lie in [row_colbase[i], row_colbase[i] + row_ncols[i]) (its
// request's width), and the passes over the bitmap cover that range only.

constexpr int kBitmapBlock = 256;
constexpr size_t kLdsBudget = 156 * 1024;  // dynamic LDS a workgroup may take (of 160 KiB)
constexpr int kBitmapLongBlock = 1024;  // rows beyond the block bins: 16 wavefronts per row
constexpr int32_t kBitmapCols = 1 << 17;

// Each thread owns 2^s consecutive bitmap words (the least s with
// BLOCK * 2^s >= words).  Word w is stored at w + (w >> s) when s > 0: one
// pad word after every thread's range, so the threads' ranges start in
// distinct LDS banks and the passes over owned words are conflict-free.
__host__ __device__ constexpr int bitmap_own_log2(int32_t words, int block) {
  int s = 0;
  while (((int64_t)block << s) < words) ++s;
  return s;
}

__host__ __device__ constexpr int32_t bitmap_padded_words(int32_t ncols, int block) {
  const int32_t words = (ncols + 31) >> 5;
  const int s = bitmap_own_log2(words, block);
  const int32_t padded = s ? words + (words >> s) + 1 : words;
  return (padded + 1) & ~1;
}

// bitmap and rank prefix (padded words each), then vcap FP64 accumulators
__host__ __device__ constexpr size_t bitmap_lds_bytes(int32_t ncols, int64_t vcap, bool values, int block) {
  return (size_t)bitmap_padded_words(ncols, block) * 8 + (values ? (size_t)vcap * 8 : 0);
}

// LDS layout of the bitmap kernels: bitmap and rank prefix (padded words
// each), then the FP64 row accumulator.
struct BitmapLds {
  int32_t padded, pshift;
  uint32_t *bm, *pre;
  double* acc;
  template <int BLOCK>
  __device__ __forceinline__ void init(double* base, int32_t ncols) {
    const int s = bitmap_own_log2((ncols + 31) >> 5, BLOCK);
    padded = bitmap_padded_words(ncols, BLOCK);
    pshift = s ? s : 31;
    bm = reinterpret_cast<uint32_t*>(base);
    pre = bm + padded;
    acc = reinterpret_cast<double*>(pre + padded);
    // the bitmap is zero between rows: the emit pass clears the words it reads
    for (int32_t w = threadIdx.x; w < padded; w += BLOCK) bm[w] = 0u;
    __syncthreads();
  }
  __device__ __forceinline__ int32_t at(int32_t w) const { return w + (w >> pshift); }
  // this thread's words of a row of ncols_row <= ncols columns
  template <int BLOCK>
  __device__ __forceinline__ void row_words(int32_t ncols_row, int32_t* rw0, int32_t* rw1) const {
    const int32_t wr = (ncols_row + 31) >> 5;
    const int s = bitmap_own_log2(wr, BLOCK);
    *rw0 = min((int32_t)threadIdx.x << s, wr);
    *rw1 = min(*rw0 + (1 << s), wr);
  }
};

// One row through the column bitmap (zero on entry and on return).
// pass1(f) and pass2(f) call f(col, v) once for every product of the row;
// pass 1 ignores v.  Values accumulate in the LDS accumulator, except with
// kGlobal when the row has more than acc_cap entries: then in the row's slot
// region with FP64 atomics.  The sorted column list is staged in the rank-prefix
// words (dead after pass 2) and stored coalesced when it fits there, else
// each thread stores its own columns.  Ends with a barrier, so the next row
// may reuse the LDS at once.
template <int BLOCK, bool kValues, bool kGlobal, class P1, class P2>
__device__ __forceinline__ void bitmap_row(int32_t i, int32_t cb, int32_t w0, int32_t w1, int64_t off,
                                           int32_t acc_cap, const BitmapLds& L, int32_t* sh,
                                           int32_t* __restrict__ slot_crd, double* __restrict__ slot_val,
                                           int32_t* __restrict__ row_nnz, P1&& pass1, P2&& pass2) {
  pass1([&](int32_t col, double) {
    const int32_t c = col - cb;
    atomicOr(&L.bm[L.at(c >> 5)], 1u << (c & 31));
  });
  __syncthreads();
  int32_t c = 0;
  for (int32_t w = w0; w < w1; ++w) c += __popc(L.bm[L.at(w)]);
  int32_t nnz;
  int32_t run = block_excl_scan<int32_t, BLOCK>(c, sh, &nnz);
  const int32_t first_rank = run;
  if (!kValues) {
    for (int32_t w = w0; w < w1; ++w) L.bm[L.at(w)] = 0u;
    if (threadIdx.x == 0) row_nnz[i] = nnz;
    __syncthreads();
    return;
  }
  for (int32_t w = w0; w < w1; ++w) {
    L.pre[L.at(w)] = run;
    run += __popc(L.bm[L.at(w)]);
  }
  const bool in_lds = !kGlobal || nnz <= acc_cap;  // workgroup-uniform
  double* dst = in_lds ? L.acc : slot_val + off;
  for (int32_t e = threadIdx.x; e < nnz; e += BLOCK) dst[e] = 0.0;
  __syncthreads();
  pass2([&](int32_t col, double v) {
    const int32_t cc = col - cb;
    const int32_t wa = L.at(cc >> 5);
    const int32_t rank = L.pre[wa] + __popc(L.bm[wa] & ((1u << (cc & 31)) - 1u));
    atomicAdd(&dst[rank], v);
  });
  __syncthreads();
  const bool staged = nnz <= L.padded;  // workgroup-uniform
  int32_t* list = reinterpret_cast<int32_t*>(L.pre);
  int32_t rk = first_rank;
  for (int32_t w = w0; w < w1; ++w) {
    const int32_t wa = L.at(w);
    uint32_t bits = L.bm[wa];
    L.bm[wa] = 0u;
    while (bits) {
      const int32_t col = cb + w * 32 + __builtin_ctz(bits);
      if (staged) list[rk] = col;
      else slot_crd[off + rk] = col;
      ++rk;
      bits &= bits - 1;
    }
  }
  if (staged) __syncthreads();
  for (int32_t e = threadIdx.x; e < nnz; e += BLOCK) {
    if (staged) slot_crd[off + e] = list[e];
    if (in_lds) slot_val[off + e] = L.acc[e];
  }
  if (threadIdx.x == 0) row_nnz[i] = nnz;
  __syncthreads();
}

template <bool kValues, bool kGlobal = false, int BLOCK = kBitmapBlock>
__global__ __launch_bounds__(BLOCK) void k_row_bitmap(const int32_t* __restrict__ bin_rows,
                                                      const int32_t* __restrict__ bmeta, int path, int32_t ncols,
                                                      const int32_t* __restrict__ Ap, const int32_t* __restrict__ Ai,
                                                      const double* __restrict__ Ax, const int32_t* __restrict__ Bp,
                                                      const int32_t* __restrict__ Bi, const double* __restrict__ Bx,
                                                      const int64_t* __restrict__ row_off,
                                                      int32_t* __restrict__ slot_crd, double* __restrict__ slot_val,
                                                      int32_t* __restrict__ row_nnz,
                                                      const int32_t* __restrict__ row_colbase,
                                                      const int32_t* __restrict__ row_ncols) {
  // declared as double so the dynamic region (which follows the static
  // arrays) starts 8-byte aligned: acc is updated with 64-bit LDS atomics
  extern __shared__ double bml_d[];
  __shared__ int32_t sh[BLOCK / kWave + 1];
  BitmapLds L;
  L.init<BLOCK>(bml_d, ncols);
  const int wave = threadIdx.x / kWave;
  const int32_t count = bmeta[kMetaCount + path];
  const int32_t* rows = bin_rows + bmeta[kMetaOff + path];
  for (int32_t r = blockIdx.x; r < count; r += gridDim.x) {
    const int32_t i = rows[r];
    const int32_t cb = row_colbase ? row_colbase[i] : 0;
    int32_t w0, w1;
    L.row_words<BLOCK>(row_ncols ? row_ncols[i] : ncols, &w0, &w1);
    const int32_t a_beg = Ap[i], a_end = Ap[i + 1];
    const int64_t off = kValues ? row_off[i] : 0;
    bitmap_row<BLOCK, kValues, kGlobal>(
        i, cb, w0, w1, off, 0, L, sh, slot_crd, slot_val, row_nnz,
        [&](auto&& f) { expand_products<false>(a_beg, a_end, wave, BLOCK / kWave, Ai, Ax, Bp, Bi, Bx, f); },
        [&](auto&& f) { expand_products<true>(a_beg, a_end, wave, BLOCK / kWave, Ai, Ax, Bp, Bi, Bx, f); });
  }
}

// Numeric rows beyond the block bins (at most kBitmapCols columns): pass 1
// gathers every product once, sets its column bit and stages (column,
// value) in the row's region [off, off + ub) of a scratch pair (the final
// output arrays, unused until the compaction), each wavefront taking slots
// from an LDS cursor; pass 2 reads the staged products back coalesced
// instead of gathering B again (B rows of power-law inputs are mostly 1-3
// entries long, so a second gather costs a cache line per product).  Values
// accumulate in LDS when the row has at most acc_cap entries (the dynamic
// LDS holds acc_cap accumulators after the bitmap), else with FP64 atomics
// in the row's slot region.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_row_bitmap_staged(
    const int32_t* __restrict__ bin_rows, const int32_t* __restrict__ bmeta, int path, int32_t ncols,
    const int32_t* __restrict__ Ap, const int32_t* __restrict__ Ai, const double* __restrict__ Ax,
    const int32_t* __restrict__ Bp, const int32_t* __restrict__ Bi, const double* __restrict__ Bx,
    const int64_t* __restrict__ row_off, int32_t* __restrict__ slot_crd, double* __restrict__ slot_val,
    int32_t* __restrict__ row_nnz, const int32_t* __restrict__ row_colbase, const int32_t* __restrict__ row_ncols,
    int32_t* __restrict__ stage_crd, double* __restrict__ stage_val, int32_t acc_cap) {
  extern __shared__ double bml_d[];
  __shared__ int32_t sh[BLOCK / kWave + 1];
  __shared__ int32_t cursor;
  if (threadIdx.x == 0) cursor = 0;
  BitmapLds L;
  L.init<BLOCK>(bml_d, ncols);  // ends with a barrier
  const int wave = threadIdx.x / kWave;
  const int lane = lane_id();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int32_t count = bmeta[kMetaCount + path];
  const int32_t* rows = bin_rows + bmeta[kMetaOff + path];
  for (int32_t r = blockIdx.x; r < count; r += gridDim.x) {
    const int32_t i = rows[r];
    const int32_t cb = row_colbase ? row_colbase[i] : 0;
    int32_t w0, w1;
    L.row_words<BLOCK>(row_ncols ? row_ncols[i] : ncols, &w0, &w1);
    const int32_t a_beg = Ap[i], a_end = Ap[i + 1];
    const int64_t off = row_off[i];
    const int32_t ub = (int32_t)(row_off[i + 1] - off);
    int32_t* sc = stage_crd + off;
    double* sv = stage_val + off;
    bitmap_row<BLOCK, true, true>(
        i, cb, w0, w1, off, acc_cap, L, sh, slot_crd, slot_val, row_nnz,
        [&](auto&& f) {
          expand_products<true>(a_beg, a_end, wave, BLOCK / kWave, Ai, Ax, Bp, Bi, Bx, [&](int32_t col, double v) {
            // the active lanes of this wavefront take consecutive slots
            const uint64_t act = __ballot(1);
            const int leader = __ffsll((long long)act) - 1;
            int32_t base = 0;
            if (lane == leader) base = atomicAdd(&cursor, __popcll(act));
            base = __shfl(base, leader, kWave);
            const int32_t slot = base + __popcll(act & lt);
            sc[slot] = col;
            sv[slot] = v;
            f(col, v);
          });
          __threadfence_block();  // the staged products are read by other wavefronts in pass 2
        },
        [&](auto&& f) {
          if (threadIdx.x == 0) cursor = 0;  // pass 1 of every wavefront is done
          constexpr int kU = 4;
          for (int32_t q0 = 0; q0 < ub; q0 += BLOCK * kU) {
            int32_t c[kU];
            double v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
              const int32_t q = q0 + u * BLOCK + (int32_t)threadIdx.x;
              if (q < ub) {
                c[u] = sc[q];
                v[u] = sv[q];
              }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
              if (q0 + u * BLOCK + (int32_t)threadIdx.x < ub) f(c[u], v[u]);
          }
        });
  }
}

// register capacity of the two block bins (products per thread), rows per
// step, and the largest grid of the numeric block-bin kernel.  The second
// bin holds rows of up to kBlock2UbMax products but keeps 2048 of them in
// registers, which keeps occupancy at the LDS limit.  One row per step:
// two rows per step overlap their load chains but cost more occupancy
// than they gain.
constexpr int kBitmapItems1 = kBlockUbMax / kBitmapBlock;
constexpr int kBitmapItems2 = 8;
constexpr int kBitmapRowsK = 1;
constexpr int kBitmapRowsGrid = 1 << 16;

// Numeric block-bin rows, K rows per step of a grid-stride loop: the
// products of every row with at most 64 A entries and at most
// kBitmapBlock * ITEMS products are gathered into registers (pass 2 reads
// no memory); the rows then go through the bitmap one after another.  Other
// rows take the two-gather expansion.
template <int K, int ITEMS>
__global__ __launch_bounds__(kBitmapBlock) void k_row_bitmap_rows(
    const int32_t* __restrict__ bin_rows, const int32_t* __restrict__ bmeta, int path, int32_t ncols,
    const int32_t* __restrict__ Ap, const int32_t* __restrict__ Ai, const double* __restrict__ Ax,
    const int32_t* __restrict__ Bp, const int32_t* __restrict__ Bi, const double* __restrict__ Bx,
    const int64_t* __restrict__ row_off, int32_t* __restrict__ slot_crd, double* __restrict__ slot_val,
    int32_t* __restrict__ row_nnz, const int32_t* __restrict__ row_colbase, const int32_t* __restrict__ row_ncols) {
  extern __shared__ double bml_d[];
  __shared__ int32_t sh[kBitmapBlock / kWave + 1];
  BitmapLds L;
  L.init<kBitmapBlock>(bml_d, ncols);
  const int wave = threadIdx.x / kWave;
  const int32_t count = bmeta[kMetaCount + path];
  const int32_t* rows = bin_rows + bmeta[kMetaOff + path];
  for (int32_t r0 = blockIdx.x * K; r0 < count; r0 += gridDim.x * K) {
    int32_t i[K], cb[K], nc[K], a_beg[K], a_end[K];
    int64_t off[K];
    bool take[K];
#pragma unroll
    for (int k = 0; k < K; ++k) i[k] = r0 + k < count ? rows[r0 + k] : -1;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      a_beg[k] = a_end[k] = cb[k] = 0;
      nc[k] = ncols;
      off[k] = 0;
      take[k] = false;
      if (i[k] >= 0) {
        a_beg[k] = Ap[i[k]];
        a_end[k] = Ap[i[k] + 1];
        off[k] = row_off[i[k]];
        // the row's upper bound is its slot-region length
        take[k] = row_off[i[k] + 1] - off[k] <= (int64_t)kBitmapBlock * ITEMS && a_end[k] - a_beg[k] <= kWave;
        cb[k] = row_colbase ? row_colbase[i[k]] : 0;
        if (row_ncols) nc[k] = row_ncols[i[k]];
      }
    }
    int32_t pc[K][ITEMS];
    double pv[K][ITEMS];
    gather_rows_products<kBitmapBlock, K, ITEMS, false>(take, a_beg, a_end, Ai, Ax, Bp, Bi, Bx, pc, pv);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (i[k] < 0) continue;
      int32_t w0, w1;
      L.row_words<kBitmapBlock>(nc[k], &w0, &w1);
      if (take[k]) {
        bitmap_row<kBitmapBlock, true, false>(
            i[k], cb[k], w0, w1, off[k], 0, L, sh, slot_crd, slot_val, row_nnz,
            [&](auto&& f) {
#pragma unroll
              for (int u = 0; u < ITEMS; ++u)
                if (pc[k][u] >= 0) f(pc[k][u], 0.0);
            },
            [&](auto&& f) {
#pragma unroll
              for (int u = 0; u < ITEMS; ++u)
                if (pc[k][u] >= 0) f(pc[k][u], pv[k][u]);
            });
      } else {
        const int32_t ab = a_beg[k], ae = a_end[k];
        bitmap_row<kBitmapBlock, true, false>(
            i[k], cb[k], w0, w1, off[k], 0, L, sh, slot_crd, slot_val, row_nnz,
            [&](auto&& f) { expand_products<false>(ab, ae, wave, kBitmapBlock / kWave, Ai, Ax, Bp, Bi, Bx, f); },
            [&](auto&& f) { expand_products<true>(ab, ae, wave, kBitmapBlock / kWave, Ai, Ax, Bp, Bi, Bx, f); });
      }
    }
  }
}

// ------------------------------------------------------------------ dense accumulator kernel
//
// Each workgroup owns dflag[ncols] (uint8) and dval[ncols] (FP64) in HBM,
// both zero between rows.  Columns are emitted in order, so no sort is
// needed; the emit pass also clears what it read.

constexpr int kDenseBlock = 256;
constexpr int kDenseItems = 16;

template <bool kValues>
__global__ __launch_bounds__(kDenseBlock) void k_row_dense(const int32_t* __restrict__ bin_rows,
                                                           const int32_t* __restrict__ bmeta, int path,
                                                           int32_t ncols, const int32_t* __restrict__ Ap,
                                                           const int32_t* __restrict__ Ai, const double* __restrict__ Ax,
                                                           const int32_t* __restrict__ Bp, const int32_t* __restrict__ Bi,
                                                           const double* __restrict__ Bx,
                                                           const int64_t* __restrict__ row_off,
                                                           int32_t* __restrict__ slot_crd, double* __restrict__ slot_val,
                                                           int32_t* __restrict__ row_nnz, uint8_t* __restrict__ dflag_all,
                                                           double* __restrict__ dval_all) {
  __shared__ int32_t sh[kDenseBlock / kWave + 1];
  uint8_t* dflag = dflag_all + (size_t)blockIdx.x * (size_t)ncols;
  double* dval = kValues ? dval_all + (size_t)blockIdx.x * (size_t)ncols : nullptr;
  const int wave = threadIdx.x / kWave;
  const int32_t count = bmeta[kMetaCount + path];
  const int32_t* rows = bin_rows + bmeta[kMetaOff + path];
  for (int32_t r = blockIdx.x; r < count; r += gridDim.x) {
    const int32_t i = rows[r];
    expand_products<kValues>(Ap[i], Ap[i + 1], wave, kDenseBlock / kWave, Ai, Ax, Bp, Bi, Bx,
                             [&](int32_t col, double v) {
                               dflag[col] = 1;
                               if (kValues) atomicAdd(&dval[col], v);
                             });
    __threadfence();
    __syncthreads();
    const int64_t off = kValues ? row_off[i] : 0;
    int32_t run = 0;
    for (int32_t base = 0; base < ncols; base += kDenseBlock * kDenseItems) {
      const int32_t mine = base + threadIdx.x * kDenseItems;
      uint32_t bits = 0;
      int32_t cnt = 0;
#pragma unroll
      for (int it = 0; it < kDenseItems; ++it) {
        const int32_t c = mine + it;
        if (c < ncols && dflag[c]) {
          bits |= 1u << it;
          ++cnt;
        }
      }
      int32_t tot;
      int32_t pre = block_excl_scan<int32_t, kDenseBlock>(cnt, sh, &tot);
      if (bits) {
        int32_t pos = run + pre;
#pragma unroll
        for (int it = 0; it < kDenseItems; ++it) {
          if (bits & (1u << it)) {
            const int32_t c = mine + it;
            if (kValues) {
              slot_crd[off + pos] = c;
              slot_val[off + pos] = dval[c];
              dval[c] = 0.0;
            }
            dflag[c] = 0;
            ++pos;
          }
        }
      }
      run += tot;
    }
    if (threadIdx.x == 0) row_nnz[i] = run;
    __threadfence();
    __syncthreads();
  }
}

// ------------------------------------------------------------------ HBM hash kernel (long rows)
//
// Each workgroup owns a table of tab_slots slots (int32 key, FP64 value);
// a row uses the first next_pow2(2*ub) of them.  After accumulation the
// occupied slots are compacted into the row's slot region and sorted there:
// in LDS when nnz <= kWideLdsSort, else with a global-memory bitonic network.

constexpr int kWideBlock = 256;

__device__ __forceinline__ void wide_insert(int32_t* keys, double* vals, uint32_t shift, uint32_t mask, int32_t col,
                                            double v, bool with_values) {
  uint32_t h = hash_col((uint32_t)col, shift);
  for (;;) {
    int32_t cur = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == -1) {
      cur = atomicCAS(&keys[h], -1, col);
      if (cur == -1) cur = col;
    }
    if (cur == col) {
      if (with_values) atomicAdd(&vals[h], v);
      return;
    }
    h = (h + 1) & mask;
  }
}

template <bool kValues>
__global__ __launch_bounds__(kWideBlock) void k_row_wide(const int32_t* __restrict__ bin_rows,
                                                         const int32_t* __restrict__ bmeta, int path,
                                                         const int32_t* __restrict__ Ap, const int32_t* __restrict__ Ai,
                                                         const double* __restrict__ Ax, const int32_t* __restrict__ Bp,
                                                         const int32_t* __restrict__ Bi, const double* __restrict__ Bx,
                                                         const int64_t* __restrict__ row_ub,
                                                         const int64_t* __restrict__ row_off,
                                                         int32_t* __restrict__ slot_crd, double* __restrict__ slot_val,
                                                         int32_t* __restrict__ row_nnz, int32_t* __restrict__ tab_keys_all,
                                                         double* __restrict__ tab_vals_all, int64_t tab_slots) {
  __shared__ int32_t skeys[kValues ? kWideLdsSort : 1];
  __shared__ double svals[kValues ? kWideLdsSort : 1];
  __shared__ int32_t counter;
  int32_t* tkeys = tab_keys_all + (size_t)blockIdx.x * (size_t)tab_slots;
  double* tvals = kValues ? tab_vals_all + (size_t)blockIdx.x * (size_t)tab_slots : nullptr;
  const int wave = threadIdx.x / kWave;
  const int32_t count = bmeta[kMetaCount + path];
  const int32_t* rows = bin_rows + bmeta[kMetaOff + path];
  for (int32_t r = blockIdx.x; r < count; r += gridDim.x) {
    const int32_t i = rows[r];
    const int64_t ub = row_ub[i];
    uint32_t ts = next_pow2_u32((uint32_t)(2 * ub));
    if (ts < 64) ts = 64;
    const uint32_t mask = ts - 1;
    const uint32_t shift = 32 - log2_pow2(ts);
    for (uint32_t s = threadIdx.x; s < ts; s += kWideBlock) {
      tkeys[s] = -1;
      if (kValues) tvals[s] = 0.0;
    }
    if (threadIdx.x == 0) counter = 0;
    __threadfence();
    __syncthreads();
    expand_products<kValues>(Ap[i], Ap[i + 1], wave, kWideBlock / kWave, Ai, Ax, Bp, Bi, Bx,
                             [&](int32_t col, double v) { wide_insert(tkeys, tvals, shift, mask, col, v, kValues); });
    __threadfence();
    __syncthreads();
    const int64_t off = kValues ? row_off[i] : 0;
    for (uint32_t s = threadIdx.x; s < ts; s += kWideBlock) {
      const int32_t key = __hip_atomic_load(&tkeys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (key != -1) {
        const int pos = atomicAdd(&counter, 1);
        if (kValues) {
          slot_crd[off + pos] = key;
          slot_val[off + pos] = __hip_atomic_load(&tvals[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    __threadfence();
    __syncthreads();
    const int32_t nnz = counter;
    if (kValues && nnz > 1) {
      int32_t* rc = slot_crd + off;
      double* rv = slot_val + off;
      if (nnz <= kWideLdsSort) {
        int P = 2;
        while (P < nnz) P <<= 1;
        for (int e = threadIdx.x; e < P; e += kWideBlock) {
          skeys[e] = e < nnz ? rc[e] : INT_MAX;
          svals[e] = e < nnz ? rv[e] : 0.0;
        }
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1) {
          for (int j = k >> 1; j >= 1; j >>= 1) {
            for (int t = threadIdx.x; t < P / 2; t += kWideBlock) {
              int lo, hi;
              if (j == (k >> 1)) {
                const int b = t / j, o = t % j;
                lo = b * k + o;
                hi = b * k + k - 1 - o;
              } else {
                const int b = t / j, o = t % j;
                lo = b * 2 * j + o;
                hi = lo + j;
              }
              if (skeys[lo] > skeys[hi]) {
                const int32_t tk = skeys[lo];
                skeys[lo] = skeys[hi];
                skeys[hi] = tk;
                const double tv = svals[lo];
                svals[lo] = svals[hi];
                svals[hi] = tv;
              }
            }
            __syncthreads();
          }
        }
        for (int e = threadIdx.x; e < nnz; e += kWideBlock) {
          rc[e] = skeys[e];
          rv[e] = svals[e];
        }
      } else {
        int P = 2;
        while (P < nnz) P <<= 1;
        for (int k = 2; k <= P; k <<= 1) {
          for (int j = k >> 1; j >= 1; j >>= 1) {
            for (int t = threadIdx.x; t < P / 2; t += kWideBlock) {
              int lo, hi;
              if (j == (k >> 1)) {
                const int b = t / j, o = t % j;
                lo = b * k + o;
                hi = b * k + k - 1 - o;
              } else {
                const int b = t / j, o = t % j;
                lo = b * 2 * j + o;
                hi = lo + j;
              }
              if (hi < nnz) {
                const int32_t x = rc[lo], y = rc[hi];
                if (x > y) {
                  rc[lo] = y;
                  rc[hi] = x;
                  const double tv = rv[lo];
                  rv[lo] = rv[hi];
                  rv[hi] = tv;
                }
              }
            }
            __threadfence();
            __syncthreads();
          }
        }
      }
    }
    if (threadIdx.x == 0) row_nnz[i] = nnz;
    __threadfence();
    __syncthreads();
  }
}

// ------------------------------------------------------------------ compaction

constexpr int kCompactUnroll = 8;

// (the batch engine compacts with its own kernel, which writes into request
// output buffers, so this one is unused there)
[[maybe_unused]] __global__ __launch_bounds__(256) void k_compact(int32_t m_host, const int32_t* __restrict__ m_dev,
                                                 const int64_t* __restrict__ row_off,
                                                 const int32_t* __restrict__ C_pos,
                                                 const int32_t* __restrict__ slot_crd,
                                                 const double* __restrict__ slot_val,
                                                 const int32_t* __restrict__ row_coloff, int32_t* __restrict__ C_crd,
                                                 double* __restrict__ C_val, int32_t* __restrict__ meta_zero) {
  // the product's metadata block, read by the host before this launch, is
  // left zero for the next product
  if (meta_zero && blockIdx.x == 0 && threadIdx.x < kMetaBlockInts) meta_zero[threadIdx.x] = 0;
  const int32_t m = dev_or(m_dev, m_host);
  const int lane = lane_id();
  const int32_t waves = gridDim.x * (blockDim.x / kWave);
  for (int32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < m; i += waves) {
    const int32_t dst = C_pos[i];
    const int32_t n = C_pos[i + 1] - dst;
    const int64_t src = row_off[i];
    const int32_t shift = row_coloff ? row_coloff[i] : 0;
    // kCompactUnroll strides loaded before any is stored: the copy is
    // bandwidth-bound, not bound by one load's latency per stride
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
          C_crd[dst + e] = c[u] - shift;
          C_val[dst + e] = v[u];
        }
      }
    }
  }
}


}  // namespace
}  // namespace spws
