// TTM by roots and column windows:
//   Y(i,j,r) = sum_k X(i,k,j) U(k,r)   for a CSF with levels (i, k, j),
// every output fiber (i, j) summed over k in increasing order -- the order in
// which the reference's hash_il workspace accumulates it
// (benchmark/TTM/COO_CSF_DCSR_hash_il.h:247-377) -- with no sort of the
// whole tensor.  The nonzeros of a root are contiguous in the CSF and come
// in (k, j) order, so only the nonzeros within each root need grouping by j.
//
// Light roots (at most kTwLight nonzeros): one wavefront per root sorts the
// root's (j, position) keys in registers (wavefront bitonic); each run of
// equal j is one output fiber, its nonzeros still in CSF order.
// Heavy roots: their nonzeros go into buckets (root, window of S columns)
// by one stable counting pass over tiles of kTwTile nonzeros -- per-tile
// window histograms, laid out (root, window, tile) so that one prefix sum
// gives every tile its bucket offsets, then a scatter ranked within each
// 64-nonzero step by ballots.  A window's nonzeros keep the CSF order.
// One wavefront per window then adds them in that order into a dense LDS
// window of S x R doubles and emits the window's columns in order.
// Fiber counts (light: runs after the sort; heavy: the columns each window
// holds) give every root its output offset by a prefix over the roots, so
// the fibers land in (i, j) order, as the sorting path lays them out.
// Every sum starts from 0.0 and adds X * U(k, r) in increasing k: the
// results equal the sorting path's bit for bit.
#include <algorithm>
#include <climits>

#include "tensor.hpp"

namespace spws {
namespace {

constexpr int kTwLight = 256;    // nonzeros of a light root at most (4 per lane)
constexpr int kTwTile = 2048;    // nonzeros per heavy tile
constexpr int kTwMaxWin = 4096;  // column windows per root at most (LDS counters)
constexpr int kTwDense = 32;     // nonzeros per heavy window for the automatic choice

// Last index p in [0, n) with key(p) <= x, for key non-decreasing and
// key(0) <= x: one wavefront, 64-way steps.
template <class F>
__device__ int32_t wave_last_le(F key, int32_t n, int64_t x) {
  const int lane = lane_id();
  int32_t lo = 0, hi = n - 1;
  while (hi - lo + 1 > kWave) {
    const int64_t span = (int64_t)(hi - lo) + 1;
    const int32_t p = lo + (int32_t)(span * lane / kWave);
    const uint64_t b = __ballot(key(p) <= x);
    const int t = 63 - __builtin_clzll(b | 1ull);
    const int32_t nlo = lo + (int32_t)(span * t / kWave);
    const int32_t nhi = t == kWave - 1 ? hi : lo + (int32_t)(span * (t + 1) / kWave) - 1;
    lo = nlo;
    hi = nhi;
  }
  const int32_t p = lo + lane;
  const bool ok = p <= hi ? key(p) <= x : false;
  const uint64_t b = __ballot(ok);
  return lo + (63 - __builtin_clzll(b | 1ull));
}

// Position of the n-th (from 0) set bit of m (n < popcount(m)).
__device__ __forceinline__ int select_bit(uint64_t m, int n) {
  int pos = 0;
#pragma unroll
  for (int sh = 32; sh > 0; sh >>= 1) {
    const uint64_t low = m & ((1ull << sh) - 1ull);
    const int c = __popcll(low);
    if (n >= c) {
      n -= c;
      m >>= sh;
      pos += sh;
    } else {
      m = low;
    }
  }
  return pos;
}

__device__ __forceinline__ uint64_t wave_or(uint64_t v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) v |= __shfl_xor(v, d, kWave);
  return v;
}

// Per root: (tiles << 32) | heavy, tiles = ceil(n / kTwTile) for a heavy
// root (n > kTwLight nonzeros), 0 otherwise.  Its exclusive prefix gives
// each heavy root its first tile (high word) and its heavy rank (low word).
__global__ __launch_bounds__(256) void k_tw_roots(int32_t nroots, const int32_t* __restrict__ l2_pos,
                                                  const int32_t* __restrict__ l3_pos, int64_t* __restrict__ pk) {
  for (int32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nroots; r += gridDim.x * blockDim.x) {
    const int32_t n = l3_pos[l2_pos[r + 1]] - l3_pos[l2_pos[r]];
    const bool heavy = n > kTwLight;
    const int64_t tiles = heavy ? (n + kTwTile - 1) / kTwTile : 0;
    pk[r] = (tiles << 32) | (heavy ? 1 : 0);
  }
}

// The heavy tile t of this workgroup: its root, the root's first tile and
// tile count, and the tile's nonzeros [e0, e1).
struct TwTile {
  int32_t r, tiles, c, e0, e1;
  int64_t tb;
};

__device__ __forceinline__ TwTile tw_tile(int64_t t, int32_t nroots, const int64_t* __restrict__ pko,
                                          const int32_t* __restrict__ l2_pos, const int32_t* __restrict__ l3_pos) {
  TwTile T;
  T.r = wave_last_le([&](int32_t p) { return pko[p] >> 32; }, nroots, t);
  T.tb = pko[T.r] >> 32;
  T.tiles = (int32_t)((pko[T.r + 1] >> 32) - T.tb);
  T.c = (int32_t)(t - T.tb);
  const int32_t lb = l3_pos[l2_pos[T.r]], le = l3_pos[l2_pos[T.r + 1]];
  T.e0 = lb + T.c * kTwTile;
  T.e1 = min(le, T.e0 + kTwTile);
  return T;
}

// Window histogram of one heavy tile: hist[W tb + w tiles + c] (layout
// root, window, tile).  The first tile of a root also records it in hroot.
__global__ __launch_bounds__(kWave) void k_tw_hist(int32_t nroots, const int64_t* __restrict__ pko,
                                                   const int32_t* __restrict__ l2_pos,
                                                   const int32_t* __restrict__ l3_pos,
                                                   const int32_t* __restrict__ l3_crd, int s, int32_t W,
                                                   int32_t* __restrict__ hist, int32_t* __restrict__ hroot) {
  __shared__ int32_t cnt[kTwMaxWin];
  const int lane = lane_id();
  const TwTile T = tw_tile(blockIdx.x, nroots, pko, l2_pos, l3_pos);
  if (T.c == 0 && lane == 0) hroot[(int32_t)(pko[T.r] & 0xffffffffll)] = T.r;
  for (int w = lane; w < W; w += kWave) cnt[w] = 0;
  __syncthreads();
  for (int32_t e = T.e0 + lane; e < T.e1; e += kWave) atomicAdd(&cnt[l3_crd[e] >> s], 1);
  __syncthreads();
  const int64_t hb = (int64_t)W * T.tb + T.c;
  for (int w = lane; w < W; w += kWave) hist[hb + (int64_t)w * T.tiles] = cnt[w];
}

// Stable scatter of one heavy tile into its (root, window) buckets: key
// (k << s) | (j mod S), value X.  Within each 64-nonzero step the nonzeros
// of one window are ranked by lane (ballots over the window bits), so every
// bucket keeps the CSF order.
__global__ __launch_bounds__(kWave) void k_tw_scatter(int32_t nroots, const int64_t* __restrict__ pko,
                                                      const int32_t* __restrict__ l2_pos,
                                                      const int32_t* __restrict__ l2_crd,
                                                      const int32_t* __restrict__ l3_pos,
                                                      const int32_t* __restrict__ l3_crd,
                                                      const double* __restrict__ vals, int s, int32_t W, int wbits,
                                                      const int32_t* __restrict__ hoff, uint32_t* __restrict__ hkey,
                                                      double* __restrict__ hval) {
  __shared__ int32_t cnt[kTwMaxWin];
  __shared__ int32_t spos[kTwTile + 1];
  __shared__ int32_t spk[kTwTile];
  const int lane = lane_id();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const TwTile T = tw_tile(blockIdx.x, nroots, pko, l2_pos, l3_pos);
  const int64_t hb = (int64_t)W * T.tb + T.c;
  for (int w = lane; w < W; w += kWave) cnt[w] = hoff[hb + (int64_t)w * T.tiles];
  // the root's k-fibers that cover the tile, staged for the owner searches
  const int32_t f0 = l2_pos[T.r], nf = l2_pos[T.r + 1] - f0;
  auto fpos = [&](int32_t p) { return (int64_t)l3_pos[f0 + p]; };
  const int32_t fa = wave_last_le(fpos, nf, T.e0);
  const int32_t fb = wave_last_le(fpos, nf, T.e1 - 1);
  const int32_t span = fb - fa + 1;
  const bool staged = span <= kTwTile;
  if (staged) {
    for (int i = lane; i <= span; i += kWave) spos[i] = l3_pos[f0 + fa + i];
    for (int i = lane; i < span; i += kWave) spk[i] = l2_crd[f0 + fa + i];
  }
  __syncthreads();
  const uint32_t smask = (1u << s) - 1u;
  // the next step's column and value load while this step is placed
  int32_t j_next = T.e0 + lane < T.e1 ? l3_crd[T.e0 + lane] : 0;
  double x_next = T.e0 + lane < T.e1 ? vals[T.e0 + lane] : 0.0;
  for (int32_t eb = T.e0; eb < T.e1; eb += kWave) {
    const int32_t e = eb + lane;
    const bool valid = e < T.e1;
    const int32_t j = j_next;
    const double xv = x_next;
    if (eb + kWave + lane < T.e1) {
      j_next = l3_crd[eb + kWave + lane];
      x_next = vals[eb + kWave + lane];
    }
    int32_t k = 0;
    if (valid) {
      int32_t lo = 0, hi = span - 1;  // last fiber of the span starting at or before e
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        const int32_t pv = staged ? spos[mid] : l3_pos[f0 + fa + mid];
        if (pv <= e) lo = mid; else hi = mid - 1;
      }
      k = staged ? spk[lo] : l2_crd[f0 + fa + lo];
    }
    const int32_t w = valid ? (j >> s) : 0;
    uint64_t eq = __ballot(valid);
    for (int b = 0; b < wbits; ++b) {
      const bool bit = (w >> b) & 1;
      const uint64_t bb = __ballot(bit);
      eq &= bit ? bb : ~bb;
    }
    const int rank = __popcll(eq & lt);
    const int gsz = __popcll(eq);
    int32_t base = 0;
    if (valid) base = cnt[w];
    __builtin_amdgcn_wave_barrier();
    if (valid && rank == 0) cnt[w] = base + gsz;
    __builtin_amdgcn_wave_barrier();
    if (valid) {
      hkey[base + rank] = ((uint32_t)k << s) | ((uint32_t)j & smask);
      hval[base + rank] = xv;
    }
  }
}

// Start of every heavy window wi = hr * W + w in the bucketed nonzeros
// (windows are contiguous there, in (root, window) order), and the total
// at wstart[nwin].
__global__ __launch_bounds__(256) void k_tw_wstart(int64_t nwin, int32_t W, const int32_t* __restrict__ hroot,
                                                   const int64_t* __restrict__ pko,
                                                   const int32_t* __restrict__ hoff, int64_t hn,
                                                   int32_t* __restrict__ wstart) {
  for (int64_t wi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; wi <= nwin;
       wi += (int64_t)gridDim.x * blockDim.x) {
    if (wi == nwin) {
      wstart[nwin] = hoff[hn];
      continue;
    }
    const int32_t hr = (int32_t)(wi / W), w = (int32_t)(wi % W);
    const int32_t r = hroot[hr];
    const int64_t tb = pko[r] >> 32;
    const int64_t tiles = (pko[r + 1] >> 32) - tb;
    wstart[wi] = hoff[(int64_t)W * tb + (int64_t)w * tiles];
  }
}

// Work item c of the heavy windows: the windows whose nonzeros start in
// [c * kTwItem, (c + 1) * kTwItem) -- about kTwItem nonzeros per item (a
// longer window is one item by itself), so the wavefronts get balanced work
// however the columns cluster.
constexpr int kTwItem = 1024;

__device__ __forceinline__ void tw_item_windows(int64_t c, int64_t nwin, const int32_t* __restrict__ wstart,
                                                int64_t* w_lo, int64_t* w_hi) {
  // first window starting at or after x: one past the last one starting before it
  auto first_ge = [&](int64_t x) -> int64_t {
    if (x <= 0) return 0;
    const int32_t last = wave_last_le([&](int32_t p) { return (int64_t)wstart[p]; }, (int32_t)nwin, x - 1);
    return last + 1;
  };
  *w_lo = first_ge(c * kTwItem);
  *w_hi = first_ge((c + 1) * kTwItem);
}

// Columns held by each heavy window: dcnt[wi].
__global__ __launch_bounds__(kWave) void k_tw_wcount(int64_t nwin, int s, const int32_t* __restrict__ wstart,
                                                     const uint32_t* __restrict__ hkey, int32_t* __restrict__ dcnt) {
  const int lane = lane_id();
  int64_t w_lo, w_hi;
  tw_item_windows(blockIdx.x, nwin, wstart, &w_lo, &w_hi);
  const uint32_t smask = (1u << s) - 1u;
  for (int64_t wb = w_lo; wb < w_hi; wb += kWave) {
    // 64 windows' bounds at once, one per lane
    const int64_t wi = wb + lane;
    const int32_t ws = wi < w_hi ? wstart[wi] : 0, we = wi < w_hi ? wstart[wi + 1] : 0;
    int32_t mine = 0;
    const int nq = (int)min((int64_t)kWave, w_hi - wb);
    for (int q = 0; q < nq; ++q) {
      const int32_t a = __shfl(ws, q, kWave), b = __shfl(we, q, kWave);
      uint64_t occ = 0;
      for (int32_t e = a + lane; e < b; e += kWave) occ |= 1ull << (hkey[e] & smask);
      occ = wave_or(occ);
      if (lane == q) mine = __popcll(occ);
    }
    if (wi < w_hi) dcnt[wi] = mine;
  }
}

// Fiber count of each root: light roots sort their columns (wavefront
// bitonic over 256 keys) and count the distinct ones; heavy roots add up
// their windows' counts (from the prefix doff).
__global__ __launch_bounds__(kWave) void k_tw_light_count(int32_t nroots, int32_t W, const int64_t* __restrict__ pko,
                                                          const int32_t* __restrict__ doff,
                                                          const int32_t* __restrict__ l2_pos,
                                                          const int32_t* __restrict__ l3_pos,
                                                          const int32_t* __restrict__ l3_crd,
                                                          int32_t* __restrict__ fcount) {
  const int lane = lane_id();
  const int32_t r = blockIdx.x;
  const int32_t lb = l3_pos[l2_pos[r]], le = l3_pos[l2_pos[r + 1]];
  const int32_t n = le - lb;
  if (n > kTwLight) {
    if (lane == 0) {
      const int64_t hr = pko[r] & 0xffffffffll;
      fcount[r] = doff[(hr + 1) * W] - doff[hr * W];
    }
    return;
  }
  if (n == 0) {
    if (lane == 0) fcount[r] = 0;
    return;
  }
  int32_t v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int idx = u * kWave + lane;
    v[u] = idx < n ? l3_crd[lb + idx] : INT_MAX;
  }
  int n_pow2 = 2;
  while (n_pow2 < n) n_pow2 <<= 1;
  wave_bitonic<4>(v, n_pow2);
  int32_t c = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int idx = u * kWave + lane;
    const int32_t up = __shfl_up(v[u], 1, kWave);
    const int32_t carry = u > 0 ? __shfl(v[u > 0 ? u - 1 : 0], kWave - 1, kWave) : INT_MIN;
    const int32_t prev = lane > 0 ? up : carry;
    if (idx < n && (idx == 0 || v[u] != prev)) ++c;
  }
  c = wave_sum(c);
  if (lane == 0) fcount[r] = c;
}

// Heavy windows, in order: a dense LDS window of S columns x RP lanes
// (8 KB); G = 64 / RP nonzeros per step, lane (g, r) computing X * U(k, r)
// of the step's g-th nonzero.  The step's nonzeros are added in order: at
// once when their columns differ, one after another when two share one.
// The U rows of a 64-nonzero stride are all loaded before the first add.
// Then the window's columns in increasing order -> fibers, and the used
// accumulators back to 0.0 for the next window.
template <int RP>
__global__ __launch_bounds__(kWave) void k_tw_heavy(int64_t nwin, int32_t W, int s, int R,
                                                    const int32_t* __restrict__ hroot,
                                                    const int64_t* __restrict__ pko,
                                                    const int32_t* __restrict__ wstart,
                                                    const uint32_t* __restrict__ hkey,
                                                    const double* __restrict__ hval,
                                                    const int32_t* __restrict__ doff,
                                                    const int32_t* __restrict__ root_base,
                                                    const int32_t* __restrict__ l1_crd,
                                                    const double* __restrict__ U, int32_t* __restrict__ out_i,
                                                    int32_t* __restrict__ out_j, double* __restrict__ Y) {
  constexpr int G = kWave / RP;
  constexpr int S = 1024 / RP;  // columns per window (2^s)
  constexpr int NU = 8;         // steps whose U rows are loaded together
  __shared__ double acc[S * RP];
  const int lane = lane_id();
  const int g = lane / RP, rr = lane % RP;
  const bool rok = rr < R;
  const int rc = rok ? rr : 0;
  for (int i = lane; i < S * RP; i += kWave) acc[i] = 0.0;
  __builtin_amdgcn_wave_barrier();
  int64_t w_lo, w_hi;
  tw_item_windows(blockIdx.x, nwin, wstart, &w_lo, &w_hi);
  for (int64_t wb = w_lo; wb < w_hi; wb += kWave) {
  // 64 windows' bounds and output places at once, one per lane
  const int64_t wi = wb + lane;
  int32_t ws = 0, we = 0, obase = 0, ii = 0, wcol = 0;
  if (wi < w_hi) {
    ws = wstart[wi];
    we = wstart[wi + 1];
    const int64_t hr = wi / W;
    const int32_t r = hroot[hr];
    obase = root_base[r] + (doff[wi] - doff[hr * W]);
    ii = l1_crd[r];
    wcol = (int32_t)(wi % W) * S;
  }
  const int nq = (int)min((int64_t)kWave, w_hi - wb);
  const uint32_t smask = (uint32_t)S - 1u;
  for (int q = 0; q < nq; ++q) {
    const int32_t a = __shfl(ws, q, kWave), b = __shfl(we, q, kWave);
    if (a >= b) continue;  // uniform
    uint64_t occ = 0;
    for (int32_t sb = a; sb < b; sb += kWave) {
      const int32_t e = sb + lane;
      const int32_t cnt = min(kWave, b - sb);
      const uint32_t kv = e < b ? hkey[e] : 0u;
      const double xv = e < b ? hval[e] : 0.0;
      if (e < b) occ |= 1ull << (kv & smask);
      for (int t0 = 0; t0 * G < cnt; t0 += NU) {
      double uv[NU];
#pragma unroll
      for (int t1 = 0; t1 < NU; ++t1) {
        const int idx = (t0 + t1) * G + g;
        const uint32_t kk = __shfl(kv, idx & (kWave - 1), kWave);
        uv[t1] = idx < cnt ? U[(int64_t)(kk >> s) * R + rc] : 0.0;
      }
#pragma unroll
      for (int t1 = 0; t1 < NU; ++t1) {
        const int t = t0 + t1;
        if (t * G < cnt) {  // uniform
          const int idx = t * G + g;
          const uint32_t kk = __shfl(kv, idx, kWave);
          const double xx = __shfl(xv, idx, kWave);
          const int jl = (int)(kk & smask);
          const bool act = idx < cnt && rok;
          const double v = xx * uv[t1];
          bool distinct = true;
          if (G > 1) {
            int cols[G];
#pragma unroll
            for (int h = 0; h < G; ++h) {
              const int ih = t * G + h;
              cols[h] = ih < cnt ? (int)(__shfl(kv, ih & (kWave - 1), kWave) & smask) : -1 - h;
            }
#pragma unroll
            for (int h1 = 0; h1 < G; ++h1)
#pragma unroll
              for (int h2 = h1 + 1; h2 < G; ++h2) distinct = distinct && cols[h1] != cols[h2];
          }
          if (distinct) {
            if (act) acc[jl * RP + rr] += v;
          } else {
#pragma unroll
            for (int h = 0; h < G; ++h) {
              if (g == h && act) acc[jl * RP + rr] += v;
              __builtin_amdgcn_wave_barrier();
            }
          }
          __builtin_amdgcn_wave_barrier();
        }
      }
      }
    }
    occ = wave_or(occ);
    const int nfib = __popcll(occ);
    const int32_t ob = __shfl(obase, q, kWave), iq = __shfl(ii, q, kWave), wc = __shfl(wcol, q, kWave);
    for (int f0 = 0; f0 < nfib; f0 += G) {
      const int fi = f0 + g;
      if (fi < nfib) {
        const int jl = select_bit(occ, fi);
        if (rok) Y[(int64_t)(ob + fi) * R + rr] = acc[jl * RP + rr];
        acc[jl * RP + rr] = 0.0;
        if (rr == 0) {
          out_i[ob + fi] = iq;
          out_j[ob + fi] = wc + jl;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  }
}

// Light roots: one wavefront per root.  The root's nonzeros (at most 256)
// with their k (owner k-fiber by binary search in the staged fiber bounds)
// are sorted by (j, position); each run of equal j is one fiber.  G
// nonzeros per step in sorted order, lane (g, r) computing X * U(k, r); the
// running sum passes from group to group within the step and on to the next
// step, restarting at 0.0 where a run starts; a run's last nonzero writes
// the fiber.
template <int RP>
__global__ __launch_bounds__(kWave) void k_tw_light(int32_t nroots, int R, const int32_t* __restrict__ l1_crd,
                                                    const int32_t* __restrict__ l2_pos,
                                                    const int32_t* __restrict__ l2_crd,
                                                    const int32_t* __restrict__ l3_pos,
                                                    const int32_t* __restrict__ l3_crd,
                                                    const double* __restrict__ vals,
                                                    const int32_t* __restrict__ root_base,
                                                    const double* __restrict__ U, int32_t* __restrict__ out_i,
                                                    int32_t* __restrict__ out_j, double* __restrict__ Y) {
  constexpr int G = kWave / RP;
  constexpr int NB = 16;  // steps whose U rows are loaded together
  __shared__ int32_t ko[kTwLight], sk[kTwLight], sj[kTwLight], sfib[kTwLight];
  __shared__ double xo[kTwLight], sx[kTwLight];
  __shared__ int32_t spos[kTwLight + 1];
  const int lane = lane_id();
  const int32_t r = blockIdx.x;
  const int32_t f0 = l2_pos[r], nf = l2_pos[r + 1] - f0;
  const int32_t lb = l3_pos[f0], le = l3_pos[f0 + nf];
  const int32_t n = le - lb;
  if (n == 0 || n > kTwLight) return;
  const bool staged = nf <= kTwLight;
  if (staged)
    for (int i = lane; i <= nf; i += kWave) spos[i] = l3_pos[f0 + i];
  __builtin_amdgcn_wave_barrier();
  int32_t v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int idx = u * kWave + lane;
    v[u] = INT_MAX;
    if (idx < n) {
      const int32_t e = lb + idx;
      int32_t lo = 0, hi = nf - 1;
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        const int32_t pv = staged ? spos[mid] : l3_pos[f0 + mid];
        if (pv <= e) lo = mid; else hi = mid - 1;
      }
      ko[idx] = l2_crd[f0 + lo];
      xo[idx] = vals[e];
      v[u] = (l3_crd[e] << 8) | idx;
    }
  }
  int n_pow2 = 2;
  while (n_pow2 < n) n_pow2 <<= 1;
  wave_bitonic<4>(v, n_pow2);
  __builtin_amdgcn_wave_barrier();
  // sorted position p: its k, X and column; run starts and fiber numbers
  int32_t before = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int p = u * kWave + lane;
    const int32_t j = v[u] >> 8;
    const int32_t up = __shfl_up(j, 1, kWave);
    const int32_t carry = u > 0 ? (__shfl(v[u > 0 ? u - 1 : 0], kWave - 1, kWave) >> 8) : INT_MIN;
    const int32_t prev = lane > 0 ? up : carry;
    const bool start = p < n && (p == 0 || j != prev);
    const uint64_t sm = __ballot(start);
    const uint64_t le_mask = (lane == kWave - 1) ? ~0ull : ((2ull << lane) - 1ull);
    if (p < n) {
      const int idx = v[u] & 0xff;
      sk[p] = ko[idx];
      sx[p] = xo[idx];
      sj[p] = j;
      sfib[p] = before + __popcll(sm & le_mask) - 1;
    }
    before += __popcll(sm);
  }
  __builtin_amdgcn_wave_barrier();
  const int32_t base = root_base[r], ii = l1_crd[r];
  const int g = lane / RP, rr = lane % RP;
  const bool rok = rr < R;
  const int rc = rok ? rr : 0;
  double prev = 0.0;  // running sum of the previous nonzero (same component)
  const int steps = (n + G - 1) / G;
  for (int s0 = 0; s0 < steps; s0 += NB) {
    double uv[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const int p = (s0 + t) * G + g;
      uv[t] = p < n ? U[(int64_t)sk[p] * R + rc] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      if (s0 + t < steps) {  // uniform
        const int p = (s0 + t) * G + g;
        const bool ok = p < n;
        const double v0 = ok ? sx[p] * uv[t] : 0.0;
        const bool st = ok && (p == 0 || sj[p] != sj[p - 1]);
        double a = 0.0;
#pragma unroll
        for (int h = 0; h < G; ++h) {
          if (g == h) a = (st ? 0.0 : prev) + v0;
          prev = __shfl(a, h * RP + rr, kWave);
        }
        const bool last = ok && (p == n - 1 || sj[p + 1] != sj[p]);
        if (last) {
          const int fi = base + sfib[p];
          if (rok) Y[(int64_t)fi * R + rr] = a;
          if (rr == 0) {
            out_i[fi] = ii;
            out_j[fi] = sj[p];
          }
        }
      }
    }
  }
}

}  // namespace

int ttm_roots_run(spws_ctx* ctx, const Csf3& x, int32_t I, int32_t J, int32_t K, int32_t R, const double* U,
                  bool forced, int64_t* nfib_out, bool* handled) {
  (void)I;
  *handled = false;
  if (R < 1 || R > 64 || x.nnz <= 0 || x.nroots <= 0 || J <= 0 || J > (1 << 23)) return SPWS_OK;
  const int RP = R <= 16 ? 16 : (R <= 32 ? 32 : 64);
  const int S = 1024 / RP;
  const int s = S == 64 ? 6 : (S == 32 ? 5 : 4);
  const int64_t W64 = ((int64_t)J + S - 1) / S;
  if (W64 > kTwMaxWin || ((int64_t)K << s) > (int64_t)UINT32_MAX || x.nnz >= ((int64_t)1 << 31)) return SPWS_OK;
  const int32_t W = (int32_t)W64;
  int wbits = 0;
  while (((int64_t)1 << wbits) < W) ++wbits;
  *handled = true;
  hipStream_t st = ctx->stream;
  int rc;
  const int32_t nroots = x.nroots;
  memset(&ctx->stats, 0, sizeof(ctx->stats));
  // per-root tiles and heavy ranks
  if ((rc = ctx->tw[0].ensure(((size_t)nroots * 2 + 1) * 8))) return rc;
  int64_t* pk = ctx->tw[0].as<int64_t>();
  int64_t* pko = pk + nroots;
  SPWS_LAUNCH(ctx, "tw_roots", k_tw_roots, dim3((unsigned)std::min<int64_t>(div_up(nroots, 256), 65536)), dim3(256),
              0, nroots, x.l2_pos, x.l3_pos, pk);
  if ((rc = exclusive_scan<int64_t, int64_t>(ctx, pk, pko, nroots))) return rc;
  int64_t* hm = ctx->host_meta.as<int64_t>() + 40;
  SPWS_HIP(hipMemcpyAsync(hm, pko + nroots, 8, hipMemcpyDeviceToHost, st));
  SPWS_HIP(hipStreamSynchronize(st));
  const int64_t NT = hm[0] >> 32, H = hm[0] & 0xffffffffll;
  const int64_t nwin = H * W;
  // chosen automatically only when the heavy roots' windows are dense
  // enough (at least kTwDense nonzeros per window on average, counting all
  // nonzeros): a window's fixed costs (its bounds, the column reduction, the
  // emission) are paid per window, and sparse windows of mid-size roots
  // make the whole-tensor sort cheaper
  if (!forced && H > 0 && (double)x.nnz < (double)kTwDense * (double)nwin) {
    *handled = false;
    return SPWS_OK;
  }
  const size_t hn = (size_t)(NT * W);
  if ((rc = ctx->tw[1].ensure((size_t)std::max<int64_t>(H, 1) * 4)) ||
      (rc = ctx->tw[2].ensure((2 * hn + 1) * 4)) || (rc = ctx->tw[3].ensure((size_t)x.nnz * 4)) ||
      (rc = ctx->tw[4].ensure((size_t)x.nnz * 8)) || (rc = ctx->tw[5].ensure(((size_t)nwin * 2 + 1) * 4)) ||
      (rc = ctx->tw[6].ensure(((size_t)nroots * 2 + 1) * 4)) || (rc = ctx->tw[7].ensure(((size_t)nwin + 1) * 4)))
    return rc;
  int32_t* hroot = ctx->tw[1].as<int32_t>();
  int32_t* hist = ctx->tw[2].as<int32_t>();
  int32_t* hoff = hist + hn;
  uint32_t* hkey = ctx->tw[3].as<uint32_t>();
  double* hval = ctx->tw[4].as<double>();
  int32_t* dcnt = ctx->tw[5].as<int32_t>();
  int32_t* doff = dcnt + nwin;
  int32_t* fcount = ctx->tw[6].as<int32_t>();
  int32_t* root_base = fcount + nroots;
  int32_t* wstart = ctx->tw[7].as<int32_t>();
  // heavy work items of about kTwItem nonzeros: a window belongs to item
  // floor(start / kTwItem) with start <= nnz, so nnz / kTwItem + 1 items
  // cover every window (the empty ones at the end included)
  const unsigned gi = (unsigned)(x.nnz / kTwItem + 1);
  if (NT > 0) {
    SPWS_LAUNCH(ctx, "tw_hist", k_tw_hist, dim3((unsigned)NT), dim3(kWave), 0, nroots, (const int64_t*)pko, x.l2_pos,
                x.l3_pos, x.l3_crd, s, W, hist, hroot);
    if ((rc = exclusive_scan<int32_t, int32_t>(ctx, hist, hoff, (int64_t)hn))) return rc;
    SPWS_LAUNCH(ctx, "tw_scatter", k_tw_scatter, dim3((unsigned)NT), dim3(kWave), 0, nroots, (const int64_t*)pko,
                x.l2_pos, x.l2_crd, x.l3_pos, x.l3_crd, x.vals, s, W, wbits, (const int32_t*)hoff, hkey, hval);
    SPWS_LAUNCH(ctx, "tw_wstart", k_tw_wstart, dim3((unsigned)std::min<int64_t>(div_up(nwin + 1, 256), 65536)),
                dim3(256), 0, nwin, W, (const int32_t*)hroot, (const int64_t*)pko, (const int32_t*)hoff,
                (int64_t)hn, wstart);
    SPWS_LAUNCH(ctx, "tw_wcount", k_tw_wcount, dim3(gi), dim3(kWave), 0, nwin, s, (const int32_t*)wstart,
                (const uint32_t*)hkey, dcnt);
    if ((rc = exclusive_scan<int32_t, int32_t>(ctx, dcnt, doff, nwin))) return rc;
  }
  SPWS_LAUNCH(ctx, "tw_light_count", k_tw_light_count, dim3((unsigned)nroots), dim3(kWave), 0, nroots, W,
              (const int64_t*)pko, (const int32_t*)doff, x.l2_pos, x.l3_pos, x.l3_crd, fcount);
  if ((rc = exclusive_scan<int32_t, int32_t>(ctx, fcount, root_base, nroots))) return rc;
  int32_t* hf = ctx->host_meta.as<int32_t>() + 100;
  SPWS_HIP(hipMemcpyAsync(hf, root_base + nroots, 4, hipMemcpyDeviceToHost, st));
  SPWS_HIP(hipStreamSynchronize(st));
  const int64_t F = *hf;
  if ((rc = ctx->out_crd.ensure((size_t)(F > 0 ? F : 1) * 4 * 2)) ||
      (rc = ctx->out_val.ensure((size_t)(F > 0 ? F : 1) * R * 8)))
    return rc;
  int32_t* out_i = ctx->out_crd.as<int32_t>();
  int32_t* out_j = out_i + F;
  double* Y = ctx->out_val.as<double>();
  if (NT > 0) {
#define SPWS_TW_HEAVY(RPV)                                                                                     \
  SPWS_LAUNCH(ctx, "tw_heavy", k_tw_heavy<RPV>, dim3(gi), dim3(kWave), 0, nwin, W, s, R, (const int32_t*)hroot, \
              (const int64_t*)pko, (const int32_t*)wstart, (const uint32_t*)hkey, (const double*)hval,         \
              (const int32_t*)doff, (const int32_t*)root_base, x.l1_crd, U, out_i, out_j, Y)
    if (RP == 16) SPWS_TW_HEAVY(16);
    else if (RP == 32) SPWS_TW_HEAVY(32);
    else SPWS_TW_HEAVY(64);
#undef SPWS_TW_HEAVY
  }
#define SPWS_TW_LIGHT(RPV)                                                                                        \
  SPWS_LAUNCH(ctx, "tw_light", k_tw_light<RPV>, dim3((unsigned)nroots), dim3(kWave), 0, nroots, R, x.l1_crd,       \
              x.l2_pos, x.l2_crd, x.l3_pos, x.l3_crd, x.vals, (const int32_t*)root_base, U, out_i, out_j, Y)
  if (RP == 16) SPWS_TW_LIGHT(16);
  else if (RP == 32) SPWS_TW_LIGHT(32);
  else SPWS_TW_LIGHT(64);
#undef SPWS_TW_LIGHT
  ctx->stats.products = x.nnz * (int64_t)R;
  *nfib_out = F;
  return SPWS_OK;
}

}  // namespace spws
