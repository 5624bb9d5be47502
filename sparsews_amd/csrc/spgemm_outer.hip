// Outer-product SpGEMM C = A*B with A in DCSC (or CSC) and B in DCSR (or
// CSR), through a coordinate workspace: expand -> radix sort -> segmented
// reduce, processed in chunks of output rows.
//
// Reference: benchmark/DCSC_DCSR/DCSC_DCSR_coord_index_chase.h:194-311 (append
// (i,k,v) per matched j, sort the workspace by (i,k), merge with reduction),
// the expand/sort/reduce shape of DCSC_DCSR_noacc_coord.h:4-121, and the
// co-iteration of A's columns with B's rows (TACO_MIN merge, :228-264).
//
// Device pipeline:
//   k_outer_match     j co-iteration as a parallel search: for each stored
//                     column of A, the matching stored row of B (or none).
//   k_outer_row_ub    products per output row: ub_i += |B(j,:)| over A(i,j).
//   -- host: prefix of ub, row chunks [r0, r1) bounded by products and by the
//      key width (32-bit keys while rows*2^kb fits).
//   per chunk:
//   k_outer_colrange  per column: rows of A(:,j) inside [r0, r1) (binary
//                     search), products m_j * |B(j,:)|, active flag.
//   scan + k_outer_compact   active columns -> dense list with product offsets.
//   k_outer_expand    one tile of products per workgroup; the owning column of
//                     each product by binary search in the tile's slice of the
//                     offsets (LDS); key = (i - r0) << kb | k, value = a*b.
//   radix sort        stable LSD on the kb + rb key bits.
//   k_outer_heads + scan     distinct keys per tile.
//   k_outer_reduce    each distinct key sums its run in sorted (= j) order,
//                     writes (k, value) and the row pointers it starts.
//   k_outer_tail      row pointers of trailing empty rows.
// Values of each entry are summed in increasing j, exactly Gustavson's
// left-to-right order.
//
// Second sort method (the default): an MSD radix sort whose first digit is
// (row i, column window k >> s).  The digit counts need no pass over the
// products: the products of A(i,j) that land in window w are the entries of
// B(j,:) in that window, a range found by binary search in the sorted row.
//   k_bk_bseg         per stored B row j and window w: first entry >= w*2^s.
//   k_bk_entries      the chunk's A entries, keyed by row, radix-sorted so
//                     each row lists its entries in increasing j.
//   k_bk_rowsize      bucket sizes, one wavefront per row, lanes over windows.
//   scan              bucket starts (the digit's prefix sum).
//   k_bk_etab         per entry and window: where its products start.
//   k_bk_expand_lds   each product goes straight to its bucket slot (the
//                     entry's window table in LDS; k_bk_expand reads it from
//                     memory when a row has more than 1024 windows); inside
//                     a bucket the products are in (j, k) order.
//   k_bk_count        distinct columns per bucket (LDS bitmap).
//   scan              output offsets of the buckets.
//   k_bk_reduce       one wavefront per bucket: products accumulate in a dense
//                     LDS window, entries with equal k in position (= j)
//                     order; the window's columns are emitted in order.
// The reduce of a chunk runs on the context's second stream while the main
// stream prepares the next chunk (entries, bucket sizes and starts, window
// table); the next chunk's expand waits for that reduce.
#include <algorithm>
#include <climits>

#include "radix.hpp"

namespace spws {

constexpr int kExpBlock = 256;
constexpr int kExpItems = 16;
constexpr int kExpTile = kExpBlock * kExpItems;

__device__ __forceinline__ int32_t lower_bound_i32(const int32_t* a, int32_t lo, int32_t hi, int32_t x) {
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// match[ja] = jb (stored row of B with the same index as A's stored column ja), or -1
__global__ __launch_bounds__(256) void k_outer_match(int32_t a_ncols, const int32_t* __restrict__ A1_crd,
                                                     int32_t b_nrows, const int32_t* __restrict__ B1_crd,
                                                     const int32_t* __restrict__ B2_pos, int32_t* __restrict__ match,
                                                     int32_t* __restrict__ blen) {
  const int32_t ja = blockIdx.x * blockDim.x + threadIdx.x;
  if (ja >= a_ncols) return;
  const int32_t j = A1_crd ? A1_crd[ja] : ja;
  int32_t jb = -1;
  if (B1_crd) {
    const int32_t p = lower_bound_i32(B1_crd, 0, b_nrows, j);
    if (p < b_nrows && B1_crd[p] == j) jb = p;
  } else if (j < b_nrows) {
    jb = j;
  }
  match[ja] = jb;
  blen[ja] = jb >= 0 ? B2_pos[jb + 1] - B2_pos[jb] : 0;
}

// One wavefront per stored column of A (columns range from 1 to ~2e4 rows).
__global__ __launch_bounds__(256) void k_outer_row_ub(int32_t a_ncols, const int32_t* __restrict__ A2_pos,
                                                      const int32_t* __restrict__ A2_crd,
                                                      const int32_t* __restrict__ blen,
                                                      unsigned long long* __restrict__ row_ub,
                                                      int32_t* __restrict__ row_pairs) {
  const int32_t nw = (int32_t)(gridDim.x * blockDim.x / kWave);
  for (int32_t ja = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave); ja < a_ncols; ja += nw) {
    const int32_t L = blen[ja];
    if (L == 0) continue;
    for (int32_t p = A2_pos[ja] + lane_id(); p < A2_pos[ja + 1]; p += kWave) {
      atomicAdd(&row_ub[A2_crd[p]], (unsigned long long)L);
      atomicAdd(&row_pairs[A2_crd[p]], 1);
    }
  }
}

__global__ __launch_bounds__(256) void k_outer_colrange(int32_t a_ncols, int32_t r0, int32_t r1,
                                                        const int32_t* __restrict__ A2_pos,
                                                        const int32_t* __restrict__ A2_crd,
                                                        const int32_t* __restrict__ blen,
                                                        int32_t* __restrict__ col_lo, int32_t* __restrict__ col_m,
                                                        int32_t* __restrict__ flag) {
  const int32_t ja = blockIdx.x * blockDim.x + threadIdx.x;
  if (ja >= a_ncols) return;
  int32_t m = 0, lo = 0;
  if (blen[ja] > 0) {
    const int32_t b = A2_pos[ja], e = A2_pos[ja + 1];
    lo = lower_bound_i32(A2_crd, b, e, r0);
    const int32_t hi = lower_bound_i32(A2_crd, lo, e, r1);
    m = hi - lo;
  }
  col_lo[ja] = lo;
  col_m[ja] = m;
  flag[ja] = m > 0 ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_outer_compact(int32_t a_ncols, const int32_t* __restrict__ flag_off,
                                                       const int32_t* __restrict__ col_lo,
                                                       const int32_t* __restrict__ col_m,
                                                       const int32_t* __restrict__ match,
                                                       const int32_t* __restrict__ blen,
                                                       const int32_t* __restrict__ B2_pos,
                                                       int32_t* __restrict__ act_astart, int32_t* __restrict__ act_L,
                                                       int32_t* __restrict__ act_bstart,
                                                       int64_t* __restrict__ act_cnt, int32_t* __restrict__ act_ja,
                                                       int32_t* __restrict__ act_m) {
  const int32_t ja = blockIdx.x * blockDim.x + threadIdx.x;
  if (ja >= a_ncols) return;
  const int32_t m = col_m[ja];
  if (m == 0) return;
  const int32_t a = flag_off[ja];
  const int32_t L = blen[ja];
  act_astart[a] = col_lo[ja];
  act_L[a] = L;
  act_bstart[a] = B2_pos[match[ja]];
  act_cnt[a] = (int64_t)m * L;
  if (act_ja) {
    act_ja[a] = ja;
    act_m[a] = m;
  }
}

template <class K>
__global__ __launch_bounds__(kExpBlock) void k_outer_expand(int64_t P, const int32_t* __restrict__ n_act_ptr,
                                                            const int64_t* __restrict__ act_off,
                                                            const int32_t* __restrict__ act_astart,
                                                            const int32_t* __restrict__ act_L,
                                                            const int32_t* __restrict__ act_bstart,
                                                            const int32_t* __restrict__ A2_crd,
                                                            const double* __restrict__ Ax,
                                                            const int32_t* __restrict__ B2_crd,
                                                            const double* __restrict__ Bx, int32_t r0, int kb,
                                                            K* __restrict__ keys, double* __restrict__ vals) {
  __shared__ int64_t soff[kExpTile + 1];
  __shared__ int32_t s_a0, s_span;
  const int64_t p0 = (int64_t)blockIdx.x * kExpTile;
  const int64_t p_end = p0 + kExpTile < P ? p0 + kExpTile : P;
  if (threadIdx.x == 0) {
    const int32_t n_act = *n_act_ptr;
    // last a with act_off[a] <= x, over a in [0, n_act)
    auto owner = [&](int64_t x) {
      int32_t lo = 0, hi = n_act - 1;
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (act_off[mid] <= x) lo = mid; else hi = mid - 1;
      }
      return lo;
    };
    const int32_t a0 = owner(p0);
    const int32_t a1 = owner(p_end - 1);
    s_a0 = a0;
    s_span = a1 - a0 + 1;
  }
  __syncthreads();
  const int32_t a0 = s_a0, span = s_span;
  for (int i = threadIdx.x; i <= span; i += kExpBlock) soff[i] = act_off[a0 + i];
  __syncthreads();
  int32_t lo_hint = 0;
#pragma unroll 4
  for (int it = 0; it < kExpItems; ++it) {
    const int64_t p = p0 + (int64_t)it * kExpBlock + threadIdx.x;
    if (p >= p_end) break;
    // last idx in [lo_hint, span) with soff[idx] <= p (p grows with it)
    int32_t lo = lo_hint, hi = span - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (soff[mid] <= p) lo = mid; else hi = mid - 1;
    }
    lo_hint = lo;
    const int32_t a = a0 + lo;
    const uint32_t local = (uint32_t)(p - soff[lo]);
    const uint32_t L = (uint32_t)act_L[a];
    const uint32_t ii = local / L;
    const uint32_t kk = local - ii * L;
    const int32_t ai = act_astart[a] + (int32_t)ii;
    const int32_t bi = act_bstart[a] + (int32_t)kk;
    const int32_t i = A2_crd[ai];
    const int32_t k = B2_crd[bi];
    keys[p] = ((K)(uint32_t)(i - r0) << kb) | (K)(uint32_t)k;
    vals[p] = Ax[ai] * Bx[bi];
  }
}

template <class K>
__global__ __launch_bounds__(kRadixBlock) void k_outer_heads(const K* __restrict__ keys, int64_t n,
                                                             int32_t* __restrict__ tile_heads) {
  __shared__ int32_t sh[kRadixBlock / kWave];
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
  int32_t c = 0;
#pragma unroll
  for (int it = 0; it < kRadixItems; ++it) {
    const int64_t p = base + (int64_t)it * kRadixBlock + threadIdx.x;
    if (p < n && (p == 0 || keys[p] != keys[p - 1])) ++c;
  }
  c = wave_sum(c);
  if (lane_id() == 0) sh[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t t = 0;
    for (int w = 0; w < kRadixBlock / kWave; ++w) t += sh[w];
    tile_heads[blockIdx.x] = t;
  }
}

// One tile of sorted keys per workgroup.  The tile is staged in LDS with
// coalesced loads; each thread scans kRadixItems consecutive keys for run
// heads, a block scan numbers the distinct keys, each head sums its run in
// order (continuing in global memory when the run leaves the tile), and the
// tile's distinct (column, value) pairs are staged in LDS and stored with
// coalesced writes.
template <class K>
__global__ __launch_bounds__(kRadixBlock) void k_outer_reduce(const K* __restrict__ keys, const double* __restrict__ vals,
                                                              int64_t n, const int32_t* __restrict__ tile_uoff, int kb,
                                                              int64_t base, int32_t r0, int32_t* __restrict__ out_crd,
                                                              double* __restrict__ out_val,
                                                              int64_t* __restrict__ out_pos) {
  __shared__ K sk[kRadixTile];
  __shared__ double sv[kRadixTile];
  __shared__ int32_t scrd[kRadixTile];
  __shared__ int32_t sh[kRadixBlock / kWave + 1];
  const K kmask = (((K)1) << kb) - 1;
  const int64_t tbase = (int64_t)blockIdx.x * kRadixTile;
  const int64_t remain = n - tbase;
  const int tile_n = remain < kRadixTile ? (int)remain : kRadixTile;
  for (int i = threadIdx.x; i < tile_n; i += kRadixBlock) {
    sk[i] = keys[tbase + i];
    sv[i] = vals[tbase + i];
  }
  const bool has_prev = tbase > 0;
  const K prev_key = has_prev ? keys[tbase - 1] : (K)0;
  __syncthreads();
  const int first = threadIdx.x * kRadixItems;
  uint32_t heads = 0;
  int32_t cnt = 0;
#pragma unroll
  for (int it = 0; it < kRadixItems; ++it) {
    const int i = first + it;
    if (i < tile_n) {
      const bool h = i == 0 ? !has_prev || sk[0] != prev_key : sk[i] != sk[i - 1];
      if (h) {
        heads |= 1u << it;
        ++cnt;
      }
    }
  }
  int32_t tot;
  const int32_t pre = block_excl_scan<int32_t, kRadixBlock>(cnt, sh, &tot);
  // run sums (registers), then stage
  double sums[kRadixItems];
#pragma unroll
  for (int it = 0; it < kRadixItems; ++it) {
    sums[it] = 0.0;
    if (!(heads & (1u << it))) continue;
    const int i = first + it;
    const K key = sk[i];
    double v = sv[i];
    int q = i + 1;
    while (q < tile_n && sk[q] == key) v += sv[q++];
    if (q == tile_n) {
      for (int64_t g = tbase + q; g < n && keys[g] == key; ++g) v += vals[g];
    }
    sums[it] = v;
  }
  const int64_t ubase = base + tile_uoff[blockIdx.x];
  __syncthreads();
  int32_t u = pre;
#pragma unroll
  for (int it = 0; it < kRadixItems; ++it) {
    if (!(heads & (1u << it))) continue;
    const int i = first + it;
    const K key = sk[i];
    scrd[u] = (int32_t)(uint32_t)(key & kmask);
    sv[u] = sums[it];
    const int32_t row = (int32_t)(key >> kb);
    const int32_t prow = i == 0 ? (has_prev ? (int32_t)(prev_key >> kb) : -1) : (int32_t)(sk[i - 1] >> kb);
    for (int32_t r = prow + 1; r <= row; ++r) out_pos[r0 + r] = ubase + u;
    ++u;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < tot; i += kRadixBlock) {
    out_crd[ubase + i] = scrd[i];
    out_val[ubase + i] = sv[i];
  }
}

template <class K>
__global__ __launch_bounds__(256) void k_outer_tail(const K* __restrict__ keys, int64_t n, int kb, int32_t r0,
                                                    int32_t R, int64_t base_end, int64_t* __restrict__ out_pos) {
  const int32_t last = n > 0 ? (int32_t)(keys[n - 1] >> kb) : -1;
  for (int32_t r = last + 1 + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); r < R;
       r += gridDim.x * blockDim.x)
    out_pos[r0 + r] = base_end;
}

__global__ void k_set_i64(int64_t* p, int64_t v) { *p = v; }

// ------------------------------------------------------------ bucket method

constexpr int kBkWaves = 2;   // wavefronts per workgroup in k_bk_count
constexpr int kBkMinLog2 = 6;
constexpr int kBkMaxLog2 = 11;

// bseg[ja*(W+1) + w]: first entry of the B row matched with A column ja
// (relative to the row start) whose column is >= w << s, for w in [0, W].
__global__ __launch_bounds__(256) void k_bk_bseg(int32_t a_ncols, int32_t W, int s, const int32_t* __restrict__ match,
                                                 const int32_t* __restrict__ blen, const int32_t* __restrict__ B2_pos,
                                                 const int32_t* __restrict__ B2_crd, int32_t* __restrict__ bseg) {
  const int64_t total = (int64_t)a_ncols * (W + 1);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t ja = (int32_t)(t / (W + 1));
    const int32_t w = (int32_t)(t - (int64_t)ja * (W + 1));
    const int32_t L = blen[ja];
    int32_t r = 0;
    if (L > 0) {
      const int32_t b = B2_pos[match[ja]];
      const int64_t x = (int64_t)w << s;
      r = x > INT32_MAX ? L : lower_bound_i32(B2_crd, b, b + L, (int32_t)x) - b;
    }
    bseg[t] = r;
  }
}

// The chunk's matched A entries, numbered e by active column then row:
// key = i - r0, value = e (bit pattern in a double), and e's A column.
__global__ __launch_bounds__(256) void k_bk_entries(int32_t En, const int32_t* __restrict__ n_act_ptr,
                                                    const int32_t* __restrict__ act_eoff,
                                                    const int32_t* __restrict__ act_astart,
                                                    const int32_t* __restrict__ act_ja,
                                                    const int32_t* __restrict__ A2_crd, int32_t r0,
                                                    uint32_t* __restrict__ pkey, double* __restrict__ pval,
                                                    int32_t* __restrict__ ent_ja, int32_t* __restrict__ ent_ai) {
  const int32_t n_act = *n_act_ptr;
  for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < En; e += gridDim.x * blockDim.x) {
    int32_t lo = 0, hi = n_act - 1;  // last active column c with act_eoff[c] <= e
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (act_eoff[mid] <= e) lo = mid; else hi = mid - 1;
    }
    const int32_t ai = act_astart[lo] + (e - act_eoff[lo]);
    pkey[e] = (uint32_t)(A2_crd[ai] - r0);
    pval[e] = __longlong_as_double((long long)e);
    ent_ja[e] = act_ja[lo];
    ent_ai[e] = ai;
  }
}

// One wavefront per chunk row, lanes over windows, entries in increasing j.
// kPass 0: bsize[r*W + w] = products of the row in window w.
// kPass 1: etab[e*W + w] = bucket start + products of the row's earlier
//          entries in w - bseg[w], so that product (e, B entry q) of window w
//          goes to etab[e*W + w] + q.
template <int kPass>
__global__ __launch_bounds__(256) void k_bk_rows(int32_t R, int32_t W, const int32_t* __restrict__ rp_off,
                                                 const double* __restrict__ pval, const int32_t* __restrict__ ent_ja,
                                                 const int32_t* __restrict__ bseg,
                                                 const int32_t* __restrict__ bstart, int32_t* __restrict__ out) {
  const int32_t nwaves = (int32_t)(gridDim.x * blockDim.x / kWave);
  const int lane = lane_id();
  for (int32_t r = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave); r < R; r += nwaves) {
    const int32_t q0 = rp_off[r], q1 = rp_off[r + 1];
    for (int32_t w0 = 0; w0 < W; w0 += kWave) {
      const int32_t w = w0 + lane;
      const bool on = w < W;
      int32_t run = 0;
      const int32_t start = (kPass == 1 && on) ? bstart[(int64_t)r * W + w] : 0;
      for (int32_t q = q0; q < q1; ++q) {
        const int32_t e = (int32_t)__double_as_longlong(pval[q]);
        const int32_t* bs = bseg + (int64_t)ent_ja[e] * (W + 1);
        if (on) {
          const int32_t lo = bs[w];
          if (kPass == 1) out[(int64_t)e * W + w] = start + run - lo;
          run += bs[w + 1] - lo;
        }
      }
      if (kPass == 0 && on) out[(int64_t)r * W + w] = run;
    }
  }
}

// Expand: one wavefront per A entry e = A(i,j) of the chunk, lanes over the
// matched row B(j,:).  The product with B entry kk goes to slot
// etab[e*W + (k >> s)] + kk of its (row, window) bucket; consecutive kk of
// one window land in consecutive slots, so the stores are runs.  The slot
// holds k's offset in the window and the product.  Entries are taken in row
// order (the sorted entry list), so the wavefronts in flight fill the
// adjacent runs of the same buckets and the stores stay within a few DRAM
// pages.
__global__ __launch_bounds__(256) void k_bk_expand(int32_t En, const double* __restrict__ order,
                                                   const int32_t* __restrict__ ent_ai,
                                                   const int32_t* __restrict__ ent_ja,
                                                   const int32_t* __restrict__ match,
                                                   const int32_t* __restrict__ blen,
                                                   const int32_t* __restrict__ B2_pos,
                                                   const double* __restrict__ Ax,
                                                   const int32_t* __restrict__ B2_crd,
                                                   const double* __restrict__ Bx, int32_t W, int s,
                                                   const int32_t* __restrict__ etab, uint16_t* __restrict__ klo,
                                                   double* __restrict__ pv) {
  constexpr int U = 8;
  const uint32_t wmask = (1u << s) - 1;
  const int lane = lane_id();
  const int32_t nw = (int32_t)(gridDim.x * blockDim.x / kWave);
  for (int32_t q = __builtin_amdgcn_readfirstlane((int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave)); q < En;
       q += nw) {
    // wave-uniform entry: its metadata comes through scalar loads
    const int32_t e = __builtin_amdgcn_readfirstlane((int32_t)__double_as_longlong(order[q]));
    const int32_t ja = __builtin_amdgcn_readfirstlane(ent_ja[e]);
    const int32_t L = __builtin_amdgcn_readfirstlane(blen[ja]);
    const int32_t bs = __builtin_amdgcn_readfirstlane(B2_pos[__builtin_amdgcn_readfirstlane(match[ja])]);
    const double av = Ax[__builtin_amdgcn_readfirstlane(ent_ai[e])];
    const int32_t* et = etab + (int64_t)e * W;
    // software pipeline: the next block of B is loaded before this block's
    // stores are issued, so waiting for those loads never waits for stores
    uint32_t kc[U];
    double bv[U];
    auto load = [&](int32_t k0, uint32_t (&kd)[U], double (&bd)[U]) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        // clamped index: every lane loads, lanes past the row are masked at the store
        const int32_t kk = min(k0 + u * kWave + lane, L - 1);
        kd[u] = (uint32_t)B2_crd[bs + kk];
        bd[u] = Bx[bs + kk];
      }
    };
    // two register sets, alternating blocks: no copies between them
    auto block = [&](int32_t k0, const uint32_t (&kd)[U], const double (&bd)[U], uint32_t (&kn)[U],
                     double (&bn)[U]) {
      int32_t dst[U];
#pragma unroll
      for (int u = 0; u < U; ++u) dst[u] = et[kd[u] >> s];
      if (k0 + kWave * U < L) load(k0 + kWave * U, kn, bn);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t kk = k0 + u * kWave + lane;
        if (kk < L) {
          const int32_t d = dst[u] + kk;
          klo[d] = (uint16_t)(kd[u] & wmask);
          pv[d] = av * bd[u];
        }
      }
    };
    uint32_t k2[U];
    double b2[U];
    load(0, kc, bv);
    for (int32_t k0 = 0; k0 < L; k0 += 2 * kWave * U) {
      block(k0, kc, bv, k2, b2);
      if (k0 + kWave * U >= L) break;
      block(k0 + kWave * U, k2, b2, kc, bv);
    }
  }
}

// Wait until every vector memory operation of the wavefront has completed.
// Loads and stores share one counter on this architecture, so once stores
// are in flight the compiler can only wait for a load by waiting for
// everything; a kernel that issues this just before a batch's loads makes
// every such wait one on operations issued a batch earlier.
__device__ __forceinline__ void vm_drain() { __builtin_amdgcn_s_waitcnt(0x0F70); }
// s_waitcnt lgkmcnt(0), vector memory and export counts left alone
__device__ __forceinline__ void lgkm_drain() { __builtin_amdgcn_s_waitcnt(0xC07F); }

constexpr int kBkExpandLdsW = 1024;  // widest window table k_bk_expand_lds keeps in LDS

// Expand with the entry's window table in LDS (W <= kBkExpandLdsW): one
// wavefront per A entry, as k_bk_expand.  The entry's W bucket slots are
// copied to LDS, so a product's slot is an LDS read of its window.  Each
// block of kWave * U products starts with a drain of the wavefront's memory
// operations (the block's B loads and the previous block's stores, both
// issued a block earlier), then issues the next block's loads, then looks up
// the slots and stores: no wait inside a block falls on memory issued
// within it.
template <int U>
__global__ __launch_bounds__(256) void k_bk_expand_lds(int32_t En, const double* __restrict__ order,
                                                       const int32_t* __restrict__ ent_ai,
                                                       const int32_t* __restrict__ ent_ja,
                                                       const int32_t* __restrict__ match,
                                                       const int32_t* __restrict__ blen,
                                                       const int32_t* __restrict__ B2_pos,
                                                       const double* __restrict__ Ax,
                                                       const int32_t* __restrict__ B2_crd,
                                                       const double* __restrict__ Bx, int32_t W, int s,
                                                       const int32_t* __restrict__ etab, uint16_t* __restrict__ klo,
                                                       double* __restrict__ pv) {
  extern __shared__ int32_t ex_lds[];
  int32_t* wt = ex_lds + (threadIdx.x / kWave) * W;
  const uint32_t wmask = (1u << s) - 1;
  const int lane = lane_id();
  const int32_t nw = (int32_t)(gridDim.x * blockDim.x / kWave);
  for (int32_t q = __builtin_amdgcn_readfirstlane((int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave)); q < En;
       q += nw) {
    const int32_t e = __builtin_amdgcn_readfirstlane((int32_t)__double_as_longlong(order[q]));
    const int32_t ja = __builtin_amdgcn_readfirstlane(ent_ja[e]);
    const int32_t L = __builtin_amdgcn_readfirstlane(blen[ja]);
    const int32_t bs = __builtin_amdgcn_readfirstlane(B2_pos[__builtin_amdgcn_readfirstlane(match[ja])]);
    const double av = Ax[__builtin_amdgcn_readfirstlane(ent_ai[e])];
    const int32_t* et = etab + (int64_t)e * W;
    __builtin_amdgcn_wave_barrier();
    for (int32_t w = lane; w < W; w += kWave) wt[w] = et[w];
    __builtin_amdgcn_wave_barrier();
    uint32_t k0[U], k1[U];
    double b0[U], b1[U];
    auto load = [&](int32_t kb, uint32_t (&kd)[U], double (&bd)[U]) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        // clamped index: every lane loads, lanes past the row are masked at the store
        const int32_t kk = min(kb + u * kWave + lane, L - 1);
        kd[u] = (uint32_t)B2_crd[bs + kk];
        bd[u] = Bx[bs + kk];
      }
    };
    auto block = [&](int32_t kb, const uint32_t (&kd)[U], const double (&bd)[U]) {
      int32_t dst[U];
#pragma unroll
      for (int u = 0; u < U; ++u) dst[u] = wt[kd[u] >> s];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t kk = kb + u * kWave + lane;
        if (kk < L) {
          const int32_t d = dst[u] + kk;
          klo[d] = (uint16_t)(kd[u] & wmask);
          pv[d] = av * bd[u];
        }
      }
    };
    constexpr int32_t span = kWave * U;
    load(0, k0, b0);
    for (int32_t kb = 0; kb < L; kb += 2 * span) {
      vm_drain();
      if (kb + span < L) load(kb + span, k1, b1);
      block(kb, k0, b0);
      if (kb + span >= L) break;
      vm_drain();
      if (kb + 2 * span < L) load(kb + 2 * span, k0, b0);
      block(kb + span, k1, b1);
    }
  }
}

// A batch of a bucket-walking wavefront: up to kSpan positions of one bucket.
struct BkBatch {
  int64_t b;         // bucket (-1: no more)
  int32_t p, pend;   // first position of the batch, end of the bucket
  int32_t p0;        // first position of the bucket
  int32_t off;       // bucket's output offset (when the walker loads them)
  bool last;         // last batch of its bucket
};

// Walks the buckets b0, b0 + step, b0 + 2 step, ... of one wavefront, in
// batches of kSpan positions, skipping empty buckets.  The bounds (and
// output offsets) of 64 buckets are loaded at once, one bucket per lane, and
// a group ahead of use, so that moving to the next bucket is register work:
// a ballot of the non-empty lanes and readlanes.  Empty buckets of a group
// get their count (zero) written by the lane that loaded them.  With kAlign
// a bucket's first batch starts at its first position rounded down to a
// multiple of 4 (for 4-element loads; p0 keeps the true start).
template <int kSpan, bool kOff, bool kAlign = false>
struct BkWalk {
  int64_t gb, step, nb;          // first bucket of the current group, stride, bucket count
  int32_t cs, ce, co, ns, ne, no;  // per lane: current / next group's bucket bounds and offset
  uint64_t mask;                 // current group's non-empty buckets not yet started
  int32_t* zero_cnt;             // when set, zero written here for empty buckets

  __device__ __forceinline__ void dload(int64_t g, const int32_t* __restrict__ bstart, const int32_t* __restrict__ boff,
                                        int32_t& ds, int32_t& de, int32_t& dofs) {
    const int64_t bi = g + (int64_t)lane_id() * step;
    ds = de = dofs = 0;
    if (bi < nb) {
      ds = bstart[bi];
      de = bstart[bi + 1];
      if (kOff) dofs = boff[bi];
    }
  }
  __device__ __forceinline__ void settle() {
    mask = __ballot(ce > cs);
    if (zero_cnt) {
      const int64_t bi = gb + (int64_t)lane_id() * step;
      if (bi < nb && ce <= cs) zero_cnt[bi] = 0;
    }
  }
  __device__ __forceinline__ void init(int64_t b0, int64_t step_, int64_t nb_, const int32_t* __restrict__ bstart,
                                       const int32_t* __restrict__ boff, int32_t* zc) {
    gb = b0;
    step = step_;
    nb = nb_;
    zero_cnt = zc;
    dload(gb, bstart, boff, cs, ce, co);
    dload(gb + step * kWave, bstart, boff, ns, ne, no);
    settle();
  }
  __device__ __forceinline__ BkBatch next(const BkBatch& c, const int32_t* __restrict__ bstart,
                                          const int32_t* __restrict__ boff) {
    BkBatch n = c;
    if (c.b >= 0 && c.p + kSpan < c.pend) {
      n.p = c.p + kSpan;
      n.last = n.p + kSpan >= n.pend;
      return n;
    }
    while (mask == 0) {
      gb += step * kWave;
      if (gb >= nb) {
        n.b = -1;
        return n;
      }
      cs = ns;
      ce = ne;
      co = no;
      dload(gb + step * kWave, bstart, boff, ns, ne, no);
      settle();
    }
    const int i = __builtin_ctzll(mask);
    mask &= mask - 1;
    n.b = gb + (int64_t)i * step;
    n.p0 = __builtin_amdgcn_readlane(cs, i);
    n.p = kAlign ? (n.p0 & ~3) : n.p0;
    n.pend = __builtin_amdgcn_readlane(ce, i);
    n.off = kOff ? __builtin_amdgcn_readlane(co, i) : 0;
    n.last = n.p + kSpan >= n.pend;
    return n;
  }
};

// Distinct columns per bucket: one wavefront per bucket, LDS bitmap of the
// window.  Each lane loads 8 bytes (four column offsets) per stride, from a
// 4-aligned position, kBkCountUnroll strides at a time and one batch ahead
// (two register sets) across bucket boundaries; offsets outside the bucket
// are masked.  The bitmap is counted and cleared at the end of each bucket.
constexpr int kBkCountUnroll = 4;

__global__ __launch_bounds__(kBkWaves * kWave) void k_bk_count(int32_t nb, int s, const int32_t* __restrict__ bstart,
                                                               const uint16_t* __restrict__ klo,
                                                               int32_t* __restrict__ bnnz) {
  constexpr int U = kBkCountUnroll;
  constexpr int kSpan = 4 * kWave * U;
  extern __shared__ uint32_t bk_bits[];
  const int words = (1 << s) >> 5;  // <= 64: one per lane
  uint32_t* bm = bk_bits + (threadIdx.x / kWave) * words;
  const int lane = lane_id();
  if (lane < words) bm[lane] = 0u;
  __builtin_amdgcn_wave_barrier();
  const uint2* kq = reinterpret_cast<const uint2*>(klo);
  auto load = [&](const BkBatch& c, uint2 (&kk)[U]) {
    // clamped to the bucket's last 4-element group: every lane loads in bounds
    const int32_t qlast = (c.pend - 1) >> 2;
#pragma unroll
    for (int u = 0; u < U; ++u) kk[u] = kq[min((c.p >> 2) + u * kWave + lane, qlast)];
  };
  auto apply = [&](const BkBatch& c, const uint2 (&kk)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t e0 = c.p + 4 * (u * kWave + lane);
      const uint32_t v[4] = {kk[u].x & 0xffffu, kk[u].x >> 16, kk[u].y & 0xffffu, kk[u].y >> 16};
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (e0 + i >= c.p0 && e0 + i < c.pend) atomicOr(&bm[v[i] >> 5], 1u << (v[i] & 31));
    }
    if (!c.last) return;
    __builtin_amdgcn_wave_barrier();
    uint32_t w = 0;
    if (lane < words) {
      w = bm[lane];
      bm[lane] = 0u;
    }
    const int cnt = wave_sum((int)__popc(w));
    if (lane == 0) bnnz[c.b] = cnt;
    __builtin_amdgcn_wave_barrier();
  };
  BkWalk<kSpan, false, true> wk;
  wk.init(__builtin_amdgcn_readfirstlane((int32_t)(blockIdx.x * kBkWaves + threadIdx.x / kWave)),
          (int64_t)gridDim.x * kBkWaves, nb, bstart, nullptr, bnnz);
  uint2 k0[U], k1[U];
  BkBatch c0;
  c0.b = -1;
  c0 = wk.next(c0, bstart, nullptr);
  if (c0.b >= 0) load(c0, k0);
  while (c0.b >= 0) {
    vm_drain();
    BkBatch c1 = wk.next(c0, bstart, nullptr);
    if (c1.b >= 0) load(c1, k1);
    apply(c0, k0);
    if (c1.b < 0) break;
    vm_drain();
    c0 = wk.next(c1, bstart, nullptr);
    if (c0.b >= 0) load(c0, k0);
    apply(c1, k1);
  }
}

// Row pointers: row r0 + r starts at its first bucket's output offset.
__global__ __launch_bounds__(256) void k_bk_rowpos(int32_t R, int32_t W, int32_t r0, int64_t base,
                                                   const int32_t* __restrict__ boff, int64_t* __restrict__ out_pos) {
  for (int32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < R; r += gridDim.x * blockDim.x)
    out_pos[r0 + r] = base + boff[(int64_t)r * W];
}

// LDS of one wavefront in k_bk_reduce, for a window of 2^s columns (s <= 11):
// acc (2^s doubles), the seen and stride bitmaps (2^s bits each), the column
// list (2^s uint16).
__host__ __device__ constexpr size_t bk_reduce_lds_per_wave(int s) {
  return ((size_t)8 << s) + 2 * (((size_t)1 << s) >> 3) + ((size_t)2 << s);
}

constexpr int kBkRedWaves = 1;   // wavefronts per workgroup in k_bk_reduce
constexpr int kBkRedUnroll = 16; // strides per batch of k_bk_reduce (two batches in flight)

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Reduce: one wavefront per bucket.  The window is a dense LDS accumulator;
// `seen` marks the columns written so far.  Products are taken in position
// order, 64 at a time (a stride).  Each lane sets its column's bit in a
// per-stride bitmap (returning the old word) while reading its accumulator
// slot; a lane that finds the bit already set shares its column with another
// lane of the stride.  Every lane starts its sum from the accumulator, which
// holds -0.0 for a column not yet seen: -0.0 + v is v, bit for bit, for
// every v but a NaN, so a sum starts from its first product exactly.  For
// each group of lanes sharing a column, the lowest lane (the earliest
// position) then adds the other members' products in lane order, taken from
// registers, and alone writes the slot.  Each entry is thus summed in
// position (= increasing j) order, starting from its first product, with one
// LDS round trip per stride.
// At the end of a bucket each lane lists the set columns of its seen word in
// LDS, at its prefix among the words, and the list is stored 64 entries at
// a time with every lane active (four accumulator reads in flight), so the
// stores are full and consecutive.  Loads run one batch ahead (two register
// sets), across bucket boundaries.  One wavefront per workgroup: the LDS of
// a 2^11-column window (about 21 KB) then allows seven per CU.
template <int kU>
__global__ __launch_bounds__(kBkRedWaves * kWave) void k_bk_reduce(int32_t nb, int32_t W, int s,
                                                                   const int32_t* __restrict__ bstart,
                                                                   const uint16_t* __restrict__ klo,
                                                                   const double* __restrict__ pv,
                                                                   const int32_t* __restrict__ boff, int64_t base,
                                                                   int32_t* __restrict__ out_crd,
                                                                   double* __restrict__ out_val) {
  extern __shared__ double bk_lds[];
  const int S = 1 << s, words = S >> 5;  // words <= 64: one per lane
  const int lane = lane_id();
  char* mine = reinterpret_cast<char*>(bk_lds) + (threadIdx.x / kWave) * bk_reduce_lds_per_wave(s);
  double* acc = reinterpret_cast<double*>(mine);
  uint32_t* seen = reinterpret_cast<uint32_t*>(acc + S);
  uint32_t* sb = seen + words;
  uint16_t* list = reinterpret_cast<uint16_t*>(sb + words);
  if (lane < words) {
    sb[lane] = 0u;
    seen[lane] = 0u;
  }
  // an untouched slot holds -0.0 (-0.0 + v is v for every v but a NaN), so
  // the first product of a column is added like the others
  for (int q = lane; q < S; q += kWave) acc[q] = -0.0;
  __builtin_amdgcn_wave_barrier();
  auto load = [&](const BkBatch& c, uint32_t (&kk)[kU], double (&vv)[kU]) {
    // clamped index: every lane loads; lanes past the bucket are masked by `ok` in apply
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int32_t q = min(c.p + u * kWave + lane, c.pend - 1);
      kk[u] = (uint32_t)klo[q];
      vv[u] = pv[q];
    }
  };
  auto apply = [&](const BkBatch& c, const uint32_t (&kk)[kU], const double (&vv)[kU]) {
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (c.p + u * kWave >= c.pend) break;
      const bool ok = c.p + u * kWave + lane < c.pend;
      const uint32_t k = kk[u];
      const double v = vv[u];
      const uint32_t wi = k >> 5, bit = 1u << (k & 31);
      uint32_t old = 0;
      double a = 0.0;
      if (ok) {
        old = atomicOr(&sb[wi], bit);
        a = acc[k];
      }
      double sum = a + v;
      bool lead = ok;
      uint64_t todo = __ballot(ok && (old & bit) != 0);
      while (todo) {
        const int l = __builtin_ctzll(todo);
        const uint32_t kl = (uint32_t)__builtin_amdgcn_readlane((int)k, l);
        const uint64_t grp = __ballot(ok && k == kl);
        todo &= ~grp;
        const int ld = __builtin_ctzll(grp);
        for (uint64_t mm = grp & (grp - 1); mm; mm &= mm - 1) {
          const double vl = readlane_f64(v, __builtin_ctzll(mm));
          if (lane == ld) sum = sum + vl;
        }
        if (lane != ld && ((grp >> lane) & 1)) lead = false;
      }
      if (lead) {
        acc[k] = sum;
        atomicOr(&seen[wi], bit);
      }
      if (ok) atomicAnd(&sb[wi], ~bit);
      __builtin_amdgcn_wave_barrier();
    }
    if (!c.last) return;
    // write the set columns in order
    const uint32_t myw = lane < words ? seen[lane] : 0u;
    __builtin_amdgcn_wave_barrier();
    if (lane < words) seen[lane] = 0u;
    const int64_t ob = base + c.off;
    const int32_t wb = (int32_t)(c.b % W) << s;
    int32_t* oc = out_crd + ob;
    double* ov = out_val + ob;
    // list every set column in order (each lane its word), then store 64 at a time
    const int cnt = __popc(myw);
    const int incl = wave_incl_scan(cnt);
    int r = incl - cnt;
    for (uint32_t bits = myw; bits; bits &= bits - 1) list[r++] = (uint16_t)(lane * 32 + __builtin_ctz(bits));
    const int total = __builtin_amdgcn_readlane(incl, kWave - 1);
    __builtin_amdgcn_wave_barrier();
    for (int r0 = 0; r0 < total; r0 += 4 * kWave) {
      uint32_t kq[4];
      double aq[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) kq[i] = list[min(r0 + i * kWave + lane, total - 1)];
#pragma unroll
      for (int i = 0; i < 4; ++i) aq[i] = acc[kq[i]];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int q = r0 + i * kWave + lane;
        if (q < total) {
          oc[q] = wb + (int32_t)kq[i];
          ov[q] = aq[i];
          acc[kq[i]] = -0.0;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  };

  BkWalk<kWave * kU, true> wk;
  wk.init(__builtin_amdgcn_readfirstlane((int32_t)(blockIdx.x * kBkRedWaves + threadIdx.x / kWave)), (int64_t)gridDim.x * kBkRedWaves,
          nb, bstart, boff, nullptr);
  uint32_t k0[kU], k1[kU];
  double v0[kU], v1[kU];
  BkBatch c0;
  c0.b = -1;
  c0 = wk.next(c0, bstart, boff);
  if (c0.b >= 0) load(c0, k0, v0);
  while (c0.b >= 0) {
    // the drain falls on this batch's loads and the last bucket's stores,
    // both issued a batch earlier, not on the loads issued next
    vm_drain();
    BkBatch c1 = wk.next(c0, bstart, boff);
    if (c1.b >= 0) load(c1, k1, v1);
    apply(c0, k0, v0);
    if (c1.b < 0) break;
    vm_drain();
    c0 = wk.next(c1, bstart, boff);
    if (c0.b >= 0) load(c0, k0, v0);
    apply(c1, k1, v1);
  }
}

// ------------------------------------------------------------ fused method
//
// Expand and reduce in one walk, with no product workspace: one wavefront
// per output row (rows handed out by a work counter), the row's column
// windows of 2^s columns in increasing order.  The products of row i in
// window w are the segments B(j, [w 2^s, (w+1) 2^s)) of its entries A(i, j),
// taken in increasing j (bseg gives each segment); concatenated they are the
// products of bucket (i, w) in the bucket method's order, gathered straight
// from B (whose rows stay in the caches) instead of from an expanded copy.
// Each stride of 64 products finds every lane's entry by a binary search
// over the lanes' segment offsets (wavefront shuffles).
//   symbolic pass: distinct columns per window in an LDS bitmap, summed per
//                  row (-> row pointers);
//   numeric pass:  the bucket reduce's LDS window accumulation (entries with
//                  equal k summed in position = increasing-j order, from
//                  -0.0), and the window's columns emitted in order at the
//                  row's running output offset.
// The loads of the next batch of strides (across window boundaries) are
// issued before the current batch is applied, and each lane's next segment
// end is loaded a window ahead.  Rows with more than 64 entries take their
// entries 64 at a time within each window, in order.

// One record per entry of a row (entries grouped by row, increasing j):
// where the matched B row starts, the A column (for bseg), the A value.
__global__ __launch_bounds__(256) void k_fz_records(int32_t En, const double* __restrict__ order,
                                                    const int32_t* __restrict__ ent_ja,
                                                    const int32_t* __restrict__ ent_ai,
                                                    const int32_t* __restrict__ match,
                                                    const int32_t* __restrict__ B2_pos, const double* __restrict__ Ax,
                                                    int32_t* __restrict__ rec_b, int32_t* __restrict__ rec_ja,
                                                    double* __restrict__ rec_a) {
  for (int32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < En; q += gridDim.x * blockDim.x) {
    const int32_t e = (int32_t)__double_as_longlong(order[q]);
    const int32_t ja = ent_ja[e];
    rec_b[q] = B2_pos[match[ja]];
    rec_ja[q] = ja;
    rec_a[q] = Ax[ent_ai[e]];
  }
}

constexpr int kFzU = 16;  // strides of 64 products per batch (numeric pass)
constexpr int kFzUSym = 8;  // the symbolic pass's (fewer registers, more wavefronts)
constexpr int kFzOwn = kFzU * kWave;  // owner marks of a batch (u16 each)
constexpr int kFzG = 4;  // numeric pass: strides whose owner searches overlap
static_assert(kFzU % kFzG == 0, "whole groups of strides per batch");
constexpr uint32_t kFzTagFirst = 32;  // (32 << 6) exceeds every window column (< 2^11)
constexpr uint32_t kFzTagEnd = 1024;  // tags stay below 2^16 with the lane in the low 6 bits
// integer compare predicates of __builtin_amdgcn_sicmp (LLVM's ICMP_SGT, SGE, SLE)
constexpr int kIcmpSgt = 38, kIcmpSge = 39, kIcmpSle = 41;

// Entries of the u16 area shared by the emitted column list and the owner
// marks of a batch.  A window of 2^11 columns can hold more entries: its
// emit then lists and writes lanes 0-31's columns, then lanes 32-63's (at
// most 32 x 32 each).
constexpr int kFzList = 32 * 32;
static_assert(kFzList >= kFzOwn, "a batch's owner marks fit the list area");

// LDS of one k_fz_rows wavefront: numeric: the window's values, the seen
// bitmap, (ordered) the stride bitmap, the list / marks, one 16-byte record
// per lane (B base, A value); symbolic: the seen bitmap, the marks, the B
// bases.  The window of 2^11 columns takes 19,712 bytes (unordered), so
// that eight wavefronts share a CU's 160 KiB.
__host__ __device__ constexpr size_t fz_lds_bytes(int s, int d, bool numeric, bool ordered) {
  return (numeric ? ((size_t)8 << s) + (ordered ? 2 : 1) * (((size_t)1 << s) >> 3) : (((size_t)1 << (s + d)) >> 3)) +
         2 * (size_t)kFzList + (numeric ? 16 : 4) * kWave;
}


// A batch: up to kFzU strides of one (window, entry chunk) of the row.
struct FzBatch {
  int32_t w;      // window (-1: the row is done)
  int32_t c;      // first entry of the chunk (relative to the row)
  int32_t q0;     // first product of the batch within the (window, chunk)
  int32_t total;  // products of the (window, chunk)
  bool win_end;   // last batch of the window
};

// B's column and value at entry idx.  kOff32 (B has fewer than 2^29
// entries): the byte offsets fit in 32 unsigned bits, and the loads take
// the array base from scalar registers with a 32-bit offset per lane
// instead of a sign-extended 64-bit address per lane.
template <bool kOff32>
__device__ __forceinline__ uint32_t fz_col(const int32_t* __restrict__ B2_crd, int32_t idx) {
  if constexpr (kOff32)
    return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(B2_crd) + ((uint32_t)idx << 2));
  else
    return (uint32_t)B2_crd[idx];
}
template <bool kOff32>
__device__ __forceinline__ double fz_val(const double* __restrict__ Bx, int32_t idx) {
  if constexpr (kOff32)
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(Bx) + ((uint32_t)idx << 3));
  else
    return Bx[idx];
}

// kOrdered (numeric pass): entries summed in position order, bit-identical
// to the other methods; otherwise every product goes into the window by an
// LDS atomic add (ds_add_f64), with no conflict resolution -- the sums are
// then reassociated within the 1e-12 bar and not reproducible bit for bit.
template <bool kNumeric, bool kOrdered = true, bool kOff32 = false>
__global__ __launch_bounds__(kWave) void k_fz_rows(int32_t m, int32_t W, int s, int d,
                                                   const int32_t* __restrict__ rp_off,
                                                   const int32_t* __restrict__ rec_b,
                                                   const int32_t* __restrict__ rec_ja,
                                                   const double* __restrict__ rec_a,
                                                   const int32_t* __restrict__ bseg,
                                                   const int32_t* __restrict__ B2_crd,
                                                   const double* __restrict__ Bx, int32_t* __restrict__ row_cnt,
                                                   const int64_t* __restrict__ C_pos, int32_t* __restrict__ C_crd,
                                                   double* __restrict__ C_val, int32_t* __restrict__ work) {
  extern __shared__ double fz_lds[];
  // strides per batch, and the owner marks of a batch
  constexpr int kU = kNumeric ? kFzU : kFzUSym;
  constexpr int kOwn = kU * kWave;
  static_assert(kOwn <= kFzList && kU % kFzG == 0, "a batch's marks fit the list area");
  // the walk's windows: 2^d segment-table windows each (the symbolic pass
  // only counts, so it takes wide windows: fewer per-window steps); the
  // numeric pass keeps d = 0 (its emit holds one bitmap word per lane)
  const int S = 1 << (s + d), words = S >> 5;
  const int32_t Wd = (W + (1 << d) - 1) >> d;
  const int lane = lane_id();
  double* acc = fz_lds;                                          // numeric: 2^s values
  uint32_t* seen = kNumeric ? reinterpret_cast<uint32_t*>(acc + S) : reinterpret_cast<uint32_t*>(fz_lds);
  constexpr bool kStrideMap = kNumeric && kOrdered;
  uint32_t* sb = seen + words;                                   // ordered numeric: this stride's columns
  // numeric: the emitted columns of a window; both passes: a batch's owner
  // marks (the two uses never overlap in time)
  uint16_t* list = reinterpret_cast<uint16_t*>(kStrideMap ? sb + words : seen + words);
  uint16_t* own = list;
  // per lane: B index - product index, and (numeric) the A value: one
  // 16-byte record per lane {base, unused, A value}, so that one address
  // serves both reads of an owner's record
  constexpr int kRec = kNumeric ? 4 : 1;  // record length in int32
  int32_t* s_base = reinterpret_cast<int32_t*>(list + kFzList);
  double* s_av = reinterpret_cast<double*>(s_base) + 1;  // numeric: A value of lane t at s_av[2 t]
  for (int q = lane; q < words; q += kWave) {
    seen[q] = 0u;
    if (kStrideMap) sb[q] = 0u;
  }
  if (kNumeric)
    for (int q = lane; q < S; q += kWave) acc[q] = -0.0;
  for (int q = lane; q < kOwn; q += kWave) own[q] = 0;
  __builtin_amdgcn_wave_barrier();
  // owner marks carry a batch tag (tagc << 6 | lane) above every column
  // value the list holds (< 2^11 = 32 << 6), so marks of earlier batches
  // and list entries never win the max-scan against this batch's
  uint32_t tagc = kFzTagFirst;

  for (;;) {
    int32_t r = 0;
    if (lane == 0) r = atomicAdd(work, 1);
    r = __builtin_amdgcn_readfirstlane(__shfl(r, 0, kWave));
    if (r >= m) break;
    const int32_t q_beg = __builtin_amdgcn_readfirstlane(rp_off[r]);
    const int32_t E = __builtin_amdgcn_readfirstlane(rp_off[r + 1]) - q_beg;
    const int64_t out0 = kNumeric ? C_pos[r] : 0;
    // the symbolic pass's count bounds the numeric pass's writes (the two
    // walk the same segments, so a window never passes it)
    const int64_t out_end = kNumeric ? C_pos[r + 1] : 0;
    int64_t run = 0;  // entries emitted (numeric) or counted (symbolic) so far
    if (E == 0) {
      if (!kNumeric && lane == 0) row_cnt[r] = 0;
      continue;
    }
    const bool one_chunk = E <= kWave;
    // per lane: its entry of the current chunk
    int32_t b_start = 0;
    double a_val = 0.0;
    const int32_t* bs_row = bseg;
    int32_t seg_lo = 0, seg_hi = 0, nxt_hi = 0;
    // Straight-line loads and selects (lanes past the chunk read entry 0 of
    // the row and keep empty segments): conditional writes of this state
    // compile to stores through a selected address in private memory.
    // Past the row's last window the clamped reads give the end of the
    // segment table's row, so nxt_hi then equals seg_hi.
    auto load_chunk = [&](int32_t c, int32_t w) {
      const int32_t q = c + lane;
      const bool in = q < E;
      const int32_t e = q_beg + (in ? q : 0);
      const int32_t rb = rec_b[e];
      const double ra = kNumeric ? rec_a[e] : 0.0;
      const int32_t* row = bseg + (int64_t)rec_ja[e] * (W + 1);
      const int32_t lo = row[min(w << d, W)];
      const int32_t hi = row[min((w + 1) << d, W)];
      const int32_t nh = row[min((w + (one_chunk ? 2 : 1)) << d, W)];
      b_start = in ? rb : 0;
      a_val = in ? ra : 0.0;
      bs_row = in ? row : bseg;
      seg_lo = in ? lo : 0;
      seg_hi = in ? hi : 0;
      nxt_hi = in ? nh : 0;
    };
    // the segment scan of the current (window, chunk): base + q = B index of
    // product q for the lane owning it
    int32_t excl = 0, base = 0;
    auto scan = [&](FzBatch& b) {
      const int32_t len = seg_hi - seg_lo;
      const int32_t incl = wave_incl_scan(len);
      excl = incl - len;
      base = b_start + seg_lo - excl;
      b.total = __builtin_amdgcn_readlane(incl, kWave - 1);
      s_base[kRec * lane] = base;
      if (kNumeric) s_av[2 * lane] = a_val;
    };
    // the batch after b (advancing chunks and windows; empty ones skipped)
    auto next = [&](const FzBatch& b) -> FzBatch {
      FzBatch n = b;
      if (b.q0 + kWave * kU < b.total) {
        n.q0 = b.q0 + kWave * kU;
        n.win_end = (n.q0 + kWave * kU >= n.total) && (n.c + kWave >= E);
        return n;
      }
      for (;;) {
        if (n.c + kWave < E) {
          n.c += kWave;  // next chunk of the same window
          load_chunk(n.c, n.w);
        } else {
          n.w += 1;
          n.c = 0;
          if (n.w >= Wd) {
            n.w = -1;
            return n;
          }
          if (one_chunk) {
            seg_lo = seg_hi;
            seg_hi = nxt_hi;
            // the next window's end, a window ahead (lanes past the row's
            // entries read the table's first row and stay empty)
            const int32_t nh = bs_row[min((n.w + 2) << d, W)];
            nxt_hi = lane < E ? nh : 0;
          } else {
            load_chunk(0, n.w);
          }
        }
        n.q0 = 0;
        scan(n);
        n.win_end = (kWave * kU >= n.total) && (n.c + kWave >= E);
        // an empty (window, chunk) still closes its window when it is the last chunk
        if (n.total > 0 || n.c + kWave >= E) return n;
      }
    };
    // loads of one batch: column and value (numeric) of each lane's product
    uint32_t kc[2][kU];
    double bv[2][kU], av[2][kU];
    auto load = [&](const FzBatch& b, int slot) {
      const int32_t len = seg_hi - seg_lo;
      // this batch's owner marks: each segment that starts inside the
      // batch marks its first product with its lane
      if (tagc >= kFzTagEnd) {
        for (int q = lane; q < kOwn; q += kWave) own[q] = 0;
        tagc = kFzTagFirst;
      }
      const uint32_t tag = tagc << 6;
      ++tagc;
      {
        const int32_t rel = excl - b.q0;
        if (len > 0 && rel >= 0 && rel < kOwn) own[rel] = (uint16_t)(tag | (uint32_t)lane);
      }
      __builtin_amdgcn_wave_barrier();
      // The owner of product q is the last lane t with excl_t <= q: the
      // latest mark at or before q (a wavefront max-scan of the stride's
      // marks), or, before the stride's first mark, the segment that holds
      // the stride's start (the carry; a segment covering the whole stride
      // is that case).
      // lanes holding a segment; with the compares below written as lane
      // masks (sicmp) rather than ballots of a bool, which widen it first
      const uint64_t live = __builtin_amdgcn_sicmp(len, 0, kIcmpSgt);
      if constexpr (kNumeric) {
        // strides in groups of kFzG: all the group's mark reads, then its
        // scans and base / A-value reads, then its B loads, so that the LDS
        // round trips of a group overlap
#pragma unroll
        for (int g = 0; g < kU; g += kFzG) {
          if (b.q0 + g * kWave >= b.total) break;  // uniform; apply stops at the same stride
          uint32_t mk[kFzG], cr[kFzG];
#pragma unroll
          for (int i = 0; i < kFzG; ++i) {
            const int32_t s0 = b.q0 + (g + i) * kWave;
            const uint64_t before = live & __builtin_amdgcn_sicmp(excl, s0, kIcmpSle);  // nonzero: total > 0
            cr[i] = tag | (uint32_t)(63 - __builtin_clzll(before));
            mk[i] = own[(g + i) * kWave + lane];
          }
          int32_t bi[kFzG];
          double ai[kFzG];
#pragma unroll
          for (int i = 0; i < kFzG; ++i) {
            const int o = (int)(max(wave_incl_max(mk[i]), cr[i]) & 63u);
            bi[i] = s_base[kRec * o];
            ai[i] = s_av[2 * o];
          }
#pragma unroll
          for (int i = 0; i < kFzG; ++i) {
            const int u = g + i;
            const int32_t s0 = b.q0 + u * kWave;  // the stride's first product (uniform)
            if (s0 < b.total) {
              const int32_t idx = bi[i] + min(s0 + lane, b.total - 1);
              kc[slot][u] = fz_col<kOff32>(B2_crd, idx);
              bv[slot][u] = fz_val<kOff32>(Bx, idx);
              av[slot][u] = ai[i];
            }
          }
        }
      } else {
        // the symbolic walk's wide windows make covering segments common:
        // they take the owner from registers, the rest search the marks
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int32_t s0 = b.q0 + u * kWave;  // the stride's first product (uniform)
          if (s0 >= b.total) break;             // apply stops at the same stride
          const int32_t q = min(s0 + lane, b.total - 1);
          const int32_t s_end = min(s0 + kWave, b.total);
          const uint64_t at_or_before = live & __builtin_amdgcn_sicmp(excl, s0, kIcmpSle);
          const uint64_t cov = at_or_before & __builtin_amdgcn_sicmp(excl + len, s_end, kIcmpSge);
          int32_t idx;
          if (cov) {
            idx = __builtin_amdgcn_readlane(base, __builtin_ctzll(cov)) + q;
          } else {
            const uint64_t before = at_or_before;
            const uint32_t carry = tag | (uint32_t)(63 - __builtin_clzll(before));
            const int o = (int)(max(wave_incl_max(own[u * kWave + lane]), carry) & 63u);
            idx = s_base[o] + q;
          }
          kc[slot][u] = fz_col<kOff32>(B2_crd, idx);
        }
      }
      // the A values read from LDS above are used only in apply: wait for
      // them here, where they have long arrived, so that apply's first use
      // does not wait for its own LDS atomics as well (LDS operations
      // retire in order)
      lgkm_drain();
    };
    auto apply = [&](const FzBatch& b, int slot) {
      const uint32_t wmask = (uint32_t)S - 1;
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (b.q0 + u * kWave >= b.total) break;
        const bool ok = b.q0 + u * kWave + lane < b.total;
        const uint32_t k = kc[slot][u] & wmask;
        const uint32_t wi = k >> 5, bit = 1u << (k & 31);
        if (!kNumeric) {
          if (ok) atomicOr(&seen[wi], bit);
          continue;
        }
        const double v = av[slot][u] * bv[slot][u];
        if (!kOrdered) {
          if (ok) {
            atomicAdd(&acc[k], v);
            atomicOr(&seen[wi], bit);
          }
          continue;
        }
        uint32_t old = 0;
        double a = 0.0;
        if (ok) {
          old = atomicOr(&sb[wi], bit);
          a = acc[k];
        }
        double sum = a + v;
        bool lead = ok;
        uint64_t todo = __ballot(ok && (old & bit) != 0);
        while (todo) {
          const int l = __builtin_ctzll(todo);
          const uint32_t kl = (uint32_t)__builtin_amdgcn_readlane((int)k, l);
          const uint64_t grp = __ballot(ok && k == kl);
          todo &= ~grp;
          const int ld = __builtin_ctzll(grp);
          for (uint64_t mm = grp & (grp - 1); mm; mm &= mm - 1) {
            const double vl = readlane_f64(v, __builtin_ctzll(mm));
            if (lane == ld) sum = sum + vl;
          }
          if (lane != ld && ((grp >> lane) & 1)) lead = false;
        }
        if (lead) {
          acc[k] = sum;
          atomicOr(&seen[wi], bit);
        }
        if (ok) atomicAnd(&sb[wi], ~bit);
        __builtin_amdgcn_wave_barrier();
      }
      if (!b.win_end) return;
      __builtin_amdgcn_wave_barrier();
      const uint32_t myw = lane < words ? seen[lane] : 0u;
      int cnt = __popc(myw);
      if (!kNumeric)  // wide windows: more than one bitmap word per lane
        for (int q = lane + kWave; q < words; q += kWave) {
          cnt += __popc(seen[q]);
          seen[q] = 0u;
        }
      __builtin_amdgcn_wave_barrier();
      if (lane < words) seen[lane] = 0u;
      const int incl = wave_incl_scan(cnt);
      const int total = __builtin_amdgcn_readlane(incl, kWave - 1);
      if (kNumeric && total > 0 && out0 + run + total <= out_end) {
        // at most kFzList entries are listed at a time: a denser window is
        // listed and written in two phases (uniform), lanes 0-31's columns
        // (the entries before split) and then lanes 32-63's
        const int split = __builtin_amdgcn_readlane(incl, 31);
        const int phases = total > kFzList ? 2 : 1;
        const int32_t wb = b.w << s;
        for (int h = 0; h < phases; ++h) {
          const int lo = h == 0 ? 0 : split;  // the phase's first entry of the window
          const int n = phases == 1 ? total : (h == 0 ? split : total - split);
          if (phases == 1 || (lane >> 5) == h) {
            int rr = incl - cnt - lo;
            for (uint32_t bits = myw; bits; bits &= bits - 1)
              list[rr++] = (uint16_t)(lane * 32 + __builtin_ctz(bits));
          }
          __builtin_amdgcn_wave_barrier();
          int32_t* oc = C_crd + out0 + run + lo;
          double* ov = C_val + out0 + run + lo;
          for (int r0 = 0; r0 < n; r0 += 4 * kWave) {
            uint32_t kq[4];
            double aq[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) kq[i] = list[min(r0 + i * kWave + lane, n - 1)];
#pragma unroll
            for (int i = 0; i < 4; ++i) aq[i] = acc[kq[i]];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int q = r0 + i * kWave + lane;
              if (q < n) {
                oc[q] = wb + (int32_t)kq[i];
                ov[q] = aq[i];
                acc[kq[i]] = -0.0;
              }
            }
          }
          // the next phase's list writes follow this phase's reads
          __builtin_amdgcn_wave_barrier();
        }
        // nothing of the emit stays in flight: without this, a list or
        // value read on a path with no lane left to store it counts as
        // pending at every stride of the next batch
        lgkm_drain();
      }
      run += total;
    };

    FzBatch c0;
    c0.w = 0;
    c0.c = 0;
    c0.q0 = 0;
    load_chunk(0, 0);
    scan(c0);
    c0.win_end = (kWave * kU >= c0.total) && (kWave >= E);
    load(c0, 0);
    while (c0.w >= 0) {
      vm_drain();
      FzBatch c1 = next(c0);
      if (c1.w >= 0) load(c1, 1);
      apply(c0, 0);
      if (c1.w < 0) break;
      vm_drain();
      c0 = next(c1);
      if (c0.w >= 0) load(c0, 0);
      apply(c1, 1);
    }
    if (!kNumeric && lane == 0) row_cnt[r] = (int32_t)run;
  }
}

static int bits_for(int64_t v) {  // bits to represent values in [0, v)
  int b = 0;
  while (b < 63 && ((int64_t)1 << b) < v) ++b;
  return b;
}

// Output columns/values: allocated at min(total products, free memory less
// 2 GiB), grown (with a copy of what is written) when a chunk needs more.
static int outer_output_init(spws_ctx* ctx, int32_t m, int64_t total_products) {
  int rc;
  if ((rc = ctx->out_pos.ensure(((size_t)m + 1) * 8))) return rc;
  size_t free_b = 0, tot_b = 0;
  SPWS_HIP(hipMemGetInfo(&free_b, &tot_b));
  const size_t cur = ctx->out_crd.bytes / 4 < ctx->out_val.bytes / 8 ? ctx->out_crd.bytes / 4 : ctx->out_val.bytes / 8;
  const size_t avail = (free_b > ((size_t)2 << 30) ? free_b - ((size_t)2 << 30) : 0) / 12 + cur;
  size_t want = (size_t)(total_products > 0 ? total_products : 1);
  if (want > avail) want = avail > 1 ? avail : 1;
  if (cur < want) {
    ctx->out_crd.release();
    ctx->out_val.release();
    if ((rc = ctx->out_crd.ensure(want * 4)) || (rc = ctx->out_val.ensure(want * 8))) return rc;
  }
  return SPWS_OK;
}

static size_t outer_output_capacity(const spws_ctx* ctx) {
  return ctx->out_crd.bytes / 4 < ctx->out_val.bytes / 8 ? ctx->out_crd.bytes / 4 : ctx->out_val.bytes / 8;
}

// Launches inside the scope go to stream s (when not null) instead of the
// context's stream, timing included.
struct StreamScope {
  spws_ctx* ctx;
  hipStream_t saved;
  StreamScope(spws_ctx* c, hipStream_t s) : ctx(c), saved(c->stream) {
    if (s) c->stream = s;
  }
  ~StreamScope() { ctx->stream = saved; }
};

static int outer_output_grow(spws_ctx* ctx, int64_t base, int64_t U) {
  hipStream_t s = ctx->stream;
  int rc;
  const size_t have = ctx->out_crd.bytes / 4 < ctx->out_val.bytes / 8 ? ctx->out_crd.bytes / 4 : ctx->out_val.bytes / 8;
  if ((size_t)(base + U) <= have) return SPWS_OK;
  const size_t want = (size_t)(base + U) + (size_t)(base + U) / 2;
  DevBuf nc, nv;
  if ((rc = nc.ensure(want * 4)) || (rc = nv.ensure(want * 8))) {
    nc.release();
    nv.release();
    return rc;
  }
  SPWS_HIP(hipMemcpyAsync(nc.ptr, ctx->out_crd.ptr, (size_t)base * 4, hipMemcpyDeviceToDevice, s));
  SPWS_HIP(hipMemcpyAsync(nv.ptr, ctx->out_val.ptr, (size_t)base * 8, hipMemcpyDeviceToDevice, s));
  SPWS_HIP(hipStreamSynchronize(s));
  ctx->out_crd.release();
  ctx->out_val.release();
  ctx->out_crd = nc;
  ctx->out_val = nv;
  nc.ptr = nv.ptr = nullptr;
  return SPWS_OK;
}

struct OuterArgs {
  int32_t m, kdim, n;
  int32_t a_ncols;
  const int32_t *A1_crd, *A2_pos, *A2_crd;
  const double* Ax;
  int32_t b_nrows;
  const int32_t *B1_crd, *B2_pos, *B2_crd;
  const double* Bx;
};

template <class K>
static int outer_chunks(spws_ctx* ctx, const OuterArgs& a, const std::vector<int64_t>& row_prefix, int kb,
                        int64_t max_products, int32_t max_rows, int64_t* nnz_out) {
  hipStream_t s = ctx->stream;
  int rc;
  const int32_t m = a.m;
  const int32_t na = a.a_ncols;
  const size_t nac = (size_t)(na > 0 ? na : 1);
  int32_t* match = ctx->obuf[0].as<int32_t>();
  int32_t* blen = ctx->obuf[1].as<int32_t>();
  if ((rc = ctx->obuf[2].ensure(nac * 4 * 3)) || (rc = ctx->obuf[3].ensure((nac + 1) * 4)) ||
      (rc = ctx->obuf[4].ensure(nac * 4 * 3)) || (rc = ctx->obuf[5].ensure((nac + 1) * 8 * 2)))
    return rc;
  int32_t* col_lo = ctx->obuf[2].as<int32_t>();
  int32_t* col_m = col_lo + nac;
  int32_t* flag = col_m + nac;
  int32_t* flag_off = ctx->obuf[3].as<int32_t>();
  int32_t* act_astart = ctx->obuf[4].as<int32_t>();
  int32_t* act_L = act_astart + nac;
  int32_t* act_bstart = act_L + nac;
  int64_t* act_cnt = ctx->obuf[5].as<int64_t>();
  int64_t* act_off = act_cnt + (nac + 1);

  // chunk boundaries on the host
  std::vector<std::pair<int32_t, int32_t>> chunks;
  int64_t max_chunk = 0;
  for (int32_t r0 = 0; r0 < m;) {
    int32_t r1 = r0 + 1;
    while (r1 < m && r1 - r0 < max_rows && row_prefix[r1 + 1] - row_prefix[r0] <= max_products) ++r1;
    chunks.push_back({r0, r1});
    max_chunk = std::max(max_chunk, row_prefix[r1] - row_prefix[r0]);
    r0 = r1;
  }
  SPWS_REQUIRE(max_chunk < ((int64_t)1 << 31), SPWS_ERR_UNSUPPORTED,
               "an output row range with %lld products exceeds the int32 workspace index", (long long)max_chunk);
  const size_t cap = (size_t)(max_chunk > 0 ? max_chunk : 1);
  if ((rc = ctx->keys[0].ensure(cap * sizeof(K))) || (rc = ctx->keys[1].ensure(cap * sizeof(K))) ||
      (rc = ctx->vals[0].ensure(cap * 8)) || (rc = ctx->vals[1].ensure(cap * 8)))
    return rc;
  if ((rc = outer_output_init(ctx, m, row_prefix[m]))) return rc;
  const int64_t ntiles_max = div_up((int64_t)cap, kRadixTile);
  DevBuf& heads = ctx->chunk_a;
  if ((rc = heads.ensure((size_t)(2 * ntiles_max + 2) * 4))) return rc;
  int32_t* tile_heads = heads.as<int32_t>();
  int32_t* tile_uoff = tile_heads + ntiles_max + 1;
  int64_t* hm = ctx->host_meta.as<int64_t>();
  int64_t base = 0;
  int64_t* out_pos = ctx->out_pos.as<int64_t>();
  const unsigned gcol = (unsigned)div_up(na > 0 ? na : 1, 256);

  for (auto& ch : chunks) {
    const int32_t r0 = ch.first, r1 = ch.second, R = r1 - r0;
    const int64_t P = row_prefix[r1] - row_prefix[r0];
    const int rb = bits_for(R);
    const int key_bits = kb + rb;
    ctx->stats.chunks++;
    ctx->stats.sort_bits = std::max<int64_t>(ctx->stats.sort_bits, key_bits);
    K* kbuf = ctx->keys[0].as<K>();
    double* vbuf = ctx->vals[0].as<double>();
    int64_t U = 0;
    if (P > 0) {
      SPWS_LAUNCH(ctx, "outer_colrange", k_outer_colrange, dim3(gcol), dim3(256), 0, na, r0, r1, a.A2_pos, a.A2_crd,
                  blen, col_lo, col_m, flag);
      if ((rc = exclusive_scan<int32_t, int32_t>(ctx, flag, flag_off, na))) return rc;
      // entries past the active count stay zero, so the scan over all na
      // entries leaves act_off[a] = P for every a >= the active count
      SPWS_HIP(hipMemsetAsync(act_cnt, 0, nac * 8, s));
      SPWS_LAUNCH(ctx, "outer_compact", k_outer_compact, dim3(gcol), dim3(256), 0, na, flag_off, col_lo, col_m,
                  match, blen, a.B2_pos, act_astart, act_L, act_bstart, act_cnt, (int32_t*)nullptr,
                  (int32_t*)nullptr);
      if ((rc = exclusive_scan<int64_t, int64_t>(ctx, act_cnt, act_off, na))) return rc;
      const unsigned gexp = (unsigned)div_up(P, kExpTile);
      SPWS_LAUNCH(ctx, "outer_expand", (k_outer_expand<K>), dim3(gexp), dim3(kExpBlock), 0, P, flag_off + na, act_off,
                  act_astart, act_L, act_bstart, a.A2_crd, a.Ax, a.B2_crd, a.Bx, r0, kb, kbuf, vbuf);
      bool in_alt = false;
      if ((rc = radix_sort_pairs<K>(ctx, kbuf, vbuf, ctx->keys[1].as<K>(), ctx->vals[1].as<double>(), P, 0, key_bits,
                                    &in_alt)))
        return rc;
      if (in_alt) {
        kbuf = ctx->keys[1].as<K>();
        vbuf = ctx->vals[1].as<double>();
      }
      const int64_t ntiles = div_up(P, kRadixTile);
      SPWS_LAUNCH(ctx, "outer_heads", (k_outer_heads<K>), dim3((unsigned)ntiles), dim3(kRadixBlock), 0, kbuf, P,
                  tile_heads);
      if ((rc = exclusive_scan<int32_t, int32_t>(ctx, tile_heads, tile_uoff, ntiles))) return rc;
      SPWS_HIP(hipMemcpyAsync(hm, tile_uoff + ntiles, 4, hipMemcpyDeviceToHost, s));
      SPWS_HIP(hipStreamSynchronize(s));
      U = *reinterpret_cast<int32_t*>(hm);
      if ((rc = outer_output_grow(ctx, base, U))) return rc;
      SPWS_LAUNCH(ctx, "outer_reduce", (k_outer_reduce<K>), dim3((unsigned)ntiles), dim3(kRadixBlock), 0, kbuf, vbuf,
                  P, tile_uoff, kb, base, r0, ctx->out_crd.as<int32_t>(), ctx->out_val.as<double>(), out_pos);
    }
    SPWS_LAUNCH(ctx, "outer_tail", (k_outer_tail<K>), dim3((unsigned)div_up(R, 256)), dim3(256), 0, kbuf, P, kb, r0,
                R, base + U, out_pos);
    base += U;
  }
  SPWS_LAUNCH(ctx, "set_i64", k_set_i64, dim3(1), dim3(1), 0, out_pos + m, base);
  SPWS_HIP(hipStreamSynchronize(s));
  *nnz_out = base;
  return SPWS_OK;
}

// Bucket method over output-row chunks (see the file comment).  Chunks are
// bounded by products (int32 slots), by the entry table (entries x windows)
// and by the bucket count (rows x windows).
static int outer_chunks_bucket(spws_ctx* ctx, const OuterArgs& a, const std::vector<int64_t>& row_prefix,
                               const std::vector<int64_t>& pair_prefix, const int32_t* row_pairs, int s,
                               int64_t max_products, int32_t max_rows, int64_t* nnz_out) {
  hipStream_t st = ctx->stream;
  int rc;
  const int32_t m = a.m;
  const int32_t na = a.a_ncols;
  const size_t nac = (size_t)(na > 0 ? na : 1);
  const int32_t W = (int32_t)std::max<int64_t>(1, div_up((int64_t)a.n, (int64_t)1 << s));
  int32_t* match = ctx->obuf[0].as<int32_t>();
  int32_t* blen = ctx->obuf[1].as<int32_t>();
  if ((rc = ctx->obuf[2].ensure(nac * 4 * 3)) || (rc = ctx->obuf[3].ensure((nac + 1) * 4)) ||
      (rc = ctx->obuf[4].ensure(nac * 4 * 5)) || (rc = ctx->obuf[5].ensure((nac + 1) * 8 * 2)) ||
      (rc = ctx->obuf[6].ensure((nac + 1) * 4)) || (rc = ctx->bk[0].ensure(nac * (size_t)(W + 1) * 4)))
    return rc;
  int32_t* col_lo = ctx->obuf[2].as<int32_t>();
  int32_t* col_m = col_lo + nac;
  int32_t* flag = col_m + nac;
  int32_t* flag_off = ctx->obuf[3].as<int32_t>();
  int32_t* act_astart = ctx->obuf[4].as<int32_t>();
  int32_t* act_L = act_astart + nac;
  int32_t* act_bstart = act_L + nac;
  int32_t* act_ja = act_bstart + nac;
  int32_t* act_m = act_ja + nac;
  int64_t* act_cnt = ctx->obuf[5].as<int64_t>();
  int64_t* act_off = act_cnt + (nac + 1);
  int32_t* act_eoff = ctx->obuf[6].as<int32_t>();
  int32_t* bseg = ctx->bk[0].as<int32_t>();

  const unsigned gcol = (unsigned)div_up(na > 0 ? na : 1, 256);
  if (na > 0) {
    const unsigned g = (unsigned)std::min<int64_t>(div_up((int64_t)na * (W + 1), 256), 16384);
    SPWS_LAUNCH(ctx, "bk_bseg", k_bk_bseg, dim3(g), dim3(256), 0, na, W, s, match, blen, a.B2_pos, a.B2_crd, bseg);
  }

  constexpr int64_t kMaxEtab = (int64_t)1 << 28;
  constexpr int64_t kMaxBuckets = (int64_t)1 << 27;
  std::vector<std::pair<int32_t, int32_t>> chunks;
  int64_t max_chunk = 0, max_en = 0, max_r = 0;
  for (int32_t r0 = 0; r0 < m;) {
    int32_t r1 = r0 + 1;
    while (r1 < m && r1 - r0 < max_rows && row_prefix[r1 + 1] - row_prefix[r0] <= max_products &&
           (pair_prefix[r1 + 1] - pair_prefix[r0]) * W <= kMaxEtab && (int64_t)(r1 + 1 - r0) * W <= kMaxBuckets)
      ++r1;
    chunks.push_back({r0, r1});
    max_chunk = std::max(max_chunk, row_prefix[r1] - row_prefix[r0]);
    max_en = std::max(max_en, pair_prefix[r1] - pair_prefix[r0]);
    max_r = std::max<int64_t>(max_r, r1 - r0);
    r0 = r1;
  }
  SPWS_REQUIRE(max_chunk < ((int64_t)1 << 31), SPWS_ERR_UNSUPPORTED,
               "an output row range with %lld products exceeds the int32 workspace index", (long long)max_chunk);
  const size_t cap = (size_t)(max_chunk > 0 ? max_chunk : 1);
  const size_t ecap = (size_t)(max_en > 0 ? max_en : 1);
  const size_t nbcap = (size_t)max_r * (size_t)W + 1;
  // The reduce of chunk c runs on the auxiliary stream while the main
  // stream prepares chunk c + 1 (column ranges, entry sort, bucket tables);
  // the expand of c + 1 waits for that reduce.  So the bucket tables come in
  // two sets and the products and column offsets in one.  Column offsets:
  // k_bk_count reads them in aligned groups of 4.
  const int nsets = chunks.size() > 1 ? 2 : 1;
  DevBuf* tabs[2] = {&ctx->bk[8], &ctx->bk[10]};
  if ((rc = ctx->keys[0].ensure(cap * 2 + 8)) || (rc = ctx->vals[0].ensure(cap * 8))) return rc;
  for (int t = 0; t < nsets; ++t)
    if ((rc = tabs[t]->ensure(nbcap * 4 * 4))) return rc;
  if ((rc = ctx->bk[1].ensure(ecap * 4)) || (rc = ctx->bk[2].ensure(ecap * 4)) ||
      (rc = ctx->bk[3].ensure(ecap * 8)) || (rc = ctx->bk[4].ensure(ecap * 8)) ||
      (rc = ctx->bk[5].ensure(ecap * 4)) || (rc = ctx->bk[6].ensure(ecap * (size_t)W * 4)) ||
      (rc = ctx->bk[7].ensure(((size_t)max_r + 1) * 4)) || (rc = ctx->bk[9].ensure(ecap * 4)))
    return rc;
  if ((rc = outer_output_init(ctx, m, row_prefix[m]))) return rc;
  hipStream_t aux = nullptr;
  if (nsets > 1) {
    if ((rc = ctx_aux_stream(ctx, &aux))) return rc;
  }
  int32_t* ent_ja = ctx->bk[5].as<int32_t>();
  int32_t* ent_ai = ctx->bk[9].as<int32_t>();
  int32_t* etab = ctx->bk[6].as<int32_t>();
  int32_t* rp_off = ctx->bk[7].as<int32_t>();
  int64_t* out_pos = ctx->out_pos.as<int64_t>();
  int32_t* hm = ctx->host_meta.as<int32_t>();
  const size_t lds_count = (size_t)kBkWaves * (((size_t)1 << s) >> 5) * 4;
  const size_t lds_reduce = (size_t)kBkRedWaves * bk_reduce_lds_per_wave(s);
  ctx->stats.outer_method = 2;
  ctx->stats.window_bits = s;
  int64_t base = 0;
  bool pending[2] = {false, false};  // set t's last reduce not yet waited for
  // on any error return, let the auxiliary stream's reduce finish first
  struct AuxDrain {
    hipStream_t s;
    ~AuxDrain() {
      if (s) (void)hipStreamSynchronize(s);
    }
  } aux_drain{aux};
  int ci = 0;

  for (auto& ch : chunks) {
    const int t = ci++ % nsets;
    uint16_t* klo = ctx->keys[0].as<uint16_t>();
    double* pv = ctx->vals[0].as<double>();
    int32_t* bsize = tabs[t]->as<int32_t>();
    int32_t* bstart = bsize + nbcap;
    int32_t* bnnz = bstart + nbcap;
    int32_t* boff = bnnz + nbcap;
    const int32_t r0 = ch.first, r1 = ch.second, R = r1 - r0;
    const int64_t P = row_prefix[r1] - row_prefix[r0];
    const int32_t En = (int32_t)(pair_prefix[r1] - pair_prefix[r0]);
    const int32_t NB = R * W;
    ctx->stats.chunks++;
    int64_t U = 0;
    if (P > 0) {
      SPWS_LAUNCH(ctx, "outer_colrange", k_outer_colrange, dim3(gcol), dim3(256), 0, na, r0, r1, a.A2_pos, a.A2_crd,
                  blen, col_lo, col_m, flag);
      if ((rc = exclusive_scan<int32_t, int32_t>(ctx, flag, flag_off, na))) return rc;
      SPWS_HIP(hipMemsetAsync(act_cnt, 0, nac * 8, st));
      SPWS_HIP(hipMemsetAsync(act_m, 0, nac * 4, st));
      SPWS_LAUNCH(ctx, "outer_compact", k_outer_compact, dim3(gcol), dim3(256), 0, na, flag_off, col_lo, col_m,
                  match, blen, a.B2_pos, act_astart, act_L, act_bstart, act_cnt, act_ja, act_m);
      if ((rc = exclusive_scan<int64_t, int64_t>(ctx, act_cnt, act_off, na))) return rc;
      if ((rc = exclusive_scan<int32_t, int32_t>(ctx, act_m, act_eoff, na))) return rc;
      uint32_t* pkey = ctx->bk[1].as<uint32_t>();
      double* pval = ctx->bk[3].as<double>();
      const unsigned ge = (unsigned)std::min<int64_t>(div_up(En, 256), 16384);
      SPWS_LAUNCH(ctx, "bk_entries", k_bk_entries, dim3(ge), dim3(256), 0, En, flag_off + na, act_eoff, act_astart,
                  act_ja, a.A2_crd, r0, pkey, pval, ent_ja, ent_ai);
      bool alt = false;
      if ((rc = radix_sort_pairs<uint32_t>(ctx, pkey, pval, ctx->bk[2].as<uint32_t>(), ctx->bk[4].as<double>(), En, 0,
                                           std::max(1, bits_for(R)), &alt)))
        return rc;
      if (alt) pval = ctx->bk[4].as<double>();
      if ((rc = exclusive_scan<int32_t, int32_t>(ctx, row_pairs + r0, rp_off, R))) return rc;
      // set t is free once the reduce that last read it has finished
      if (pending[t]) {
        SPWS_HIP(hipStreamWaitEvent(st, ctx->aux_done[t], 0));
        pending[t] = false;
      }
      const unsigned gr = (unsigned)std::min<int64_t>(div_up((int64_t)R * kWave, 256), 16384);
      SPWS_LAUNCH(ctx, "bk_rowsize", k_bk_rows<0>, dim3(gr), dim3(256), 0, R, W, rp_off, pval, ent_ja, bseg,
                  (const int32_t*)nullptr, bsize);
      if ((rc = exclusive_scan<int32_t, int32_t>(ctx, bsize, bstart, NB))) return rc;
      SPWS_LAUNCH(ctx, "bk_etab", k_bk_rows<1>, dim3(gr), dim3(256), 0, R, W, rp_off, pval, ent_ja, bseg, bstart,
                  etab);
      // the expand waits for the previous chunk's reduce: only this chunk's
      // preparation (column ranges, entry sort, bucket tables) overlaps it,
      // so the two memory-bound kernels never share the GPU
      if (pending[1 - t]) {
        SPWS_HIP(hipStreamWaitEvent(st, ctx->aux_done[1 - t], 0));
        pending[1 - t] = false;
      }
      const unsigned gx = (unsigned)std::min<int64_t>(div_up((int64_t)En * kWave, 256), 32768);
      if (W <= kBkExpandLdsW) {
        SPWS_LAUNCH(ctx, "bk_expand", k_bk_expand_lds<8>, dim3(gx), dim3(256), (size_t)(256 / kWave) * W * 4, En, pval,
                    ent_ai, ent_ja, match, blen, a.B2_pos, a.Ax, a.B2_crd, a.Bx, W, s, etab, klo, pv);
      } else {
        SPWS_LAUNCH(ctx, "bk_expand", k_bk_expand, dim3(gx), dim3(256), 0, En, pval, ent_ai, ent_ja, match, blen,
                    a.B2_pos, a.Ax, a.B2_crd, a.Bx, W, s, etab, klo, pv);
      }
      const unsigned gb = (unsigned)std::min<int64_t>(div_up(NB, kBkWaves), 65536);
      SPWS_LAUNCH(ctx, "bk_count", k_bk_count, dim3(gb), dim3(kBkWaves * kWave), lds_count, NB, s, bstart, klo,
                  bnnz);
      if ((rc = exclusive_scan<int32_t, int32_t>(ctx, bnnz, boff, NB))) return rc;
      SPWS_HIP(hipMemcpyAsync(hm, boff + NB, 4, hipMemcpyDeviceToHost, st));
      SPWS_HIP(hipStreamSynchronize(st));
      U = *hm;
      if (outer_output_capacity(ctx) < (size_t)(base + U) && aux) {
        // growing moves the output: the reduce in flight must finish first
        SPWS_HIP(hipStreamSynchronize(aux));
      }
      if ((rc = outer_output_grow(ctx, base, U))) return rc;
      // row pointers and the reduce: on the auxiliary stream when there is one
      StreamScope on(ctx, aux);
      if (aux) {
        SPWS_HIP(hipEventRecord(ctx->aux_ready, st));
        SPWS_HIP(hipStreamWaitEvent(aux, ctx->aux_ready, 0));
      }
      SPWS_LAUNCH(ctx, "bk_rowpos", k_bk_rowpos, dim3((unsigned)std::min<int64_t>(div_up(R, 256), 4096)), dim3(256), 0,
                  R, W, r0, base, boff, out_pos);
      const unsigned gred = (unsigned)std::min<int64_t>(div_up(NB, kBkRedWaves), 65536);
      SPWS_LAUNCH(ctx, "bk_reduce", k_bk_reduce<kBkRedUnroll>, dim3(gred), dim3(kBkRedWaves * kWave), lds_reduce, NB, W,
                  s, bstart, klo, pv, boff, base, ctx->out_crd.as<int32_t>(), ctx->out_val.as<double>());
      if (aux) {
        SPWS_HIP(hipEventRecord(ctx->aux_done[t], aux));
        pending[t] = true;
      }
    } else {
      SPWS_LAUNCH(ctx, "outer_tail", (k_outer_tail<uint32_t>), dim3((unsigned)div_up(R, 256)), dim3(256), 0,
                  (const uint32_t*)nullptr, (int64_t)0, 0, r0, R, base, out_pos);
    }
    base += U;
  }
  if (aux) SPWS_HIP(hipStreamSynchronize(aux));
  SPWS_LAUNCH(ctx, "set_i64", k_set_i64, dim3(1), dim3(1), 0, out_pos + m, base);
  SPWS_HIP(hipStreamSynchronize(st));
  *nnz_out = base;
  return SPWS_OK;
}

constexpr int kFzSymWavesPerCu = 24;  // k_fz_rows<false>: one wave per workgroup, 76 VGPRs, 6.4 KB of LDS
constexpr int kFzSymWindowLog2 = 15;  // widest window of the symbolic walk (a 4 KiB bitmap)
constexpr int kFzNumWavesPerCu = 8;   // k_fz_rows<true>: bounded by the LDS window

// Fused method over all rows at once (see the fused-method comment): A's
// matched entries grouped by row (sorted by row, stable, so increasing j
// within a row), one record per entry, the symbolic pass and its prefix
// (exact row pointers, and an output allocated to the exact size), then the
// numeric pass writing every entry in place.
static int outer_fused(spws_ctx* ctx, const OuterArgs& a, const std::vector<int64_t>& pair_prefix,
                       const int32_t* row_pairs, int s, bool ordered, bool off32, int64_t* nnz_out) {
  hipStream_t st = ctx->stream;
  int rc;
  const int32_t m = a.m;
  const int32_t na = a.a_ncols;
  const size_t nac = (size_t)(na > 0 ? na : 1);
  const int32_t W = (int32_t)std::max<int64_t>(1, div_up((int64_t)a.n, (int64_t)1 << s));
  const int64_t En64 = pair_prefix[m];
  SPWS_REQUIRE(En64 < ((int64_t)1 << 31), SPWS_ERR_UNSUPPORTED, "%lld matched A entries exceed the int32 range",
               (long long)En64);
  const int32_t En = (int32_t)En64;
  const size_t ecap = (size_t)(En > 0 ? En : 1);
  const size_t mm = (size_t)(m > 0 ? m : 1);
  int32_t* match = ctx->obuf[0].as<int32_t>();
  int32_t* blen = ctx->obuf[1].as<int32_t>();
  if ((rc = ctx->obuf[2].ensure(nac * 4 * 3)) || (rc = ctx->obuf[3].ensure((nac + 1) * 4)) ||
      (rc = ctx->obuf[4].ensure(nac * 4 * 5)) || (rc = ctx->obuf[5].ensure((nac + 1) * 8 * 2)) ||
      (rc = ctx->obuf[6].ensure((nac + 1) * 4)) || (rc = ctx->bk[0].ensure(nac * (size_t)(W + 1) * 4)) ||
      (rc = ctx->bk[1].ensure(ecap * 4)) || (rc = ctx->bk[2].ensure(ecap * 4)) ||
      (rc = ctx->bk[3].ensure(ecap * 8)) || (rc = ctx->bk[4].ensure(ecap * 8)) ||
      (rc = ctx->bk[5].ensure(ecap * 4)) || (rc = ctx->bk[9].ensure(ecap * 4)) ||
      (rc = ctx->bk[6].ensure(ecap * 16)) || (rc = ctx->bk[7].ensure((mm + 1) * 4)) ||
      (rc = ctx->bk[8].ensure((mm + 64) * 4)) || (rc = ctx->out_pos.ensure((mm + 1) * 8)))
    return rc;
  int32_t* col_lo = ctx->obuf[2].as<int32_t>();
  int32_t* col_m = col_lo + nac;
  int32_t* flag = col_m + nac;
  int32_t* flag_off = ctx->obuf[3].as<int32_t>();
  int32_t* act_astart = ctx->obuf[4].as<int32_t>();
  int32_t* act_L = act_astart + nac;
  int32_t* act_bstart = act_L + nac;
  int32_t* act_ja = act_bstart + nac;
  int32_t* act_m = act_ja + nac;
  int64_t* act_cnt = ctx->obuf[5].as<int64_t>();
  int32_t* act_eoff = ctx->obuf[6].as<int32_t>();
  int32_t* bseg = ctx->bk[0].as<int32_t>();
  int32_t* rp_off = ctx->bk[7].as<int32_t>();
  int32_t* row_cnt = ctx->bk[8].as<int32_t>();
  int32_t* work = row_cnt + mm + 16;  // 256-byte apart from the counts
  int32_t* rec_b = ctx->bk[6].as<int32_t>();
  int32_t* rec_ja = rec_b + ecap;
  double* rec_a = reinterpret_cast<double*>(rec_ja + ecap);
  int64_t* out_pos = ctx->out_pos.as<int64_t>();
  ctx->stats.outer_method = ordered ? 3 : 4;
  ctx->stats.window_bits = s;
  ctx->stats.chunks = 1;
  const unsigned gcol = (unsigned)div_up(na > 0 ? na : 1, 256);
  if (na > 0) {
    const unsigned g = (unsigned)std::min<int64_t>(div_up((int64_t)na * (W + 1), 256), 16384);
    SPWS_LAUNCH(ctx, "bk_bseg", k_bk_bseg, dim3(g), dim3(256), 0, na, W, s, match, blen, a.B2_pos, a.B2_crd, bseg);
  }
  // A's matched entries of every row, grouped by row (increasing j within a row)
  if (En > 0) {
    SPWS_LAUNCH(ctx, "outer_colrange", k_outer_colrange, dim3(gcol), dim3(256), 0, na, 0, m, a.A2_pos, a.A2_crd, blen,
                col_lo, col_m, flag);
    if ((rc = exclusive_scan<int32_t, int32_t>(ctx, flag, flag_off, na))) return rc;
    SPWS_HIP(hipMemsetAsync(act_cnt, 0, nac * 8, st));
    SPWS_HIP(hipMemsetAsync(act_m, 0, nac * 4, st));
    SPWS_LAUNCH(ctx, "outer_compact", k_outer_compact, dim3(gcol), dim3(256), 0, na, flag_off, col_lo, col_m, match,
                blen, a.B2_pos, act_astart, act_L, act_bstart, act_cnt, act_ja, act_m);
    if ((rc = exclusive_scan<int32_t, int32_t>(ctx, act_m, act_eoff, na))) return rc;
    uint32_t* pkey = ctx->bk[1].as<uint32_t>();
    double* pval = ctx->bk[3].as<double>();
    const unsigned ge = (unsigned)std::min<int64_t>(div_up(En, 256), 16384);
    SPWS_LAUNCH(ctx, "bk_entries", k_bk_entries, dim3(ge), dim3(256), 0, En, flag_off + na, act_eoff, act_astart,
                act_ja, a.A2_crd, 0, pkey, pval, ctx->bk[5].as<int32_t>(), ctx->bk[9].as<int32_t>());
    bool alt = false;
    if ((rc = radix_sort_pairs<uint32_t>(ctx, pkey, pval, ctx->bk[2].as<uint32_t>(), ctx->bk[4].as<double>(), En, 0,
                                         std::max(1, bits_for(m)), &alt)))
      return rc;
    if (alt) pval = ctx->bk[4].as<double>();
    SPWS_LAUNCH(ctx, "fz_records", k_fz_records, dim3(ge), dim3(256), 0, En, pval, ctx->bk[5].as<int32_t>(),
                ctx->bk[9].as<int32_t>(), match, a.B2_pos, a.Ax, rec_b, rec_ja, rec_a);
  }
  if ((rc = exclusive_scan<int32_t, int32_t>(ctx, row_pairs, rp_off, m))) return rc;
  int dev_cus = 256;
  (void)hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
  // symbolic pass: entries per row
  SPWS_HIP(hipMemsetAsync(work, 0, 4, st));
  const unsigned gsym = (unsigned)std::max<int64_t>(1, std::min<int64_t>(m, (int64_t)dev_cus * kFzSymWavesPerCu));
  // the symbolic walk's windows: 2^dsym segment-table windows each
  int dsym = 0;
  while (s + dsym < kFzSymWindowLog2 && ((int64_t)1 << dsym) < W) ++dsym;
  if (off32)
    SPWS_LAUNCH(ctx, "fz_count", (k_fz_rows<false, true, true>), dim3(gsym), dim3(kWave),
                fz_lds_bytes(s, dsym, false, false), m, W, s, dsym, rp_off, rec_b, rec_ja, rec_a, bseg, a.B2_crd, a.Bx, row_cnt, (const int64_t*)nullptr,
                (int32_t*)nullptr, (double*)nullptr, work);
  else
    SPWS_LAUNCH(ctx, "fz_count", (k_fz_rows<false, true, false>), dim3(gsym), dim3(kWave),
                fz_lds_bytes(s, dsym, false, false), m, W, s, dsym, rp_off, rec_b, rec_ja, rec_a, bseg, a.B2_crd, a.Bx, row_cnt, (const int64_t*)nullptr,
                (int32_t*)nullptr, (double*)nullptr, work);
  if ((rc = exclusive_scan<int32_t, int64_t>(ctx, row_cnt, out_pos, m))) return rc;
  int64_t* hm = ctx->host_meta.as<int64_t>();
  SPWS_HIP(hipMemcpyAsync(hm, out_pos + m, 8, hipMemcpyDeviceToHost, st));
  SPWS_HIP(hipStreamSynchronize(st));
  const int64_t total = *hm;
  const size_t cap = (size_t)(total > 0 ? total : 1);
  if (outer_output_capacity(ctx) < cap) {
    ctx->out_crd.release();
    ctx->out_val.release();
    if ((rc = ctx->out_crd.ensure(cap * 4)) || (rc = ctx->out_val.ensure(cap * 8))) return rc;
  }
  // numeric pass: every entry written at its final position
  SPWS_HIP(hipMemsetAsync(work, 0, 4, st));
  const unsigned gnum = (unsigned)std::max<int64_t>(1, std::min<int64_t>(m, (int64_t)dev_cus * kFzNumWavesPerCu));
#define SPWS_FZ_REDUCE(ORD, OFF)                                                                              \
  SPWS_LAUNCH(ctx, "fz_reduce", (k_fz_rows<true, ORD, OFF>), dim3(gnum), dim3(kWave), fz_lds_bytes(s, 0, true, ORD), m, \
              W, s, 0, rp_off, rec_b, rec_ja, rec_a, bseg, a.B2_crd, a.Bx, (int32_t*)nullptr,                     \
              (const int64_t*)out_pos, ctx->out_crd.as<int32_t>(), ctx->out_val.as<double>(), work)
  if (ordered && off32)
    SPWS_FZ_REDUCE(true, true);
  else if (ordered)
    SPWS_FZ_REDUCE(true, false);
  else if (off32)
    SPWS_FZ_REDUCE(false, true);
  else
    SPWS_FZ_REDUCE(false, false);
#undef SPWS_FZ_REDUCE
  SPWS_HIP(hipStreamSynchronize(st));
  *nnz_out = total;
  return SPWS_OK;
}

static int outer_run(spws_ctx* ctx, const OuterArgs& a, int64_t* nnz_out) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_REQUIRE(a.m >= 0 && a.n >= 0 && a.kdim >= 0 && a.a_ncols >= 0 && a.b_nrows >= 0, SPWS_ERR_INVALID,
               "negative dimension");
  SPWS_REQUIRE(nnz_out && (a.a_ncols == 0 || (a.A2_pos && a.A2_crd && a.Ax)) &&
                   (a.b_nrows == 0 || (a.B2_pos && a.B2_crd && a.Bx)),
               SPWS_ERR_INVALID, "null pointer argument");
  hipStream_t s = ctx->stream;
  memset(&ctx->stats, 0, sizeof(ctx->stats));
  const int32_t m = a.m;
  const size_t nac = (size_t)(a.a_ncols > 0 ? a.a_ncols : 1);
  const size_t mm = (size_t)(m > 0 ? m : 1);
  if ((rc = ctx->obuf[0].ensure(nac * 4)) || (rc = ctx->obuf[1].ensure(nac * 4)) ||
      (rc = ctx->row_ub.ensure(mm * 8)) || (rc = ctx->obuf[7].ensure(mm * 4)))
    return rc;
  int32_t* match = ctx->obuf[0].as<int32_t>();
  int32_t* blen = ctx->obuf[1].as<int32_t>();
  int32_t* row_pairs = ctx->obuf[7].as<int32_t>();
  SPWS_HIP(hipMemsetAsync(ctx->row_ub.ptr, 0, mm * 8, s));
  SPWS_HIP(hipMemsetAsync(row_pairs, 0, mm * 4, s));
  if (a.a_ncols > 0) {
    const unsigned g = (unsigned)div_up(a.a_ncols, 256);
    SPWS_LAUNCH(ctx, "outer_match", k_outer_match, dim3(g), dim3(256), 0, a.a_ncols, a.A1_crd, a.b_nrows, a.B1_crd,
                a.B2_pos, match, blen);
    const unsigned g2 = (unsigned)std::min<int64_t>(div_up(a.a_ncols, 4), 8192);
    SPWS_LAUNCH(ctx, "outer_row_ub", k_outer_row_ub, dim3(g2), dim3(256), 0, a.a_ncols, a.A2_pos, a.A2_crd, blen,
                (unsigned long long*)ctx->row_ub.ptr, row_pairs);
  }
  std::vector<int64_t> ub(mm), prefix((size_t)m + 1, 0), pair_prefix((size_t)m + 1, 0);
  std::vector<int32_t> pairs(mm);
  int32_t b_nnz = 0;  // B's entries (the fused method's offset width)
  if (m > 0) {
    SPWS_HIP(hipMemcpyAsync(ub.data(), ctx->row_ub.ptr, (size_t)m * 8, hipMemcpyDeviceToHost, s));
    SPWS_HIP(hipMemcpyAsync(pairs.data(), row_pairs, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    if (a.b_nrows > 0) SPWS_HIP(hipMemcpyAsync(&b_nnz, a.B2_pos + a.b_nrows, 4, hipMemcpyDeviceToHost, s));
    SPWS_HIP(hipStreamSynchronize(s));
  }
  int64_t busy_rows = 0;
  for (int32_t i = 0; i < m; ++i) {
    prefix[i + 1] = prefix[i] + ub[i];
    pair_prefix[i + 1] = pair_prefix[i] + pairs[i];
    busy_rows += ub[i] > 0;
  }
  ctx->stats.products = prefix[m];
  const int64_t max_products = ctx->outer_max_products > 0 ? ctx->outer_max_products : ((int64_t)1 << 30);
  const int64_t user_rows = ctx->outer_max_rows > 0 ? ctx->outer_max_rows : INT32_MAX;

  // method: the fused row windows with atomic adds (4) unless the windows
  // would hold too few products each (very sparse, very wide outputs) --
  // then the LSD sort over whole chunks.  The fused method needs no product
  // workspace; on the 200k workload its step is about 7% shorter than the
  // bucket method's in order (3), and about 10% shorter again with the
  // unordered adds (within the 1e-12 bar; 3 keeps the sums reproducible).
  int method = ctx->outer_method;
  int wlog = ctx->outer_window_log2;
  if (wlog == 0) {
    // about 1024 products per (row, window) bucket on average
    const double avg = busy_rows > 0 ? (double)prefix[m] / (double)busy_rows : 1.0;
    const double want = 1024.0 * (double)std::max(a.n, 1) / std::max(avg, 1.0);
    wlog = kBkMinLog2;
    while (wlog < kBkMaxLog2 && (double)((int64_t)1 << wlog) * 1.41421356 < want) ++wlog;
  }
  const int64_t W = std::max<int64_t>(1, div_up((int64_t)a.n, (int64_t)1 << wlog));
  if (method == 0) {
    const double per_bucket = busy_rows > 0 ? (double)prefix[m] / ((double)busy_rows * (double)W) : 0.0;
    int32_t max_pairs = 0;
    for (int32_t i = 0; i < m; ++i) max_pairs = std::max(max_pairs, pairs[i]);
    const int fused = ctx->outer_auto_ordered ? 3 : 4;
    method = (per_bucket >= 64.0 && W <= 4096 && (int64_t)max_pairs * W <= ((int64_t)1 << 28)) ? fused : 1;
  }
  if (method == 3 || method == 4) {
    // 32-bit byte offsets into B's values while B has fewer than 2^29
    // entries; SPWS_OUTER_WIDE_OFFSETS=1 takes the 64-bit addresses anyway
    const char* wide = getenv("SPWS_OUTER_WIDE_OFFSETS");
    const bool off32 = b_nnz < (1 << 29) && !(wide && wide[0] == '1');
    rc = outer_fused(ctx, a, pair_prefix, row_pairs, wlog, method == 3, off32, nnz_out);
  } else if (method == 2) {
    rc = outer_chunks_bucket(ctx, a, prefix, pair_prefix, row_pairs, wlog, max_products,
                             (int32_t)std::min<int64_t>(user_rows, INT32_MAX), nnz_out);
  } else {
    ctx->stats.outer_method = 1;
    const int kb = std::max(1, bits_for(a.n));
    int64_t max_rows = user_rows;
    if (kb <= 24) {
      max_rows = std::min<int64_t>(max_rows, (int64_t)1 << (32 - kb));
      rc = outer_chunks<uint32_t>(ctx, a, prefix, kb, max_products, (int32_t)max_rows, nnz_out);
    } else {
      max_rows = std::min<int64_t>(max_rows, (int64_t)1 << std::min(31, 63 - kb));
      rc = outer_chunks<uint64_t>(ctx, a, prefix, kb, max_products, (int32_t)max_rows, nnz_out);
    }
  }
  if (rc) return rc;
  ctx->stats.nnz = *nnz_out;
  return SPWS_OK;
}

}  // namespace spws

using namespace spws;

extern "C" {

int spws_ctx_set_outer_chunking(spws_ctx* ctx, int64_t max_products, int32_t max_rows) {
  SPWS_REQUIRE(ctx != nullptr, SPWS_ERR_INVALID, "null context");
  SPWS_REQUIRE(max_products >= 0 && max_rows >= 0, SPWS_ERR_INVALID, "negative chunk limit");
  ctx->outer_max_products = max_products;
  ctx->outer_max_rows = max_rows;
  return SPWS_OK;
}

int spws_ctx_set_outer_method(spws_ctx* ctx, int32_t method, int32_t window_log2) {
  SPWS_REQUIRE(ctx != nullptr, SPWS_ERR_INVALID, "null context");
  SPWS_REQUIRE(method >= 0 && method <= 4, SPWS_ERR_INVALID, "outer method must be 0, 1, 2, 3 or 4");
  SPWS_REQUIRE(window_log2 == 0 || (window_log2 >= kBkMinLog2 && window_log2 <= kBkMaxLog2), SPWS_ERR_INVALID,
               "window_log2 must be 0 or in [%d, %d]", kBkMinLog2, kBkMaxLog2);
  ctx->outer_method = method;
  ctx->outer_window_log2 = window_log2;
  return SPWS_OK;
}

int spws_spgemm_outer(spws_ctx* ctx, int32_t m, int32_t k, int32_t n, int32_t a_ncols, const int32_t* A1_crd,
                      const int32_t* A2_pos, const int32_t* A2_crd, const double* A_vals, int32_t b_nrows,
                      const int32_t* B1_crd, const int32_t* B2_pos, const int32_t* B2_crd, const double* B_vals,
                      int64_t* C_nnz, int64_t** C_pos, int32_t** C_crd, double** C_vals) {
  OuterArgs a{m, k, n, a_ncols, A1_crd, A2_pos, A2_crd, A_vals, b_nrows, B1_crd, B2_pos, B2_crd, B_vals};
  int64_t nnz = 0;
  int rc = outer_run(ctx, a, &nnz);
  if (rc) return rc;
  if (C_nnz) *C_nnz = nnz;
  if (C_pos) *C_pos = ctx->out_pos.as<int64_t>();
  if (C_crd) *C_crd = ctx->out_crd.as<int32_t>();
  if (C_vals) *C_vals = ctx->out_val.as<double>();
  return SPWS_OK;
}

}  // extern "C"
