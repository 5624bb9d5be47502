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

__global__ __launch_bounds__(256) void k_outer_row_ub(int32_t a_ncols, const int32_t* __restrict__ A2_pos,
                                                      const int32_t* __restrict__ A2_crd,
                                                      const int32_t* __restrict__ blen,
                                                      unsigned long long* __restrict__ row_ub) {
  for (int32_t ja = blockIdx.x * blockDim.x + threadIdx.x; ja < a_ncols; ja += gridDim.x * blockDim.x) {
    const int32_t L = blen[ja];
    if (L == 0) continue;
    for (int32_t p = A2_pos[ja]; p < A2_pos[ja + 1]; ++p) atomicAdd(&row_ub[A2_crd[p]], (unsigned long long)L);
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
                                                       int64_t* __restrict__ act_cnt) {
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

static int bits_for(int64_t v) {  // bits to represent values in [0, v)
  int b = 0;
  while (b < 63 && ((int64_t)1 << b) < v) ++b;
  return b;
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
  // output: start at min(total products, what fits), grow when a chunk needs more
  const int64_t total_products = row_prefix[m];
  if ((rc = ctx->out_pos.ensure(((size_t)m + 1) * 8))) return rc;
  {
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
  }
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
                  match, blen, a.B2_pos, act_astart, act_L, act_bstart, act_cnt);
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
      // grow the output if this chunk does not fit
      const size_t have = ctx->out_crd.bytes / 4 < ctx->out_val.bytes / 8 ? ctx->out_crd.bytes / 4
                                                                          : ctx->out_val.bytes / 8;
      if ((size_t)(base + U) > have) {
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
      }
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
      (rc = ctx->row_ub.ensure(mm * 8)))
    return rc;
  int32_t* match = ctx->obuf[0].as<int32_t>();
  int32_t* blen = ctx->obuf[1].as<int32_t>();
  SPWS_HIP(hipMemsetAsync(ctx->row_ub.ptr, 0, mm * 8, s));
  if (a.a_ncols > 0) {
    const unsigned g = (unsigned)div_up(a.a_ncols, 256);
    SPWS_LAUNCH(ctx, "outer_match", k_outer_match, dim3(g), dim3(256), 0, a.a_ncols, a.A1_crd, a.b_nrows, a.B1_crd,
                a.B2_pos, match, blen);
    const unsigned g2 = g < 4096 ? g : 4096;
    SPWS_LAUNCH(ctx, "outer_row_ub", k_outer_row_ub, dim3(g2), dim3(256), 0, a.a_ncols, a.A2_pos, a.A2_crd, blen,
                (unsigned long long*)ctx->row_ub.ptr);
  }
  std::vector<int64_t> ub(mm), prefix((size_t)m + 1, 0);
  if (m > 0) {
    SPWS_HIP(hipMemcpyAsync(ub.data(), ctx->row_ub.ptr, (size_t)m * 8, hipMemcpyDeviceToHost, s));
    SPWS_HIP(hipStreamSynchronize(s));
  }
  for (int32_t i = 0; i < m; ++i) prefix[i + 1] = prefix[i] + ub[i];
  ctx->stats.products = prefix[m];
  const int kb = std::max(1, bits_for(a.n));
  const int64_t max_products = ctx->outer_max_products > 0 ? ctx->outer_max_products : ((int64_t)1 << 30);
  int64_t max_rows = ctx->outer_max_rows > 0 ? ctx->outer_max_rows : INT32_MAX;
  if (kb <= 24) {
    max_rows = std::min<int64_t>(max_rows, (int64_t)1 << (32 - kb));
    rc = outer_chunks<uint32_t>(ctx, a, prefix, kb, max_products, (int32_t)max_rows, nnz_out);
  } else {
    max_rows = std::min<int64_t>(max_rows, (int64_t)1 << std::min(31, 63 - kb));
    rc = outer_chunks<uint64_t>(ctx, a, prefix, kb, max_products, (int32_t)max_rows, nnz_out);
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
