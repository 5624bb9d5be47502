// Sparse tensor kernels over 3-level CSF inputs: MTTKRP and TTM.
//
// MTTKRP  M(i,r) = sum_{j,k} X(i,j,k) * Bf(j,r) * Cf(k,r)   (output mode i)
//   Input CSF levels (j, k, i): the output mode is the leaf, as in the
//   reference (B(k,l,i) in benchmark/MTTKRP/COO_CSF_DCSR_DCSR_hash_free.h:233;
//   its C and D factors are Bf and Cf here, given dense, row-major R wide).
//   Bucket workspace over output rows:
//     1. k_seg_tile_first + k_seg_owner   fiber -> root owners, then the
//                           leaves' owner walk writes each nonzero's packed
//                           key (i << (bj+bk) | j << bk | k) directly
//                           (CsfKeyOut); with factor masks, leaf -> fiber
//                           owners and k_mttkrp_keys instead.
//     2. radix sort         stable on the i bits only: the buckets are the
//                           output rows, CSF (j,k) order kept inside each.
//     3. k_mttkrp_accum     each half-wavefront (32 lanes = 32 ranks) walks a
//                           segment of the bucketed nonzeros, loading factor
//                           rows Bf(j,:), Cf(k,:) as coalesced 256-byte rows
//                           (Bf(j,:) only when j changes: within a row the
//                           nonzeros come in (j, k) order) and accumulating
//                           the rank-R fiber in registers;
//                           a fiber is added into M when its row changes (only
//                           segment-boundary rows are shared, so the atomics
//                           are few and uncontended).
//     4. k_rows_present + scan + k_rows_compact: the sorted list of rows that
//                           hold a nonzero (the reference's COO output rows).
//   The product is (X * Bf) * Cf, the reference's B_vals * C_vals * D_vals
//   (:296), and each segment sums in CSF order.
//
// TTM  Y(i,j,r) = sum_k X(i,k,j) * U(k,r)   (contracted mode k)
//   Input CSF levels (i, k, j): the contracted coordinate is the middle
//   level, as in the reference's hash_il (B(k,i,j) in
//   benchmark/TTM/COO_CSF_DCSR_hash_il.h:247), whose hash workspace is keyed
//   on the output fiber (i,j).  Here the fibers are grouped by sorting:
//     1. k_seg_tile_first + k_seg_owner   fiber -> root owners, then the
//                           leaves' owner walk writes each nonzero's packed
//                           (i, j, k) key directly (CsfKeyOut); with a U row
//                           mask, leaf -> fiber owners and k_ttm_direct_keys
//                           instead.  The sort reads the values straight
//                           from the CSF.
//     2. radix sort         stable, on the (i, j) bits only: each fiber's
//                           nonzeros end up together and keep the CSF order,
//                           increasing k -- the order hash_il sums them.
//     3. k_ttm_fiber_count + scan of the tile counts + k_ttm_fiber_starts:
//                           fiber heads ranked within their tile, fiber
//                           offsets and the output coordinates (i, j).
//     4. k_ttm_accum        one 16-lane group per work item (at most 128
//                           nonzeros of one fiber) sums X * U(k,:) in order;
//                           a fiber of one item writes its vector, longer
//                           fibers write partial sums that k_ttm_combine adds
//                           in item order (no atomics; deterministic).
//   When i, j and k need more than 64 bits together, the fibers are found
//   through an open-addressing table in HBM (k_ttm_insert, 64-bit keys
//   (i << bj) | j, atomicCAS), collected and sorted (k_ttm_occupied,
//   k_ttm_collect, radix sort, k_ttm_rank), and the nonzeros are then
//   sorted by fiber rank (k_ttm_keys, k_ttm_heads) before step 5.
//   Output: semi-sparse -- the sorted fiber list (i,j) with a dense length-R
//   vector each.
#include <algorithm>
#include <climits>

#include "radix.hpp"
#include "tensor.hpp"

namespace spws {

constexpr int kOwnBlock = 256;
constexpr int kOwnItems = 8;
constexpr int kOwnTile = kOwnBlock * kOwnItems;

// owner[e] = s for e in [pos[s], pos[s+1]), s in [0, nseg); pos[nseg] = nitems.
// First the owner of every tile's first item (one thread per tile, a binary
// search over pos: the last s with pos[s] <= e, which is the non-empty
// segment holding e); then one tile of items per workgroup, each item's
// owner by a binary search in the LDS slice of pos between its tile's first
// owner and the next tile's (the owner of the tile's last item is at most
// the latter).
__global__ __launch_bounds__(256) void k_seg_tile_first(const int32_t* __restrict__ pos, int32_t nseg, int64_t nitems,
                                                        int64_t ntiles, int32_t* __restrict__ first) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  if (t == ntiles) {
    first[t] = nseg - 1;
    return;
  }
  const int64_t x = t * kOwnTile;
  int32_t lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (pos[mid] <= x) lo = mid; else hi = mid - 1;
  }
  first[t] = lo;
}

// What k_seg_owner does with item e of segment s: store s.
struct OwnerOut {
  int32_t* __restrict__ owner;
  __device__ void operator()(int64_t e, int32_t s) const { owner[e] = s; }
};

// Or, over a CSF's leaves (s = the nonzero's fiber): its packed sort key
// (hi << (bj + bk)) | (mid << bk) | k with k = the fiber's coordinate and,
// for TTM, hi = the root's and mid = the leaf's coordinate; for MTTKRP the
// other way round.
template <bool kMttkrp>
struct CsfKeyOut {
  const int32_t* __restrict__ fiber_root;
  const int32_t* __restrict__ l1_crd;
  const int32_t* __restrict__ l2_crd;
  const int32_t* __restrict__ l3_crd;
  int bj, bk;
  uint64_t* __restrict__ keys;
  __device__ void operator()(int64_t e, int32_t f) const {
    const uint64_t root = (uint32_t)l1_crd[fiber_root[f]];
    const uint64_t k = (uint32_t)l2_crd[f];
    const uint64_t leaf = (uint32_t)l3_crd[e];
    const uint64_t hi = kMttkrp ? leaf : root, mid = kMttkrp ? root : leaf;
    keys[e] = (hi << (bj + bk)) | (mid << bk) | k;
  }
};

template <class Out>
__global__ __launch_bounds__(kOwnBlock) void k_seg_owner(const int32_t* __restrict__ pos, int64_t nitems,
                                                         const int32_t* __restrict__ first, Out out) {
  __shared__ int32_t spos[kOwnTile + 1];
  const int64_t e0 = (int64_t)blockIdx.x * kOwnTile;
  const int64_t e_end = e0 + kOwnTile < nitems ? e0 + kOwnTile : nitems;
  const int32_t lo0 = first[blockIdx.x];
  const int32_t span = first[blockIdx.x + 1] - lo0 + 1;  // candidate owners lo0 .. lo0 + span - 1
  const bool fits = span <= kOwnTile;
  if (fits)
    for (int i = threadIdx.x; i < span; i += kOwnBlock) spos[i] = pos[lo0 + i];
  __syncthreads();
  for (int it = 0; it < kOwnItems; ++it) {
    const int64_t e = e0 + (int64_t)it * kOwnBlock + threadIdx.x;
    if (e >= e_end) break;
    int32_t lo = 0, hi = span - 1;
    if (fits) {
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (spos[mid] <= e) lo = mid; else hi = mid - 1;
      }
    } else {
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (pos[lo0 + mid] <= e) lo = mid; else hi = mid - 1;
      }
    }
    out(e, lo0 + lo);
  }
}

// first: ntiles + 1 entries of scratch (ntiles = tiles of kOwnTile items).
template <class Out>
static int seg_owner(spws_ctx* ctx, const int32_t* pos, int32_t nseg, int64_t nitems, Out out, int32_t* first,
                     const char* name = "seg_owner") {
  if (nitems <= 0) return SPWS_OK;
  const int64_t ntiles = div_up(nitems, kOwnTile);
  SPWS_LAUNCH(ctx, "seg_owner", k_seg_tile_first, dim3((unsigned)div_up(ntiles + 1, 256)), dim3(256), 0, pos, nseg,
              nitems, ntiles, first);
  SPWS_LAUNCH(ctx, name, k_seg_owner<Out>, dim3((unsigned)ntiles), dim3(kOwnBlock), 0, pos, nitems, first, out);
  return SPWS_OK;
}

static int bits_of(int64_t v) {  // bits for values in [0, v)
  int b = 0;
  while (b < 63 && ((int64_t)1 << b) < v) ++b;
  return b;
}

// ------------------------------------------------------------------ MTTKRP

// With factor row masks (bit r: column r of the factor row is stored), a
// nonzero whose two rows share no stored column takes no part (the
// reference's co-iteration never reaches it): it gets the row key `drop_i`,
// past every output row, so the sort moves it to the end, and is counted.
template <bool kMask>
__global__ __launch_bounds__(256) void k_mttkrp_keys(int64_t nnz, const int32_t* __restrict__ leaf_fiber,
                                                     const int32_t* __restrict__ fiber_root,
                                                     const int32_t* __restrict__ l1_crd,
                                                     const int32_t* __restrict__ l2_crd,
                                                     const int32_t* __restrict__ l3_crd, int bj, int bk,
                                                     uint64_t* __restrict__ keys, const uint64_t* __restrict__ mB,
                                                     const uint64_t* __restrict__ mC, uint64_t drop_i,
                                                     unsigned long long* __restrict__ dropped) {
  unsigned long long nd = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t f = leaf_fiber[e];
    const uint64_t j = (uint32_t)l1_crd[fiber_root[f]];
    const uint64_t k = (uint32_t)l2_crd[f];
    uint64_t i = (uint32_t)l3_crd[e];
    if (kMask && (mB[j] & mC[k]) == 0) {
      i = drop_i;
      ++nd;
    }
    keys[e] = (i << (bj + bk)) | (j << bk) | k;
  }
  if (kMask) {
    nd = wave_sum(nd);
    if (lane_id() == 0 && nd) atomicAdd(dropped, nd);
  }
}

// Each half-wavefront owns `seg` consecutive bucketed nonzeros.  Lanes cover
// ranks r = lane32, lane32 + 32, ... (R <= 64 supported with two slices).
// The half-wavefront loads 32 keys and values at once (coalesced), then
// walks them kAccU at a time: the factor rows of kAccU nonzeros are loaded
// before any is used, so their latencies overlap; the sums still run in
// bucket order.
constexpr int kAccBlock = 256;
constexpr int kAccU = 4;

// kMask: per row, the OR of its nonzeros' shared stored factor columns
// (mB[j] & mC[k]) is or-ed into rmask[i], the row's stored output columns.
template <bool kTwo, bool kMask>
__global__ __launch_bounds__(kAccBlock) void k_mttkrp_accum(int64_t nnz, int64_t seg, const uint64_t* __restrict__ keys,
                                                            const double* __restrict__ vals, int bj, int bk, int R,
                                                            const double* __restrict__ Bf,
                                                            const double* __restrict__ Cf, double* __restrict__ M,
                                                            int32_t* __restrict__ present,
                                                            const uint64_t* __restrict__ mB,
                                                            const uint64_t* __restrict__ mC,
                                                            unsigned long long* __restrict__ rmask) {
  const int half = threadIdx.x >> 5;
  const int r0 = threadIdx.x & 31;
  const int64_t hid = (int64_t)blockIdx.x * (kAccBlock / 32) + half;
  const int64_t e0 = hid * seg;
  if (e0 >= nnz) return;  // uniform per half-wavefront
  const int64_t e1 = e0 + seg < nnz ? e0 + seg : nnz;
  const uint64_t jm = (1ull << bj) - 1, km = (1ull << bk) - 1;
  constexpr bool two = kTwo;  // R > 32: a second slice of ranks per lane
  const int r1 = r0 + 32;
  const int rr0 = r0 < R ? r0 : 0, rr1 = r1 < R ? r1 : 0;  // in-range rank for the loads
  int64_t cur = -1;
  double acc0 = 0.0, acc1 = 0.0;
  uint64_t rowm = 0;
  // Bf(j,:) of the last nonzero: within a row the nonzeros come in (j, k)
  // order, so runs of equal j reuse it instead of gathering it again
  int64_t jprev = -1;
  double bp0 = 0.0, bp1 = 0.0;
  for (int64_t base = e0; base < e1; base += 32) {
    const int cnt = (int)(e1 - base < 32 ? e1 - base : 32);
    const int64_t me = base + (r0 < cnt ? r0 : cnt - 1);
    const uint64_t mykey = keys[me];
    const double myval = vals[me];
    for (int t = 0; t < cnt; t += kAccU) {
      uint64_t kk[kAccU];
      double xv[kAccU], b0[kAccU], c0[kAccU], b1[kAccU], c1[kAccU];
#pragma unroll
      for (int u = 0; u < kAccU; ++u) {
        const int src = t + u < cnt ? t + u : cnt - 1;
        kk[u] = __shfl(mykey, src, 32);
        xv[u] = __shfl(myval, src, 32);
      }
      bool nb[kAccU];  // Bf(j,:) differs from the previous nonzero's (uniform per half-wavefront)
#pragma unroll
      for (int u = 0; u < kAccU; ++u) {
        const int64_t j = (int64_t)((kk[u] >> bk) & jm);
        const int64_t k = (int64_t)(kk[u] & km);
        const int64_t jbefore = u == 0 ? jprev : (int64_t)((kk[u - 1] >> bk) & jm);
        nb[u] = j != jbefore;
        b0[u] = b1[u] = 0.0;
        if (nb[u]) {
          b0[u] = Bf[j * R + rr0];
          if (two) b1[u] = Bf[j * R + rr1];
        }
        c0[u] = Cf[k * R + rr0];
        if (two) c1[u] = Cf[k * R + rr1];
      }
      jprev = (int64_t)((kk[kAccU - 1] >> bk) & jm);
#pragma unroll
      for (int u = 0; u < kAccU; ++u) {
        if (nb[u]) {
          bp0 = b0[u];
          if (two) bp1 = b1[u];
        }
        b0[u] = bp0;
        if (two) b1[u] = bp1;
      }
#pragma unroll
      for (int u = 0; u < kAccU; ++u) {
        if (t + u >= cnt) break;  // uniform per half-wavefront
        const int64_t i = (int64_t)(kk[u] >> (bj + bk));
        if (i != cur) {
          if (cur >= 0) {
            if (r0 < R) atomicAdd(&M[cur * R + r0], acc0);
            if (two && r1 < R) atomicAdd(&M[cur * R + r1], acc1);
            if (r0 == 0) present[cur] = 1;
            if (kMask && r0 == 0) atomicOr(&rmask[cur], (unsigned long long)rowm);
          }
          cur = i;
          acc0 = acc1 = 0.0;
          rowm = 0;
        }
        if (kMask) rowm |= mB[(kk[u] >> bk) & jm] & mC[kk[u] & km];
        acc0 += (xv[u] * b0[u]) * c0[u];
        if (two) acc1 += (xv[u] * b1[u]) * c1[u];
      }
    }
  }
  if (cur >= 0) {
    if (r0 < R) atomicAdd(&M[cur * R + r0], acc0);
    if (two && r1 < R) atomicAdd(&M[cur * R + r1], acc1);
    if (r0 == 0) present[cur] = 1;
    if (kMask && r0 == 0) atomicOr(&rmask[cur], (unsigned long long)rowm);
  }
}

__global__ __launch_bounds__(256) void k_rows_compact(int32_t n, const int32_t* __restrict__ present,
                                                      const int32_t* __restrict__ off, int32_t* __restrict__ rows,
                                                      const unsigned long long* __restrict__ rmask,
                                                      uint64_t* __restrict__ rows_mask) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && present[i]) {
    rows[off[i]] = i;
    if (rmask) rows_mask[off[i]] = rmask[i];
  }
}

// ------------------------------------------------------------------ TTM

constexpr uint64_t kEmptyKey = ~0ull;

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 31;
  x *= 0x7fb5d329728ea185ull;
  x ^= x >> 27;
  x *= 0x81dadef4bc2dd44dull;
  x ^= x >> 33;
  return x;
}

__global__ __launch_bounds__(256) void k_ttm_insert(int64_t nnz, const int32_t* __restrict__ leaf_fiber,
                                                    const int32_t* __restrict__ fiber_root,
                                                    const int32_t* __restrict__ l1_crd,
                                                    const int32_t* __restrict__ l3_crd, int bj,
                                                    unsigned long long* __restrict__ table, uint64_t mask,
                                                    int32_t* __restrict__ leaf_slot) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t i = (uint32_t)l1_crd[fiber_root[leaf_fiber[e]]];
    const uint64_t j = (uint32_t)l3_crd[e];
    const uint64_t key = (i << bj) | j;
    uint64_t h = mix64(key) & mask;
    for (;;) {
      unsigned long long cur = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cur == kEmptyKey) {
        cur = atomicCAS(&table[h], (unsigned long long)kEmptyKey, (unsigned long long)key);
        if (cur == kEmptyKey) cur = key;
      }
      if (cur == key) break;
      h = (h + 1) & mask;
    }
    leaf_slot[e] = (int32_t)h;
  }
}

__global__ __launch_bounds__(256) void k_ttm_occupied(int64_t T, const unsigned long long* __restrict__ table,
                                                      int32_t* __restrict__ flag) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < T; s += (int64_t)gridDim.x * blockDim.x)
    flag[s] = table[s] != kEmptyKey ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_ttm_collect(int64_t T, const unsigned long long* __restrict__ table,
                                                     const int32_t* __restrict__ off, uint64_t* __restrict__ keys,
                                                     double* __restrict__ slot_payload) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < T; s += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long key = table[s];
    if (key != kEmptyKey) {
      const int32_t q = off[s];
      keys[q] = key;
      slot_payload[q] = __longlong_as_double((long long)s);
    }
  }
}

__global__ __launch_bounds__(256) void k_ttm_rank(int64_t F, const uint64_t* __restrict__ keys,
                                                  const double* __restrict__ slot_payload, int bj,
                                                  int32_t* __restrict__ rank_of_slot, int32_t* __restrict__ out_i,
                                                  int32_t* __restrict__ out_j) {
  const uint64_t jm = (1ull << bj) - 1;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < F; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = (int64_t)__double_as_longlong(slot_payload[q]);
    rank_of_slot[s] = (int32_t)q;
    out_i[q] = (int32_t)(keys[q] >> bj);
    out_j[q] = (int32_t)(keys[q] & jm);
  }
}

// Nonzeros keyed by their fiber's rank and their contracted coordinate:
// key = q << bk | k, value = X.  A stable sort on the q bits groups each
// fiber's nonzeros, still in CSF order (increasing k) inside the fiber, the
// order in which the reference's hash workspace accumulates them.
__global__ __launch_bounds__(256) void k_ttm_keys(int64_t nnz, const int32_t* __restrict__ leaf_fiber,
                                                  const int32_t* __restrict__ l2_crd,
                                                  const double* __restrict__ xvals,
                                                  const int32_t* __restrict__ leaf_slot,
                                                  const int32_t* __restrict__ rank_of_slot, int bk,
                                                  uint64_t* __restrict__ keys, double* __restrict__ vals) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t q = (uint32_t)rank_of_slot[leaf_slot[e]];
    const uint64_t k = (uint32_t)l2_crd[leaf_fiber[e]];
    keys[e] = (q << bk) | k;
    vals[e] = xvals[e];
  }
}

// foff[q] = first sorted position of fiber q (every fiber has a nonzero).
__global__ __launch_bounds__(256) void k_ttm_heads(int64_t nnz, const uint64_t* __restrict__ keys, int bk, int64_t F,
                                                   int32_t* __restrict__ foff) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t q = keys[p] >> bk;
    if (p == 0 || (keys[p - 1] >> bk) != q) foff[q] = (int32_t)p;
    if (p == 0) foff[F] = (int32_t)nnz;
  }
}

// Work items of at most kTtmChunk sorted nonzeros: a fiber of n nonzeros
// has ceil(n / kTtmChunk) items (fibers reach 2,000 nonzeros, one per k).
constexpr int kTtmChunk = 128;
constexpr int kTtmLanes = 16;  // one DPP row per group (broadcasts by row_newbcast)
constexpr int kTtmU = 16;
constexpr int kTtmBlocksPerCu = 5;  // workgroups per CU in the accumulation grid (5 waves per SIMD: 96 VGPRs)

__global__ __launch_bounds__(256) void k_ttm_chunks(int64_t F, const int32_t* __restrict__ foff,
                                                    int32_t* __restrict__ nchunk) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < F; q += (int64_t)gridDim.x * blockDim.x)
    nchunk[q] = (foff[q + 1] - foff[q] + kTtmChunk - 1) / kTtmChunk;
}

// Work item records {first nonzero, end, destination}: the destination is
// the fiber q for a fiber of one item (its vector goes to Y), else ~c for
// item c (its partial sum goes to the partial-sum rows).
__global__ __launch_bounds__(256) void k_ttm_itemmap(int64_t F, const int32_t* __restrict__ foff,
                                                     const int32_t* __restrict__ coff, int4* __restrict__ items) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < F; q += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c0 = coff[q], c1 = coff[q + 1], f1 = foff[q + 1];
    for (int32_t c = c0; c < c1; ++c) {
      const int32_t p0 = foff[q] + (c - c0) * kTtmChunk;
      items[c] = make_int4(p0, min(p0 + kTtmChunk, f1), c1 - c0 == 1 ? (int32_t)q : ~c, 0);
    }
  }
}

// acc + the in-order sum over nonzeros [base, p1) of an item, loaded
// kTtmLanes at a time (only items of more than kTtmLanes nonzeros take it).
template <bool kOff32>
__device__ __forceinline__ double ttm_load_u(const double* __restrict__ U, uint32_t k, int R, int rr) {
  // with kOff32 (K * R * 8 < 2^32) the byte offset is a 32-bit value: the
  // load takes U's address from scalar registers and a one-register offset
  if (kOff32)
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(U) +
                                            (k * (uint32_t)R + (uint32_t)rr) * 8u);
  return U[(int64_t)k * R + rr];
}

template <bool kOff32>
__device__ __forceinline__ double ttm_item_tail(double acc, int32_t base, int32_t p1, int l, int rr, int R, uint64_t km,
                                             const uint64_t* __restrict__ keys, const double* __restrict__ vals,
                                             const double* __restrict__ U) {
  for (; base < p1; base += kTtmLanes) {
    const int cnt2 = p1 - base < kTtmLanes ? p1 - base : kTtmLanes;
    const int32_t m2 = base + (l < cnt2 ? l : cnt2 - 1);
    const uint32_t k2 = (uint32_t)(keys[m2] & km);
    const double v2 = vals[m2];
    double uv[kTtmU];
#pragma unroll
    for (int u = 0; u < kTtmU; ++u) {
      const uint32_t k = row_lane(k2, u);  // lanes past cnt2 hold the last entry
      uv[u] = ttm_load_u<kOff32>(U, k, R, rr);
    }
#pragma unroll
    for (int u = 0; u < kTtmU; ++u) {
      const double x = row_lane(v2, u);
      if (u < cnt2) acc += x * uv[u];
    }
  }
  return acc;
}

// One group of kTtmLanes lanes per work item (ranks r = lane, lane +
// kTtmLanes, ...), each group walking items c, c + G, ... of a grid of G
// groups.  The group holds its item's first kTtmLanes sorted keys/values (a
// lane each), gathers the rows U(k,:) of all of them, then sums in order;
// longer items load their further nonzeros kTtmLanes at a time.  While an
// item's rows of U are in flight, the next item's keys/values and the record
// after it load too, so each item costs about one memory latency instead of
// three dependent ones.  A single-item fiber writes its vector; the items of
// a longer fiber write partial sums, which k_ttm_combine adds in item order
// (so the result does not depend on timing).
template <bool kOff32>
__global__ __launch_bounds__(256) void k_ttm_accum(const int32_t* __restrict__ nitems_p, int R,
                                                   const int4* __restrict__ items,
                                                   const uint64_t* __restrict__ keys, const double* __restrict__ vals,
                                                   int bk, const double* __restrict__ U, double* __restrict__ Y,
                                                   double* __restrict__ part) {
  const int groups = 256 / kTtmLanes;
  const int l = threadIdx.x % kTtmLanes;
  const uint64_t km = (1ull << bk) - 1;
  const uint32_t km32 = (uint32_t)km;  // k < 2^31: the key's low word holds it
  const uint32_t* __restrict__ keys_lo = reinterpret_cast<const uint32_t*>(keys);  // little-endian low words
  const int64_t nitems = *nitems_p;  // the items' count, on the device (the grid is sized for at most that many)
  const int64_t stride = (int64_t)gridDim.x * groups;
  int64_t c = (int64_t)blockIdx.x * groups + threadIdx.x / kTtmLanes;
  if (c >= nitems) return;  // uniform per group
  // current item: record and first keys/values; next item: record
  int4 cur = items[c];
  int4 nxt = make_int4(0, 1, 0, 0);
  if (c + stride < nitems) nxt = items[c + stride];
  int32_t me = cur.x + min(l, cur.y - cur.x - 1);
  uint32_t kraw = keys_lo[2 * (int64_t)me];
  double myval = vals[me];
  for (;;) {
    const int64_t cn = c + stride;
    const bool has_next = cn < nitems;  // uniform per group
    // prefetches, left unused until the next iteration (no wait here)
    uint32_t nkraw = 0;
    double nval = 0.0;
    int4 nn = nxt;
    if (has_next) {
      const int32_t nme = nxt.x + min(l, nxt.y - nxt.x - 1);
      nkraw = keys_lo[2 * (int64_t)nme];
      nval = vals[nme];
      if (cn + stride < nitems) nn = items[cn + stride];
    }
    const uint32_t myk = kraw & km32;
    const int32_t p0 = cur.x, p1 = cur.y;
    const int cnt = min(p1 - p0, kTtmLanes);
    for (int r0 = 0; r0 < R; r0 += kTtmLanes) {
      const int r = r0 + l;
      const int rr = r < R ? r : R - 1;
      double acc = 0.0;
      {
        double uv[kTtmU];
#pragma unroll
        for (int u = 0; u < kTtmU; ++u) {
          const uint32_t k = row_lane(myk, u);  // lanes past cnt hold the last entry
          uv[u] = ttm_load_u<kOff32>(U, k, R, rr);
        }
#pragma unroll
        for (int u = 0; u < kTtmU; ++u) {
          const double x = row_lane(myval, u);
          if (u < cnt) acc += x * uv[u];
        }
      }
      // the item's further nonzeros (items of more than kTtmLanes)
      if (p1 - p0 > kTtmLanes) acc = ttm_item_tail<kOff32>(acc, p0 + kTtmLanes, p1, l, rr, R, km, keys, vals, U);
      if (r < R) {
        if (cur.z >= 0) Y[(int64_t)cur.z * R + r] = acc;
        else part[(int64_t)(~cur.z) * R + r] = acc;
      }
    }
    if (!has_next) break;
    c = cn;
    cur = nxt;
    nxt = nn;
    kraw = nkraw;
    myval = nval;
  }
}

// Fibers of several items: Y(q,:) = sum of their partial sums in item order.
__global__ __launch_bounds__(256) void k_ttm_combine(int64_t F, int R, const int32_t* __restrict__ coff,
                                                     const double* __restrict__ part, double* __restrict__ Y) {
  const int groups = 256 / kTtmLanes;
  const int g = threadIdx.x / kTtmLanes;
  const int l = threadIdx.x % kTtmLanes;
  for (int64_t q = (int64_t)blockIdx.x * groups + g; q < F; q += (int64_t)gridDim.x * groups) {
    const int32_t c0 = coff[q], c1 = coff[q + 1];
    if (c1 - c0 <= 1) continue;
    for (int r = l; r < R; r += kTtmLanes) {
      double acc = part[(int64_t)c0 * R + r];
      for (int32_t c = c0 + 1; c < c1; ++c) acc += part[(int64_t)c * R + r];
      Y[q * R + r] = acc;
    }
  }
}

// ------------------------------------------------------------------ drivers

// Direct TTM grouping: the nonzeros carry packed (i, j, k) keys, and one
// stable radix sort on the (i, j) bits puts each output fiber together, in
// (i, j) order, with its nonzeros in increasing k (the CSF order (i, k, j)
// has them so within a root) -- the order hash_il accumulates them
// (COO_CSF_DCSR_hash_il.h:247-377).
// With a U row mask (bit r: U(k, r) stored), a nonzero whose U row stores
// nothing takes no part (the reference's co-iteration of the contracted
// level with U's rows skips it): it gets the fiber key of row `drop_i`, past
// every root, so the sort moves it to the end, and is counted.
template <bool kMask>
__global__ __launch_bounds__(256) void k_ttm_direct_keys(int64_t nnz, const int32_t* __restrict__ leaf_fiber,
                                                         const int32_t* __restrict__ fiber_root,
                                                         const int32_t* __restrict__ l1_crd,
                                                         const int32_t* __restrict__ l2_crd,
                                                         const int32_t* __restrict__ l3_crd, int bj, int bk,
                                                         uint64_t* __restrict__ keys,
                                                         const uint64_t* __restrict__ umask, uint64_t drop_i,
                                                         unsigned long long* __restrict__ dropped) {
  unsigned long long nd = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t f = leaf_fiber[e];
    uint64_t i = (uint32_t)l1_crd[fiber_root[f]];
    const uint64_t k = (uint32_t)l2_crd[f];
    const uint64_t j = (uint32_t)l3_crd[e];
    if (kMask && umask[k] == 0) {
      i = drop_i;
      ++nd;
    }
    keys[e] = (i << (bj + bk)) | (j << bk) | k;
  }
  if (kMask) {
    nd = wave_sum(nd);
    if (lane_id() == 0 && nd) atomicAdd(dropped, nd);
  }
}

// per fiber: the OR of its nonzeros' U row masks (its stored output columns)
__global__ __launch_bounds__(256) void k_ttm_fiber_mask(int64_t F, const int32_t* __restrict__ foff,
                                                        const uint64_t* __restrict__ keys, int bk,
                                                        const uint64_t* __restrict__ umask, uint64_t* __restrict__ fmask) {
  const uint64_t km = (1ull << bk) - 1;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < F; f += (int64_t)gridDim.x * blockDim.x) {
    uint64_t m = 0;
    for (int32_t p = foff[f]; p < foff[f + 1]; ++p) m |= umask[keys[p] & km];
    fmask[f] = m;
  }
}

// Fiber starts of the sorted nonzeros (p starts a fiber when its (i, j)
// differs from p - 1's), in tiles of kFibTile: a count per tile, then (with
// the counts' prefix) each tile's starts ranked in order, with no flag or
// rank array per nonzero.  A tile is walked in kFibItems coalesced rounds of
// kFibBlock nonzeros; ballots give the ranks within a round's wavefronts.
constexpr int kFibBlock = 256;
constexpr int kFibItems = 16;
constexpr int kFibTile = kFibBlock * kFibItems;
constexpr int kFibWaves = kFibBlock / kWave;

// Whether nonzero p starts a fiber (false past the end).
__device__ __forceinline__ bool ttm_fiber_start(int64_t nnz, const uint64_t* __restrict__ keys, int bk, int64_t p) {
  if (p >= nnz) return false;
  return p == 0 || (keys[p - 1] >> bk) != (keys[p] >> bk);
}

__global__ __launch_bounds__(kFibBlock) void k_ttm_fiber_count(int64_t nnz, const uint64_t* __restrict__ keys, int bk,
                                                                int32_t* __restrict__ tile_cnt) {
  __shared__ int32_t wsum[kFibWaves];
  const int64_t base = (int64_t)blockIdx.x * kFibTile + threadIdx.x;
  int32_t c = 0;
#pragma unroll
  for (int it = 0; it < kFibItems; ++it) c += __popcll(__ballot(ttm_fiber_start(nnz, keys, bk, base + it * kFibBlock)));
  if (lane_id() == 0) wsum[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t t = 0;
#pragma unroll
    for (int w = 0; w < kFibWaves; ++w) t += wsum[w];
    tile_cnt[blockIdx.x] = t;
  }
}

// tile_off = exclusive prefix of the tile counts (tile_off[ntiles] = F):
// fiber q starts at foff[q]; its coordinates come from the key.
__global__ __launch_bounds__(kFibBlock) void k_ttm_fiber_starts(int64_t nnz, const uint64_t* __restrict__ keys, int bj,
                                                                int bk, const int32_t* __restrict__ tile_off,
                                                                int64_t ntiles, int32_t* __restrict__ foff,
                                                                int32_t* __restrict__ out_i,
                                                                int32_t* __restrict__ out_j) {
  // starts per (round, wavefront), then their exclusive prefix in that order
  __shared__ int32_t cnt[kFibItems * kFibWaves];
  static_assert(kFibItems * kFibWaves == kWave, "one wavefront scans the per-round counts");
  const uint64_t jm = (1ull << bj) - 1;
  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  const int64_t base = (int64_t)blockIdx.x * kFibTile + threadIdx.x;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  uint32_t mine = 0;  // bit it: this thread's nonzero of round it starts a fiber
  int32_t before[kFibItems];  // starts of this wavefront's lanes below this one, per round
#pragma unroll
  for (int it = 0; it < kFibItems; ++it) {
    const bool st = ttm_fiber_start(nnz, keys, bk, base + it * kFibBlock);
    const uint64_t m = __ballot(st);
    mine |= (st ? 1u : 0u) << it;
    before[it] = __popcll(m & below);
    if (lane == 0) cnt[it * kFibWaves + w] = __popcll(m);
  }
  __syncthreads();
  if (w == 0) {
    const int32_t v = cnt[lane];
    cnt[lane] = wave_incl_scan(v) - v;
  }
  __syncthreads();
  const int32_t t0 = tile_off[blockIdx.x];
#pragma unroll
  for (int it = 0; it < kFibItems; ++it) {
    if ((mine >> it) & 1u) {
      const int64_t p = base + it * kFibBlock;
      const int32_t q = t0 + cnt[it * kFibWaves + w] + before[it];
      const uint64_t key = keys[p];
      foff[q] = (int32_t)p;
      out_i[q] = (int32_t)(key >> (bj + bk));
      out_j[q] = (int32_t)((key >> bk) & jm);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) foff[tile_off[ntiles]] = (int32_t)nnz;
}

static int csf_owners(spws_ctx* ctx, const Csf3& x, int32_t** leaf_fiber, int32_t** fiber_root) {
  int rc;
  // obuf[0]: the nonzeros' fibers, then the tile owners both passes use in
  // turn (nfibers <= nnz)
  const size_t nz = (size_t)(x.nnz > 0 ? x.nnz : 1);
  const size_t tiles = (size_t)div_up((int64_t)nz, kOwnTile) + 1;
  if ((rc = ctx->obuf[0].ensure((nz + tiles) * 4)) ||
      (rc = ctx->obuf[1].ensure((size_t)(x.nfibers > 0 ? x.nfibers : 1) * 4)))
    return rc;
  *leaf_fiber = ctx->obuf[0].as<int32_t>();
  *fiber_root = ctx->obuf[1].as<int32_t>();
  int32_t* first = *leaf_fiber + nz;
  if ((rc = seg_owner(ctx, x.l3_pos, x.nfibers, x.nnz, OwnerOut{*leaf_fiber}, first))) return rc;
  return seg_owner(ctx, x.l2_pos, x.nroots, x.nfibers, OwnerOut{*fiber_root}, first);
}

// The nonzeros' packed sort keys straight from the CSF (CsfKeyOut): the
// fibers' roots first, then the keys from the leaves' owner walk, with no
// per-nonzero fiber array.
template <bool kMttkrp>
static int csf_keys(spws_ctx* ctx, const Csf3& x, int bj, int bk, uint64_t* keys) {
  int rc;
  const size_t nz = (size_t)(x.nnz > 0 ? x.nnz : 1);
  const size_t tiles = (size_t)div_up((int64_t)nz, kOwnTile) + 1;
  if ((rc = ctx->obuf[0].ensure(tiles * 4)) ||
      (rc = ctx->obuf[1].ensure((size_t)(x.nfibers > 0 ? x.nfibers : 1) * 4)))
    return rc;
  int32_t* first = ctx->obuf[0].as<int32_t>();
  int32_t* fiber_root = ctx->obuf[1].as<int32_t>();
  if ((rc = seg_owner(ctx, x.l2_pos, x.nroots, x.nfibers, OwnerOut{fiber_root}, first))) return rc;
  return seg_owner(ctx, x.l3_pos, x.nfibers, x.nnz,
                   CsfKeyOut<kMttkrp>{fiber_root, x.l1_crd, x.l2_crd, x.l3_crd, bj, bk, keys}, first,
                   kMttkrp ? "mttkrp_keys" : "ttm_keys");
}

// mB / mC (device, J and K entries; both null or both set): per factor row,
// its stored columns as bits (the factors are given dense, zero where not
// stored).  With masks, the rows' stored output columns go to
// ctx->obuf[7] (one uint64 per listed row).
static int mttkrp_run(spws_ctx* ctx, const Csf3& x, int32_t I, int32_t J, int32_t K, int32_t R, const double* Bf,
                      const double* Cf, const uint64_t* mB, const uint64_t* mC, int32_t* nrows_out) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_REQUIRE(I >= 0 && J >= 0 && K >= 0 && R >= 1 && R <= 64, SPWS_ERR_INVALID,
               "MTTKRP needs non-negative dimensions and 1 <= R <= 64");
  SPWS_REQUIRE((mB == nullptr) == (mC == nullptr), SPWS_ERR_INVALID, "give both factor masks or neither");
  const bool masked = mB != nullptr;
  // with masks, row I is the key of the nonzeros that take no part
  const int bi = bits_of(masked ? (int64_t)I + 1 : (I > 1 ? I : 2)), bj = bits_of(J > 1 ? J : 2),
            bk = bits_of(K > 1 ? K : 2);
  SPWS_REQUIRE(bi + bj + bk <= 64, SPWS_ERR_UNSUPPORTED, "tensor dimensions exceed the 64-bit packed key");
  hipStream_t s = ctx->stream;
  memset(&ctx->stats, 0, sizeof(ctx->stats));
  const size_t nz = (size_t)(x.nnz > 0 ? x.nnz : 1);
  const size_t ii = (size_t)(I > 0 ? I : 1);
  if ((rc = ctx->out_val.ensure(ii * R * 8)) || (rc = ctx->out_aux.ensure(ii * 4 * 2 + 8)) ||
      (rc = ctx->out_crd.ensure(ii * 4)) || (rc = ctx->keys[0].ensure(nz * 8)) || (rc = ctx->keys[1].ensure(nz * 8)) ||
      (rc = ctx->vals[0].ensure(nz * 8)) || (rc = ctx->vals[1].ensure(nz * 8)))
    return rc;
  if (masked && ((rc = ctx->obuf[6].ensure(ii * 8 + 256)) || (rc = ctx->obuf[7].ensure(ii * 8)))) return rc;
  double* M = ctx->out_val.as<double>();
  int32_t* present = ctx->out_aux.as<int32_t>();
  int32_t* poff = present + ii;
  unsigned long long* rmask = masked ? ctx->obuf[6].as<unsigned long long>() : nullptr;
  unsigned long long* dropped = masked ? rmask + ii : nullptr;
  SPWS_HIP(hipMemsetAsync(M, 0, ii * R * 8, s));
  SPWS_HIP(hipMemsetAsync(present, 0, ii * 4, s));
  if (masked) SPWS_HIP(hipMemsetAsync(rmask, 0, ii * 8 + 8, s));
  int64_t kept = x.nnz;
  if (x.nnz > 0) {
    uint64_t* keys = ctx->keys[0].as<uint64_t>();
    double* vals = ctx->vals[0].as<double>();
    if (masked) {
      int32_t *leaf_fiber, *fiber_root;
      if ((rc = csf_owners(ctx, x, &leaf_fiber, &fiber_root))) return rc;
      const unsigned g = (unsigned)std::min<int64_t>(div_up(x.nnz, 256), 65536);
      SPWS_LAUNCH(ctx, "mttkrp_keys", k_mttkrp_keys<true>, dim3(g), dim3(256), 0, x.nnz, leaf_fiber, fiber_root,
                  x.l1_crd, x.l2_crd, x.l3_crd, bj, bk, keys, mB, mC, (uint64_t)I, dropped);
    } else if ((rc = csf_keys<true>(ctx, x, bj, bk, keys))) {
      return rc;
    }
    bool in_alt = false;
    if ((rc = radix_sort_pairs<uint64_t>(ctx, keys, vals, ctx->keys[1].as<uint64_t>(), ctx->vals[1].as<double>(),
                                         x.nnz, bj + bk, bj + bk + bi, &in_alt, x.vals)))
      return rc;
    if (in_alt) {
      keys = ctx->keys[1].as<uint64_t>();
      vals = ctx->vals[1].as<double>();
    }
    if (masked) {
      // the nonzeros that take part come first after the sort
      int64_t* hd = ctx->host_meta.as<int64_t>();
      SPWS_HIP(hipMemcpyAsync(hd, dropped, 8, hipMemcpyDeviceToHost, s));
      SPWS_HIP(hipStreamSynchronize(s));
      kept = x.nnz - *hd;
    }
    // segments of 128 nonzeros per half-wavefront
    const int64_t seg = 128;
    const int64_t halves = div_up(kept, seg);
    const unsigned ga = (unsigned)std::max<int64_t>(1, div_up(halves, kAccBlock / 32));
    if (kept > 0) {
      if (R > 32 && masked) {
        SPWS_LAUNCH(ctx, "mttkrp_accum", (k_mttkrp_accum<true, true>), dim3(ga), dim3(kAccBlock), 0, kept, seg, keys,
                    vals, bj, bk, R, Bf, Cf, M, present, mB, mC, rmask);
      } else if (R > 32) {
        SPWS_LAUNCH(ctx, "mttkrp_accum", (k_mttkrp_accum<true, false>), dim3(ga), dim3(kAccBlock), 0, kept, seg, keys,
                    vals, bj, bk, R, Bf, Cf, M, present, mB, mC, rmask);
      } else if (masked) {
        SPWS_LAUNCH(ctx, "mttkrp_accum", (k_mttkrp_accum<false, true>), dim3(ga), dim3(kAccBlock), 0, kept, seg, keys,
                    vals, bj, bk, R, Bf, Cf, M, present, mB, mC, rmask);
      } else {
        SPWS_LAUNCH(ctx, "mttkrp_accum", (k_mttkrp_accum<false, false>), dim3(ga), dim3(kAccBlock), 0, kept, seg,
                    keys, vals, bj, bk, R, Bf, Cf, M, present, mB, mC, rmask);
      }
    }
    ctx->stats.products = kept * (int64_t)R;
  }
  if ((rc = exclusive_scan<int32_t, int32_t>(ctx, present, poff, I))) return rc;
  if (I > 0)
    SPWS_LAUNCH(ctx, "rows_compact", k_rows_compact, dim3((unsigned)div_up(I, 256)), dim3(256), 0, I, present, poff,
                ctx->out_crd.as<int32_t>(), (const unsigned long long*)rmask,
                masked ? ctx->obuf[7].as<uint64_t>() : (uint64_t*)nullptr);
  int32_t* hm = ctx->host_meta.as<int32_t>();
  SPWS_HIP(hipMemcpyAsync(hm, poff + I, 4, hipMemcpyDeviceToHost, s));
  SPWS_HIP(hipStreamSynchronize(s));
  *nrows_out = *hm;
  ctx->stats.nnz = (int64_t)(*hm) * R;
  return SPWS_OK;
}

// Per-fiber sums over nonzeros grouped by fiber (foff[F + 1], keys' low bk
// bits = k): work items of at most kTtmChunk nonzeros, in-order sums, then
// the partial sums of long fibers combined in item order.
static int ttm_accumulate(spws_ctx* ctx, int64_t nnz, int64_t F, int32_t R, int bk, const uint64_t* nkeys,
                          const double* nvals, const int32_t* foff, const double* U, double* Y) {
  int rc;
  // coff = prefix of items per fiber
  if ((rc = ctx->obuf[2].ensure(((size_t)F + 1) * 4 * 2))) return rc;
  int32_t* nchunk = ctx->obuf[2].as<int32_t>();
  int32_t* coff = nchunk + (F + 1);
  const unsigned gf2 = (unsigned)std::min<int64_t>(div_up(F, 256), 65536);
  SPWS_LAUNCH(ctx, "ttm_chunks", k_ttm_chunks, dim3(gf2), dim3(256), 0, F, foff, nchunk);
  if ((rc = exclusive_scan<int32_t, int32_t>(ctx, nchunk, coff, F))) return rc;
  // items <= F + nnz / kTtmChunk
  const int64_t max_items = F + nnz / kTtmChunk + 1;
  if ((rc = ctx->obuf[3].ensure((size_t)max_items * R * 8)) || (rc = ctx->obuf[4].ensure((size_t)max_items * 16)))
    return rc;
  // the item count stays on the device (coff[F]): no wait for it here
  const int32_t* nitems_dev = coff + F;
  int4* items = ctx->obuf[4].as<int4>();
  SPWS_LAUNCH(ctx, "ttm_itemmap", k_ttm_itemmap, dim3(gf2), dim3(256), 0, F, foff, coff, items);
  // one workgroup per slot of a few per CU: each group walks many items
  int cus = 0;
  SPWS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const unsigned ga = (unsigned)std::max<int64_t>(1, std::min<int64_t>(div_up(max_items, 256 / kTtmLanes),
                                                                        (int64_t)std::max(cus, 1) * kTtmBlocksPerCu));
  const bool off32 = ((int64_t)1 << bk) * R * 8 <= (int64_t)UINT32_MAX;  // every U offset fits in 32 bits
  if (off32)
    SPWS_LAUNCH(ctx, "ttm_accum", k_ttm_accum<true>, dim3(ga), dim3(256), 0, nitems_dev, R, items, nkeys, nvals, bk,
                U, Y, ctx->obuf[3].as<double>());
  else
    SPWS_LAUNCH(ctx, "ttm_accum", k_ttm_accum<false>, dim3(ga), dim3(256), 0, nitems_dev, R, items, nkeys, nvals, bk,
                U, Y, ctx->obuf[3].as<double>());
  const unsigned gc = (unsigned)std::min<int64_t>(div_up(F, 256 / kTtmLanes), 65536);
  SPWS_LAUNCH(ctx, "ttm_combine", k_ttm_combine, dim3(gc), dim3(256), 0, F, R, coff, ctx->obuf[3].as<double>(), Y);
  return SPWS_OK;
}

// umask (device, K entries, or null): per U row, its stored columns as bits
// (U given dense, zero where not stored); the fibers' stored output columns
// then go to ctx->obuf[6] after its first 256 bytes (one uint64 per fiber).
static int ttm_run(spws_ctx* ctx, const Csf3& x, int32_t I, int32_t J, int32_t K, int32_t R, const double* U,
                   const uint64_t* umask, int64_t* nfib_out) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_REQUIRE(I >= 0 && J >= 0 && K >= 0 && R >= 1 && R <= 256, SPWS_ERR_INVALID,
               "TTM needs non-negative dimensions and 1 <= R <= 256");
  const bool masked = umask != nullptr;
  SPWS_REQUIRE(!masked || R <= 64, SPWS_ERR_UNSUPPORTED, "a U row mask needs R <= 64");
  // the roots-and-windows path (no sort of the whole tensor) unless the
  // sorting path is asked for or the shapes are outside what it supports
  if (!masked && ctx->ttm_method != 1) {
    bool handled = false;
    if ((rc = ttm_roots_run(ctx, x, I, J, K, R, U, ctx->ttm_method == 2, nfib_out, &handled))) return rc;
    if (handled) return SPWS_OK;
    SPWS_REQUIRE(ctx->ttm_method != 2, SPWS_ERR_UNSUPPORTED,
                 "the roots-and-windows TTM needs 1 <= R <= 64, at most 4,096 column windows of 1024 / R' "
                 "columns (R' = R rounded up to 16, 32 or 64), J < 2^23 and K times the window below 2^32");
  }
  // with a mask, root I is the key of the nonzeros that take no part
  const int bi = bits_of(masked ? (int64_t)I + 1 : (I > 1 ? I : 2)), bj = bits_of(J > 1 ? J : 2);
  SPWS_REQUIRE(bi + bj <= 63, SPWS_ERR_UNSUPPORTED, "tensor dimensions exceed the 64-bit fiber key");
  SPWS_REQUIRE(x.nnz < ((int64_t)1 << 30), SPWS_ERR_UNSUPPORTED, "TTM supports fewer than 2^30 nonzeros");
  hipStream_t s = ctx->stream;
  memset(&ctx->stats, 0, sizeof(ctx->stats));
  const size_t nz = (size_t)(x.nnz > 0 ? x.nnz : 1);
  const int bk_all = bits_of(K > 1 ? K : 2);
  const bool direct = bi + bj + bk_all <= 64;
  SPWS_REQUIRE(!masked || direct, SPWS_ERR_UNSUPPORTED,
               "a U row mask needs the root, leaf and contracted coordinates within 64 bits");
  // hash path: a table of 2 * nnz (power of two) slots
  uint64_t T = 1024;
  while (!direct && T < 2 * (uint64_t)nz) T <<= 1;
  const size_t nflag = direct ? nz + 1 : T + 1;
  if ((rc = ctx->row_nnz.ensure(nflag * 4 * 2)) || (rc = ctx->keys[0].ensure(nz * 8)) ||
      (rc = ctx->keys[1].ensure(nz * 8)) || (rc = ctx->vals[0].ensure(nz * 8)) || (rc = ctx->vals[1].ensure(nz * 8)))
    return rc;
  if (!direct && ((rc = ctx->wide_tab.ensure(T * 8)) || (rc = ctx->slot_crd.ensure(nz * 4)))) return rc;
  unsigned long long* table = ctx->wide_tab.as<unsigned long long>();
  int32_t* leaf_slot = ctx->slot_crd.as<int32_t>();
  int32_t* flag = ctx->row_nnz.as<int32_t>();
  int32_t* off = flag + nflag;
  int32_t* hm = ctx->host_meta.as<int32_t>();
  int64_t F = 0;
  if (x.nnz > 0 && direct) {
    // one sort of the nonzeros on (i, j, k)
    const int bk = bk_all;
    uint64_t* nkeys = ctx->keys[0].as<uint64_t>();
    double* nvals = ctx->vals[0].as<double>();
    const unsigned g = (unsigned)std::min<int64_t>(div_up(x.nnz, 256), 65536);
    unsigned long long* dropped = nullptr;
    if (masked) {
      int32_t *leaf_fiber, *fiber_root;
      if ((rc = csf_owners(ctx, x, &leaf_fiber, &fiber_root))) return rc;
      if ((rc = ctx->obuf[6].ensure(256))) return rc;
      dropped = ctx->obuf[6].as<unsigned long long>();
      SPWS_HIP(hipMemsetAsync(dropped, 0, 8, s));
      SPWS_LAUNCH(ctx, "ttm_keys", k_ttm_direct_keys<true>, dim3(g), dim3(256), 0, x.nnz, leaf_fiber, fiber_root,
                  x.l1_crd, x.l2_crd, x.l3_crd, bj, bk, nkeys, umask, (uint64_t)I, dropped);
    } else if ((rc = csf_keys<false>(ctx, x, bj, bk, nkeys))) {
      return rc;
    }
    // the sort covers (i, j) only: within an i slice the CSF order already
    // has each (i, j)'s nonzeros in increasing k, and the sort is stable
    bool alt = false;
    if ((rc = radix_sort_pairs<uint64_t>(ctx, nkeys, nvals, ctx->keys[1].as<uint64_t>(), ctx->vals[1].as<double>(),
                                         x.nnz, bk, bi + bj + bk, &alt, x.vals)))
      return rc;
    if (alt) {
      nkeys = ctx->keys[1].as<uint64_t>();
      nvals = ctx->vals[1].as<double>();
    }
    // the nonzeros that take part come first after the sort
    int64_t kept = x.nnz;
    if (masked) {
      int64_t* hd = ctx->host_meta.as<int64_t>() + 1;
      SPWS_HIP(hipMemcpyAsync(hd, dropped, 8, hipMemcpyDeviceToHost, s));
      SPWS_HIP(hipStreamSynchronize(s));
      kept = x.nnz - *hd;
    }
    // fiber starts counted per tile and the counts' prefix (flag / off hold
    // nnz + 1 entries here, more than the tiles)
    const int64_t ftiles = div_up(kept, kFibTile);
    if (kept > 0)
      SPWS_LAUNCH(ctx, "ttm_fibcount", k_ttm_fiber_count, dim3((unsigned)ftiles), dim3(kFibBlock), 0, kept, nkeys, bk,
                  flag);
    if ((rc = exclusive_scan<int32_t, int32_t>(ctx, flag, off, ftiles))) return rc;
    SPWS_HIP(hipMemcpyAsync(hm, off + ftiles, 4, hipMemcpyDeviceToHost, s));
    SPWS_HIP(hipStreamSynchronize(s));
    F = *hm;
    if ((rc = ctx->out_crd.ensure((size_t)(F > 0 ? F : 1) * 4 * 2)) ||
        (rc = ctx->out_val.ensure((size_t)(F > 0 ? F : 1) * R * 8)) ||
        (rc = ctx->obuf[5].ensure(((size_t)F + 1) * 4)))
      return rc;
    int32_t* out_i = ctx->out_crd.as<int32_t>();
    int32_t* out_j = out_i + F;
    int32_t* foff = ctx->obuf[5].as<int32_t>();
    if (kept > 0) {
      SPWS_LAUNCH(ctx, "ttm_starts", k_ttm_fiber_starts, dim3((unsigned)ftiles), dim3(kFibBlock), 0, kept, nkeys, bj,
                  bk, off, ftiles, foff, out_i, out_j);
      if ((rc = ttm_accumulate(ctx, kept, F, R, bk, nkeys, nvals, foff, U, ctx->out_val.as<double>()))) return rc;
    }
    if (masked && F > 0) {
      if ((rc = ctx->obuf[6].ensure(256 + (size_t)F * 8))) return rc;
      const unsigned gfm = (unsigned)std::min<int64_t>(div_up(F, 256), 65536);
      SPWS_LAUNCH(ctx, "ttm_fiber_mask", k_ttm_fiber_mask, dim3(gfm), dim3(256), 0, F, foff, nkeys, bk, umask,
                  reinterpret_cast<uint64_t*>(ctx->obuf[6].as<char>() + 256));
    }
    ctx->stats.products = kept * (int64_t)R;
  } else if (x.nnz > 0) {
    int32_t *leaf_fiber, *fiber_root;
    if ((rc = csf_owners(ctx, x, &leaf_fiber, &fiber_root))) return rc;
    SPWS_HIP(hipMemsetAsync(table, 0xff, T * 8, s));
    const unsigned g = (unsigned)std::min<int64_t>(div_up(x.nnz, 256), 65536);
    SPWS_LAUNCH(ctx, "ttm_insert", k_ttm_insert, dim3(g), dim3(256), 0, x.nnz, leaf_fiber, fiber_root, x.l1_crd,
                x.l3_crd, bj, table, (uint64_t)(T - 1), leaf_slot);
    const unsigned gt = (unsigned)std::min<int64_t>(div_up((int64_t)T, 256), 65536);
    SPWS_LAUNCH(ctx, "ttm_occupied", k_ttm_occupied, dim3(gt), dim3(256), 0, (int64_t)T, table, flag);
    if ((rc = exclusive_scan<int32_t, int32_t>(ctx, flag, off, (int64_t)T))) return rc;
    SPWS_HIP(hipMemcpyAsync(hm, off + T, 4, hipMemcpyDeviceToHost, s));
    SPWS_HIP(hipStreamSynchronize(s));
    F = *hm;
    uint64_t* keys = ctx->keys[0].as<uint64_t>();
    double* pay = ctx->vals[0].as<double>();
    SPWS_LAUNCH(ctx, "ttm_collect", k_ttm_collect, dim3(gt), dim3(256), 0, (int64_t)T, table, off, keys, pay);
    bool in_alt = false;
    if ((rc = radix_sort_pairs<uint64_t>(ctx, keys, pay, ctx->keys[1].as<uint64_t>(), ctx->vals[1].as<double>(), F, 0,
                                         bi + bj, &in_alt)))
      return rc;
    if (in_alt) {
      keys = ctx->keys[1].as<uint64_t>();
      pay = ctx->vals[1].as<double>();
    }
    // rank_of_slot reuses the flag array (T entries)
    int32_t* rank_of_slot = flag;
    if ((rc = ctx->out_crd.ensure((size_t)(F > 0 ? F : 1) * 4 * 2)) ||
        (rc = ctx->out_val.ensure((size_t)(F > 0 ? F : 1) * R * 8)))
      return rc;
    int32_t* out_i = ctx->out_crd.as<int32_t>();
    int32_t* out_j = out_i + F;
    const unsigned gf = (unsigned)std::min<int64_t>(div_up(F, 256), 65536);
    SPWS_LAUNCH(ctx, "ttm_rank", k_ttm_rank, dim3(gf), dim3(256), 0, F, keys, pay, bj, rank_of_slot, out_i, out_j);
    double* Y = ctx->out_val.as<double>();
    // group each fiber's nonzeros (stable sort on the fiber rank), then one
    // pass of in-order sums per fiber
    const int bq = bits_of(F > 1 ? F : 2);
    const int bk = bits_of(K > 1 ? K : 2);
    SPWS_REQUIRE(bq + bk <= 64, SPWS_ERR_UNSUPPORTED, "fiber count and K exceed the 64-bit sort key");
    uint64_t* nkeys = ctx->keys[0].as<uint64_t>();
    double* nvals = ctx->vals[0].as<double>();
    SPWS_LAUNCH(ctx, "ttm_keys", k_ttm_keys, dim3(g), dim3(256), 0, x.nnz, leaf_fiber, x.l2_crd, x.vals, leaf_slot,
                rank_of_slot, bk, nkeys, nvals);
    bool alt2 = false;
    if ((rc = radix_sort_pairs<uint64_t>(ctx, nkeys, nvals, ctx->keys[1].as<uint64_t>(), ctx->vals[1].as<double>(),
                                         x.nnz, bk, bk + bq, &alt2)))
      return rc;
    if (alt2) {
      nkeys = ctx->keys[1].as<uint64_t>();
      nvals = ctx->vals[1].as<double>();
    }
    // fiber offsets reuse the flag array past rank_of_slot's T entries
    int32_t* foff = off;
    SPWS_LAUNCH(ctx, "ttm_heads", k_ttm_heads, dim3(g), dim3(256), 0, x.nnz, nkeys, bk, F, foff);
    if ((rc = ttm_accumulate(ctx, x.nnz, F, R, bk, nkeys, nvals, foff, U, Y))) return rc;
    ctx->stats.products = x.nnz * (int64_t)R;
  } else {
    if ((rc = ctx->out_crd.ensure(8)) || (rc = ctx->out_val.ensure(8))) return rc;
  }
  SPWS_HIP(hipStreamSynchronize(s));
  *nfib_out = F;
  ctx->stats.nnz = F * R;
  return SPWS_OK;
}

}  // namespace spws

using namespace spws;

extern "C" {

int spws_ctx_set_ttm_method(spws_ctx* ctx, int32_t method) {
  SPWS_REQUIRE(ctx != nullptr, SPWS_ERR_INVALID, "null context");
  SPWS_REQUIRE(method >= 0 && method <= 2, SPWS_ERR_INVALID, "TTM method %d is not 0, 1 or 2", method);
  ctx->ttm_method = method;
  return SPWS_OK;
}

int spws_mttkrp(spws_ctx* ctx, int32_t I, int32_t J, int32_t K, int32_t R, int32_t nroots, const int32_t* l1_crd,
                const int32_t* l2_pos, const int32_t* l2_crd, int32_t nfibers, const int32_t* l3_pos,
                const int32_t* l3_crd, const double* vals, int64_t nnz, const double* Bf, const double* Cf,
                int32_t* nrows, int32_t** rows, double** M) {
  return spws_mttkrp_masked(ctx, I, J, K, R, nroots, l1_crd, l2_pos, l2_crd, nfibers, l3_pos, l3_crd, vals, nnz, Bf,
                            Cf, nullptr, nullptr, nrows, rows, nullptr, M);
}

int spws_mttkrp_masked(spws_ctx* ctx, int32_t I, int32_t J, int32_t K, int32_t R, int32_t nroots,
                       const int32_t* l1_crd, const int32_t* l2_pos, const int32_t* l2_crd, int32_t nfibers,
                       const int32_t* l3_pos, const int32_t* l3_crd, const double* vals, int64_t nnz, const double* Bf,
                       const double* Cf, const uint64_t* Bmask, const uint64_t* Cmask, int32_t* nrows,
                       int32_t** rows, uint64_t** row_mask, double** M) {
  SPWS_REQUIRE(nnz >= 0 && nroots >= 0 && nfibers >= 0, SPWS_ERR_INVALID, "negative size");
  SPWS_REQUIRE(nnz == 0 || (l1_crd && l2_pos && l2_crd && l3_pos && l3_crd && vals && Bf && Cf), SPWS_ERR_INVALID,
               "null pointer argument");
  SPWS_REQUIRE(nnz < ((int64_t)1 << 31), SPWS_ERR_UNSUPPORTED, "MTTKRP supports fewer than 2^31 nonzeros");
  Csf3 x{nroots, l1_crd, l2_pos, l2_crd, nfibers, l3_pos, l3_crd, vals, nnz};
  int32_t nr = 0;
  int rc = mttkrp_run(ctx, x, I, J, K, R, Bf, Cf, Bmask, Cmask, &nr);
  if (rc) return rc;
  if (nrows) *nrows = nr;
  if (rows) *rows = ctx->out_crd.as<int32_t>();
  if (row_mask) *row_mask = Bmask ? ctx->obuf[7].as<uint64_t>() : nullptr;
  if (M) *M = ctx->out_val.as<double>();
  return SPWS_OK;
}

int spws_ttm(spws_ctx* ctx, int32_t I, int32_t J, int32_t K, int32_t R, int32_t nroots, const int32_t* l1_crd,
             const int32_t* l2_pos, const int32_t* l2_crd, int32_t nfibers, const int32_t* l3_pos,
             const int32_t* l3_crd, const double* vals, int64_t nnz, const double* U, int64_t* nfib,
             int32_t** fib_i, int32_t** fib_j, double** Y) {
  return spws_ttm_masked(ctx, I, J, K, R, nroots, l1_crd, l2_pos, l2_crd, nfibers, l3_pos, l3_crd, vals, nnz, U,
                         nullptr, nfib, fib_i, fib_j, nullptr, Y);
}

int spws_ttm_masked(spws_ctx* ctx, int32_t I, int32_t J, int32_t K, int32_t R, int32_t nroots,
                    const int32_t* l1_crd, const int32_t* l2_pos, const int32_t* l2_crd, int32_t nfibers,
                    const int32_t* l3_pos, const int32_t* l3_crd, const double* vals, int64_t nnz, const double* U,
                    const uint64_t* Umask, int64_t* nfib, int32_t** fib_i, int32_t** fib_j, uint64_t** fib_mask,
                    double** Y) {
  SPWS_REQUIRE(nnz >= 0 && nroots >= 0 && nfibers >= 0, SPWS_ERR_INVALID, "negative size");
  SPWS_REQUIRE(nnz == 0 || (l1_crd && l2_pos && l2_crd && l3_pos && l3_crd && vals && U), SPWS_ERR_INVALID,
               "null pointer argument");
  Csf3 x{nroots, l1_crd, l2_pos, l2_crd, nfibers, l3_pos, l3_crd, vals, nnz};
  int64_t F = 0;
  int rc = ttm_run(ctx, x, I, J, K, R, U, Umask, &F);
  if (rc) return rc;
  if (nfib) *nfib = F;
  if (fib_i) *fib_i = ctx->out_crd.as<int32_t>();
  if (fib_j) *fib_j = ctx->out_crd.as<int32_t>() + F;
  if (fib_mask) *fib_mask = Umask ? reinterpret_cast<uint64_t*>(ctx->obuf[6].as<char>() + 256) : nullptr;
  if (Y) *Y = ctx->out_val.as<double>();
  return SPWS_OK;
}

}  // extern "C"
