// Row-wise CSR x CSR SpGEMM with per-row workspaces (Gustavson order).
//
// Pipeline (one stream; two host synchronisations, both tiny reads):
//   1. k_rowstat      per-row upper bound ub_i = sum_{j in A_i} len(B_j) and
//                     the path of each row (bin histogram, max ub per bin,
//                     total); the metadata block starts zero (the last
//                     kernel of the previous product leaves it so, else a
//                     clear kernel runs first).
//   -- sync: read the metadata block
//   If every row has ub <= 64: k_row_wave over the rows in order into fixed
//   64-entry slots, then k_compact_fixed forms C (no bins, no scans).  When
//   the previous product on the context took this path, both kernels are
//   launched before the sync (rows over 64 products skipped) and their
//   result is kept if the metadata allows the path: one sync in all.
//   Otherwise:
//   2. scan(ub)       int64 offsets of each row's slot region (ub_i entries).
//   3. k_binscatter   row ids grouped by path.
//   4. numeric        one kernel per non-empty path; each row accumulates in
//                     its workspace and is written in column order into its
//                     slot region:
//        wave   / wave2   one wavefront per row, LDS open-addressing table
//                         (2*ub slots), LDS atomicCAS keys + ds_add_f64
//                         values, register bitonic sort across the wavefront;
//        block  / block2  one workgroup per row: with at most 2^17 columns,
//                         the row's products gathered once into registers,
//                         an LDS bitmap of the columns (rank = popcount
//                         prefix) and ds_add_f64 at each column's rank, no
//                         sort; wider outputs use an LDS table and an LDS
//                         bitonic sort;
//        dense / wide     longer rows: with at most 2^17 columns, the column
//                         bitmap with the products staged in the output
//                         arrays and read back coalesced, accumulated in LDS
//                         (or with FP64 atomics into the slot region when the
//                         row has too many entries for LDS); wider outputs:
//                         the densest rows use a dense FP64 accumulator +
//                         flag array per workgroup in HBM, the others a
//                         per-row HBM hash table (2*ub slots), then an LDS or
//                         global bitonic sort.
//   5. scan(nnz)      C_pos.
//   6. k_compact      slot regions -> contiguous C_crd / C_vals.
//   -- sync: read nnz(C)
//
// The symbolic variant (kValues = false) stops after counting the distinct
// columns per row (steps 1-5), which is the TACO assemble() phase.
//
// Reference: benchmark/Parallel/CSR_CSR_hash_sequential.h:232-353 (hash
// workspace, sorted output), seperate/CSR_CSR_1.h:69-167 (assemble),
// evaluate/CSR_CSR_1.h:153-171 (per-row upper bound).
#include "rowwise_kernels.hpp"

namespace spws {

// ------------------------------------------------------------------ host driver

static int validate_tuning(const spws_rowwise_tuning& t) {
  SPWS_REQUIRE(t.wave_ub >= 1 && t.wave_ub <= kWaveUbMax, SPWS_ERR_INVALID, "wave_ub must be in [1, %d]",
               kWaveUbMax);
  SPWS_REQUIRE(t.wave2_ub >= t.wave_ub && t.wave2_ub <= kWave2UbMax, SPWS_ERR_INVALID,
               "wave2_ub must be in [wave_ub, %d]", kWave2UbMax);
  SPWS_REQUIRE(t.block_ub >= t.wave2_ub && t.block_ub <= kBlockUbMax, SPWS_ERR_INVALID,
               "block_ub must be in [wave2_ub, %d]", kBlockUbMax);
  SPWS_REQUIRE(t.block2_ub >= t.block_ub && t.block2_ub <= kBlock2UbMax, SPWS_ERR_INVALID,
               "block2_ub must be in [block_ub, %d]", kBlock2UbMax);
  SPWS_REQUIRE(t.dense_ratio >= 1, SPWS_ERR_INVALID, "dense_ratio must be >= 1");
  SPWS_REQUIRE(t.force_path >= -1 && t.force_path <= P_WIDE, SPWS_ERR_INVALID, "force_path must be in [-1, 5]");
  return SPWS_OK;
}

// The short-row product in two launches (k_short_rows, k_short_compact)
// into the context's C arrays.  *ok is false when it does not apply (more
// wavefronts than the compaction sums) or a row does not fit: the caller
// then runs the general pipeline, which recomputes C.
static int run_short(spws_ctx* ctx, int32_t m, int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax,
                     const int32_t* Bp, const int32_t* Bi, const double* Bx, bool* ok, int64_t* nnz,
                     int64_t* products) {
  hipStream_t s = ctx->stream;
  int rc;
  *ok = false;
  if (m <= 0) return SPWS_OK;
  int cus = 0;
  SPWS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  // about 16 wavefronts per CU, at most kShortRows rows each
  const int32_t q = (int32_t)std::min<int64_t>(kShortRows, std::max<int64_t>(1, div_up(m, (int64_t)cus * 16)));
  const int64_t nwaves = div_up(m, q);
  if (nwaves > kShortMaxWaves) return SPWS_OK;
  const size_t slots = (size_t)nwaves * q * kWave;
  const size_t nw = (size_t)nwaves;
  const size_t nwg = (size_t)div_up(nwaves, kShortWaves);  // workgroups of either kernel
  if ((rc = ctx->slot_crd.ensure(slots * 4)) || (rc = ctx->slot_val.ensure(slots * 8)) ||
      (rc = ctx->row_nnz.ensure((size_t)m * 4)) || (rc = ctx->short_meta.ensure((nw + 2 * nwg) * 4)) ||
      (rc = ctx->out_pos.ensure(((size_t)m + 1) * 4)) || (rc = ctx->out_crd.ensure(slots * 4)) ||
      (rc = ctx->out_val.ensure(slots * 8)))
    return rc;
  // the compaction's last workgroup writes products, the fit flag and
  // nnz(C) straight into pinned host memory (no copy after the kernels)
  unsigned long long* h = ctx->host_meta.as<unsigned long long>() + 64;
  int32_t* wave_tot = ctx->short_meta.as<int32_t>();
  int32_t* wg_tot = wave_tot + nw;
  int32_t* wg_prod = wg_tot + nwg;
  const unsigned g1 = (unsigned)nwg;
  // one register per row of a wavefront (R = q), packed 32-bit sort keys
  // when the columns allow them
  const bool packed = n <= (1 << 26) - 1;
#define SPWS_SHORT_ROWS(R)                                                                                  \
  do {                                                                                                      \
    if (packed)                                                                                             \
      SPWS_LAUNCH(ctx, "short_rows", (k_short_rows<R, true>), dim3(g1), dim3(kShortBlock), 0, m, q,        \
                  (int32_t)nwaves, Ap, Ai, Ax, Bp, Bi, Bx, ctx->slot_crd.as<int32_t>(),                    \
                  ctx->slot_val.as<double>(), ctx->row_nnz.as<int32_t>(), wave_tot, wg_tot, wg_prod);     \
    else                                                                                                    \
      SPWS_LAUNCH(ctx, "short_rows", (k_short_rows<R, false>), dim3(g1), dim3(kShortBlock), 0, m, q,       \
                  (int32_t)nwaves, Ap, Ai, Ax, Bp, Bi, Bx, ctx->slot_crd.as<int32_t>(),                    \
                  ctx->slot_val.as<double>(), ctx->row_nnz.as<int32_t>(), wave_tot, wg_tot, wg_prod);     \
  } while (0)
  switch (q) {
    case 1: SPWS_SHORT_ROWS(1); break;
    case 2: SPWS_SHORT_ROWS(2); break;
    case 3: SPWS_SHORT_ROWS(3); break;
    case 4: SPWS_SHORT_ROWS(4); break;
    case 5: SPWS_SHORT_ROWS(5); break;
    case 6: SPWS_SHORT_ROWS(6); break;
    case 7: SPWS_SHORT_ROWS(7); break;
    default: SPWS_SHORT_ROWS(8); break;
  }
#undef SPWS_SHORT_ROWS
  SPWS_LAUNCH(ctx, "short_compact", k_short_compact, dim3(g1), dim3(kShortBlock), 0, m, q, (int32_t)nwaves,
              (const int32_t*)ctx->row_nnz.as<int32_t>(), (const int32_t*)wave_tot, (const int32_t*)wg_tot,
              (const int32_t*)wg_prod,
              (const int32_t*)ctx->slot_crd.as<int32_t>(), (const double*)ctx->slot_val.as<double>(),
              ctx->out_pos.as<int32_t>(), ctx->out_crd.as<int32_t>(), ctx->out_val.as<double>(), h);
  SPWS_HIP(hipStreamSynchronize(s));
  *ok = h[1] == 0;
  *nnz = (int64_t)h[2];
  *products = (int64_t)h[0];
  return SPWS_OK;
}

template <bool kValues>
static int rowwise_run(spws_ctx* ctx, int32_t m, int32_t kdim, int32_t n, const int32_t* Ap, const int32_t* Ai,
                       const double* Ax, const int32_t* Bp, const int32_t* Bi, const double* Bx, int64_t* nnz_out) {
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  SPWS_REQUIRE(m >= 0 && kdim >= 0 && n >= 0, SPWS_ERR_INVALID, "negative dimension");
  SPWS_REQUIRE(Ap && Bp && nnz_out, SPWS_ERR_INVALID, "null pointer argument");
  hipStream_t s = ctx->stream;
  const spws_rowwise_tuning& t = ctx->rw;
  // the forced paths must still fit the compiled LDS tables
  PathParams pp{t.wave_ub, t.wave2_ub, t.block_ub, t.block2_ub, t.dense_ratio, t.force_path};
  memset(&ctx->stats, 0, sizeof(ctx->stats));

  // the previous product had only short rows: try the short-row path first,
  // with no upper-bound pass (its kernels check every row themselves)
  if (kValues && t.force_path < 0 && m > 0 && ctx->rw_fixed_last) {
    bool ok = false;
    int64_t nnz = 0, prods = 0;
    if ((rc = run_short(ctx, m, n, Ap, Ai, Ax, Bp, Bi, Bx, &ok, &nnz, &prods))) return rc;
    if (ok) {
      ctx->stats.products = prods;
      ctx->stats.nnz = nnz;
      ctx->stats.rows_per_path[P_WAVE] = m;
      *nnz_out = nnz;
      return SPWS_OK;
    }
    ctx->rw_fixed_last = false;
  }

  const size_t mm = (size_t)(m > 0 ? m : 1);
  if ((rc = ctx->row_ub.ensure(mm * 8)) || (rc = ctx->row_off.ensure((mm + 1) * 8)) ||
      (rc = ctx->row_nnz.ensure(mm * 4)) || (rc = ctx->bin_rows.ensure(mm * 4)) ||
      (rc = ctx->bin_meta.ensure(64 * 8)) || (rc = ctx->out_pos.ensure((mm + 1) * 4)))
    return rc;
  int32_t* bmeta = ctx->bin_meta.as<int32_t>();
  unsigned long long* maxub_d = reinterpret_cast<unsigned long long*>(bmeta + 32);
  // the last kernel of a completed product leaves the block zero
  if (!ctx->rw_meta_clean) SPWS_LAUNCH(ctx, "meta_clear", k_meta_clear, dim3(1), dim3(64), 0, bmeta, maxub_d);
  ctx->rw_meta_clean = false;
  if (m > 0) {
    SPWS_LAUNCH(ctx, "rowstat", k_rowstat, dim3((unsigned)std::min<int64_t>(div_up(m, kRsRows), 65536)),
                dim3(kRsBlock), 0,
                m, nullptr, n, Ap, Ai, Bp,
                ctx->row_ub.as<int64_t>(), bmeta, maxub_d, pp);
  }
  // one small read of the metadata block: counts (ints [0, 8)), then the
  // max ub per path and the total products (uint64 from int 32 on)
  int32_t* hc = ctx->host_meta.as<int32_t>();
  const int64_t* hm = reinterpret_cast<const int64_t*>(hc + 32);
  int32_t* h32 = hc + 100;  // nnz(C) of the fixed-slot path, past the metadata block
  SPWS_HIP(hipMemcpyAsync(hc, bmeta, 32 * 4 + 9 * 8, hipMemcpyDeviceToHost, s));

  // The fixed-slot path (every row fits one wavefront: rows straight into
  // fixed 64-entry slots, then one compaction that also forms C_pos; no
  // bins, no scans).  When the last product on this context took it, it is
  // launched now, before the metadata reaches the host: the wave kernel
  // skips (count 0) any row whose upper bound exceeds 64, and the result is
  // kept only if the metadata shows that every row fits, which is exactly
  // when the path would have been chosen; otherwise the general path below
  // recomputes C, after the host's copy of the metadata block is put back
  // (the compaction left it zero for the next product).
  auto launch_fixed = [&](size_t cap, int32_t* meta_zero) -> int {
    const size_t fixed = (size_t)m * kWave;
    int rc2;
    if ((kValues && ((rc2 = ctx->slot_crd.ensure(fixed * 4)) || (rc2 = ctx->slot_val.ensure(fixed * 8)) ||
                     (rc2 = ctx->out_crd.ensure(cap * 4)) || (rc2 = ctx->out_val.ensure(cap * 8)))))
      return rc2;
    SPWS_LAUNCH(ctx, "row_wave", (k_row_wave<128, 1, kValues, kWaveRowsK>),
                dim3((unsigned)std::min<int64_t>(div_up(m, kWaveRowsK), 1 << 20)), dim3(64), 0,
                (const int32_t*)nullptr, bmeta, (int)P_WAVE, Ap, Ai, Ax, Bp, Bi, Bx, (const int64_t*)nullptr,
                kValues ? ctx->slot_crd.as<int32_t>() : nullptr, kValues ? ctx->slot_val.as<double>() : nullptr,
                ctx->row_nnz.as<int32_t>(), m, (const int64_t*)ctx->row_ub.as<int64_t>());
    SPWS_LAUNCH(ctx, "compact_fixed", (k_compact_fixed<kValues>), dim3((unsigned)div_up(m, kCfRows)),
                dim3(kCfBlock), 0, m, ctx->row_nnz.as<int32_t>(), kValues ? ctx->slot_crd.as<int32_t>() : nullptr,
                kValues ? ctx->slot_val.as<double>() : nullptr, ctx->out_pos.as<int32_t>(),
                kValues ? ctx->out_crd.as<int32_t>() : nullptr, kValues ? ctx->out_val.as<double>() : nullptr,
                meta_zero);
    SPWS_HIP(hipMemcpyAsync(h32, ctx->out_pos.as<int32_t>() + m, 4, hipMemcpyDeviceToHost, s));
    return SPWS_OK;
  };
  // (the numeric product took the one-launch path above when this was set)
  const bool spec = !kValues && ctx->rw_fixed_last && t.force_path < 0 && m > 0 && m <= kCfMaxRows;
  // at most 64 entries per row whichever rows the wave kernel skips
  if (spec && (rc = launch_fixed((size_t)m * kWave, bmeta))) return rc;
  SPWS_HIP(hipStreamSynchronize(s));
  int64_t counts[8], maxub[8];
  for (int p = 0; p < 8; ++p) {
    counts[p] = hc[p];
    maxub[p] = hm[p];
  }
  const int64_t total_ub = hm[8];
  ctx->stats.products = total_ub;
  for (int p = 0; p < kPaths; ++p) ctx->stats.rows_per_path[p] = counts[p];
  SPWS_REQUIRE(total_ub <= (int64_t)INT32_MAX, SPWS_ERR_UNSUPPORTED,
               "row-wise SpGEMM: %lld products exceed the int32 output index range", (long long)total_ub);
  if (t.force_path >= 0) {
    const int64_t caps[4] = {kWaveUbMax, kWave2UbMax, kBlockUbMax, kBlock2UbMax};
    if (t.force_path <= P_BLOCK2)
      SPWS_REQUIRE(maxub[t.force_path] <= caps[t.force_path], SPWS_ERR_INVALID,
                   "forced path %d cannot hold a row of %lld products", t.force_path,
                   (long long)maxub[t.force_path]);
  }

  // every row has at most 64 products: the two-launch short-row path
  // (numeric), or the fixed-slot path (symbolic, or when the short-row path
  // declines)
  const bool short_rows = t.force_path < 0 && m > 0 && counts[P_WAVE] + counts[P_EMPTY] == m &&
                          maxub[P_WAVE] <= kWaveUbMax;
  const bool fixed_ok = short_rows && m <= kCfMaxRows;
  ctx->rw_fixed_last = short_rows;
  if (kValues && short_rows) {
    bool ok = false;
    int64_t nnz = 0, prods = 0;
    // the metadata block is zero again for the next product
    SPWS_LAUNCH(ctx, "meta_clear", k_meta_clear, dim3(1), dim3(64), 0, bmeta, maxub_d);
    ctx->rw_meta_clean = true;
    if ((rc = run_short(ctx, m, n, Ap, Ai, Ax, Bp, Bi, Bx, &ok, &nnz, &prods))) return rc;
    if (ok) {
      *nnz_out = nnz;
      ctx->stats.nnz = nnz;
      return SPWS_OK;
    }
    ctx->rw_meta_clean = false;
    SPWS_HIP(hipMemcpyAsync(bmeta, hc, 32 * 4 + 9 * 8, hipMemcpyHostToDevice, s));
  }
  if (fixed_ok) {
    if (!spec) {
      if ((rc = launch_fixed((size_t)(total_ub > 0 ? total_ub : 1), bmeta))) return rc;
      SPWS_HIP(hipStreamSynchronize(s));
    }
    ctx->rw_meta_clean = true;
    *nnz_out = *h32;
    ctx->stats.nnz = *h32;
    return SPWS_OK;
  }
  if (spec) SPWS_HIP(hipMemcpyAsync(bmeta, hc, 32 * 4 + 9 * 8, hipMemcpyHostToDevice, s));

  rc = exclusive_scan<int64_t, int64_t>(ctx, ctx->row_ub.as<int64_t>(), ctx->row_off.as<int64_t>(), m);
  if (rc) return rc;
  SPWS_LAUNCH(ctx, "bin_offsets", k_bin_offsets, dim3(1), dim3(64), 0, bmeta);

  const size_t slots = (size_t)(total_ub > 0 ? total_ub : 1);
  if (kValues) {
    if ((rc = ctx->slot_crd.ensure(slots * 4)) || (rc = ctx->slot_val.ensure(slots * 8)) ||
        (rc = ctx->out_crd.ensure(slots * 4)) || (rc = ctx->out_val.ensure(slots * 8)))
      return rc;
  }
  if (m > 0) {
    SPWS_LAUNCH(ctx, "binscatter", k_binscatter, dim3((unsigned)div_up(m, kRsBlock)), dim3(kRsBlock), 0, m, nullptr, n,
                ctx->row_ub.as<int64_t>(), bmeta, ctx->bin_rows.as<int32_t>(), ctx->row_nnz.as<int32_t>(), pp);
  }
  const int32_t* rows = ctx->bin_rows.as<int32_t>();
  const int64_t* roff = ctx->row_off.as<int64_t>();
  int32_t* scrd = kValues ? ctx->slot_crd.as<int32_t>() : nullptr;
  double* sval = kValues ? ctx->slot_val.as<double>() : nullptr;
  int32_t* rnnz = ctx->row_nnz.as<int32_t>();
  auto grid_for = [](int64_t cnt, int64_t cap) { return (unsigned)(cnt < cap ? cnt : cap); };

  if (counts[P_WAVE]) {
    SPWS_LAUNCH(ctx, "row_wave", (k_row_wave<128, 1, kValues, kWaveRowsK>),
                dim3(grid_for(div_up(counts[P_WAVE], kWaveRowsK), 1 << 20)), dim3(64), 0, rows, bmeta, (int)P_WAVE,
                Ap, Ai, Ax, Bp, Bi, Bx, roff, scrd, sval, rnnz, 0);
  }
  if (counts[P_WAVE2]) {
    SPWS_LAUNCH(ctx, "row_wave", (k_row_wave<512, 4, kValues>), dim3(grid_for(counts[P_WAVE2], 1 << 20)), dim3(64),
                0, rows, bmeta, (int)P_WAVE2, Ap, Ai, Ax, Bp, Bi, Bx, roff, scrd, sval, rnnz, 0);
  }
  // block bins: the LDS bitmap kernel while all columns fit the bitmap, else
  // the LDS hash table and bitonic sort
  for (int path : {(int)P_BLOCK, (int)P_BLOCK2}) {
    if (!counts[path]) continue;
    if (n <= kBitmapCols) {
      const int64_t vcap = std::min<int64_t>(maxub[path], n);
      const unsigned lds = (unsigned)bitmap_lds_bytes(n, vcap, kValues, kBitmapBlock);
      const unsigned grid_k = grid_for(div_up(counts[path], kBitmapRowsK), kBitmapRowsGrid);
      if (!kValues)
        SPWS_LAUNCH(ctx, "row_bitmap", (k_row_bitmap<false>), dim3(grid_for(counts[path], 1 << 16)),
                    dim3(kBitmapBlock), lds, rows, bmeta, path, n, Ap, Ai, Ax, Bp, Bi, Bx, roff, scrd, sval, rnnz,
                    (const int32_t*)nullptr, (const int32_t*)nullptr);
      else if (path == P_BLOCK)
        SPWS_LAUNCH(ctx, "row_bitmap", (k_row_bitmap_rows<kBitmapRowsK, kBitmapItems1>), dim3(grid_k),
                    dim3(kBitmapBlock), lds, rows, bmeta, path, n, Ap, Ai, Ax, Bp, Bi, Bx, roff, scrd, sval, rnnz,
                    (const int32_t*)nullptr, (const int32_t*)nullptr);
      else
        SPWS_LAUNCH(ctx, "row_bitmap", (k_row_bitmap_rows<kBitmapRowsK, kBitmapItems2>), dim3(grid_k),
                    dim3(kBitmapBlock), lds, rows, bmeta, path, n, Ap, Ai, Ax, Bp, Bi, Bx, roff, scrd, sval, rnnz,
                    (const int32_t*)nullptr, (const int32_t*)nullptr);
    } else if (path == P_BLOCK) {
      SPWS_LAUNCH(ctx, "row_block", (k_row_block<2048, 1024, 256, kValues>), dim3(grid_for(counts[path], 1 << 16)),
                  dim3(256), 0, rows, bmeta, path, Ap, Ai, Ax, Bp, Bi, Bx, roff, scrd, sval, rnnz);
    } else {
      SPWS_LAUNCH(ctx, "row_block", (k_row_block<8192, 4096, 512, kValues>), dim3(grid_for(counts[path], 1 << 16)),
                  dim3(512), 0, rows, bmeta, path, Ap, Ai, Ax, Bp, Bi, Bx, roff, scrd, sval, rnnz);
    }
  }
  // rows longer than the block bins: while the columns fit the LDS bitmap,
  // the bitmap kernel accumulating in the rows' slot regions replaces both
  // the dense accumulator and the HBM hash table
  if (n <= kBitmapCols) {
    for (int path : {(int)P_DENSE, (int)P_WIDE}) {
      if (!counts[path]) continue;
      // staged in the output arrays, which the compaction fills later; rows
      // of at most acc_cap entries accumulate in LDS
      const size_t bm_bytes = bitmap_lds_bytes(n, 0, false, kBitmapLongBlock);
      const int32_t acc_cap = (int32_t)std::min<int64_t>(std::min<int64_t>(maxub[path], n),
                                                         (int64_t)((kLdsBudget - bm_bytes) / 8));
      if (kValues)
        SPWS_LAUNCH(ctx, "row_bitmap_long", k_row_bitmap_staged<kBitmapLongBlock>,
                    dim3(grid_for(counts[path], 4096)), dim3(kBitmapLongBlock),
                    bitmap_lds_bytes(n, acc_cap, true, kBitmapLongBlock), rows, bmeta, path, n, Ap, Ai, Ax, Bp, Bi,
                    Bx, roff, scrd, sval, rnnz, (const int32_t*)nullptr, (const int32_t*)nullptr,
                    ctx->out_crd.as<int32_t>(), ctx->out_val.as<double>(), acc_cap);
      else
        SPWS_LAUNCH(ctx, "row_bitmap_long", (k_row_bitmap<false, true, kBitmapLongBlock>),
                    dim3(grid_for(counts[path], 4096)), dim3(kBitmapLongBlock),
                    bitmap_lds_bytes(n, 0, false, kBitmapLongBlock), rows, bmeta, path, n, Ap, Ai, Ax, Bp, Bi, Bx,
                    roff, scrd, sval, rnnz, (const int32_t*)nullptr, (const int32_t*)nullptr);
      counts[path] = 0;
    }
  }
  if (counts[P_DENSE]) {
    // bounded scratch: at most ~2 GiB of dense accumulators
    const size_t per = (size_t)(n > 0 ? n : 1) * (kValues ? 9 : 1);
    int64_t g = (int64_t)((size_t)2 << 30) / (int64_t)per;
    if (g > 512) g = 512;
    if (g < 1) g = 1;
    if (g > counts[P_DENSE]) g = counts[P_DENSE];
    const size_t nn = (size_t)(n > 0 ? n : 1);
    if ((rc = ctx->dense_flag.ensure((size_t)g * nn))) return rc;
    if (kValues && (rc = ctx->dense_val.ensure((size_t)g * nn * 8))) return rc;
    SPWS_HIP(hipMemsetAsync(ctx->dense_flag.ptr, 0, (size_t)g * nn, s));
    if (kValues) SPWS_HIP(hipMemsetAsync(ctx->dense_val.ptr, 0, (size_t)g * nn * 8, s));
    SPWS_LAUNCH(ctx, "row_dense", (k_row_dense<kValues>), dim3((unsigned)g), dim3(kDenseBlock), 0, rows, bmeta,
                (int)P_DENSE, n, Ap, Ai, Ax, Bp, Bi, Bx, roff, scrd, sval, rnnz, ctx->dense_flag.as<uint8_t>(),
                kValues ? ctx->dense_val.as<double>() : nullptr);
  }
  if (counts[P_WIDE]) {
    int64_t tab = (int64_t)next_pow2_u32((uint32_t)(2 * maxub[P_WIDE]));
    if (tab < 64) tab = 64;
    const size_t per = (size_t)tab * (kValues ? 12 : 4);
    int64_t g = (int64_t)((size_t)2 << 30) / (int64_t)per;
    if (g > 256) g = 256;
    if (g < 1) g = 1;
    if (g > counts[P_WIDE]) g = counts[P_WIDE];
    if ((rc = ctx->wide_tab.ensure((size_t)g * per))) return rc;
    int32_t* tk = ctx->wide_tab.as<int32_t>();
    double* tv = kValues ? reinterpret_cast<double*>(ctx->wide_tab.as<char>() + (size_t)g * tab * 4) : nullptr;
    SPWS_LAUNCH(ctx, "row_wide", (k_row_wide<kValues>), dim3((unsigned)g), dim3(kWideBlock), 0, rows, bmeta,
                (int)P_WIDE, Ap, Ai, Ax, Bp, Bi, Bx, ctx->row_ub.as<int64_t>(), roff, scrd, sval, rnnz, tk, tv, tab);
  }
  rc = exclusive_scan<int32_t, int32_t>(ctx, rnnz, ctx->out_pos.as<int32_t>(), m);
  if (rc) return rc;
  if (kValues && m > 0) {
    unsigned g = (unsigned)div_up(m, 4);
    if (g > 65536) g = 65536;
    SPWS_LAUNCH(ctx, "compact", k_compact, dim3(g), dim3(256), 0, m, nullptr, roff, ctx->out_pos.as<int32_t>(), scrd,
                sval, nullptr, ctx->out_crd.as<int32_t>(), ctx->out_val.as<double>(), bmeta);
  }
  SPWS_HIP(hipMemcpyAsync(h32, ctx->out_pos.as<int32_t>() + m, 4, hipMemcpyDeviceToHost, s));
  SPWS_HIP(hipStreamSynchronize(s));
  ctx->rw_meta_clean = kValues && m > 0;
  *nnz_out = *h32;
  ctx->stats.nnz = *h32;
  return SPWS_OK;
}

}  // namespace spws

using namespace spws;

extern "C" {

int spws_ctx_set_rowwise_tuning(spws_ctx* ctx, const spws_rowwise_tuning* t) {
  SPWS_REQUIRE(ctx && t, SPWS_ERR_INVALID, "null argument");
  int rc = validate_tuning(*t);
  if (rc) return rc;
  ctx->rw = *t;
  return SPWS_OK;
}

int spws_ctx_get_rowwise_tuning(spws_ctx* ctx, spws_rowwise_tuning* t) {
  SPWS_REQUIRE(ctx && t, SPWS_ERR_INVALID, "null argument");
  *t = ctx->rw;
  return SPWS_OK;
}

int spws_spgemm_rowwise(spws_ctx* ctx, int32_t m, int32_t k, int32_t n, const int32_t* A_pos,
                        const int32_t* A_crd, const double* A_vals, const int32_t* B_pos, const int32_t* B_crd,
                        const double* B_vals, int64_t* C_nnz, int32_t** C_pos, int32_t** C_crd, double** C_vals) {
  int64_t nnz = 0;
  int rc = rowwise_run<true>(ctx, m, k, n, A_pos, A_crd, A_vals, B_pos, B_crd, B_vals, &nnz);
  if (rc) return rc;
  if (C_nnz) *C_nnz = nnz;
  if (C_pos) *C_pos = ctx->out_pos.as<int32_t>();
  if (C_crd) *C_crd = ctx->out_crd.as<int32_t>();
  if (C_vals) *C_vals = ctx->out_val.as<double>();
  return SPWS_OK;
}

int spws_spgemm_rowwise_symbolic(spws_ctx* ctx, int32_t m, int32_t k, int32_t n, const int32_t* A_pos,
                                 const int32_t* A_crd, const int32_t* B_pos, const int32_t* B_crd, int64_t* C_nnz,
                                 int32_t** C_pos) {
  int64_t nnz = 0;
  int rc = rowwise_run<false>(ctx, m, k, n, A_pos, A_crd, nullptr, B_pos, B_crd, nullptr, &nnz);
  if (rc) return rc;
  if (C_nnz) *C_nnz = nnz;
  if (C_pos) *C_pos = ctx->out_pos.as<int32_t>();
  return SPWS_OK;
}

}  // extern "C"
