// TACO-shaped host entry points: host taco_tensor_t in, device pipeline,
// malloc'd host arrays out (so the caller's free() works, as with the
// reference's compute()).  Each thread gets its own default context on the
// device named by SPWS_DEVICE (default 0).
#include <climits>
#include <cstdlib>
#include <vector>

#include "common.hpp"

namespace spws {

struct DefaultCtx {
  spws_ctx* ctx = nullptr;
  ~DefaultCtx() {
    if (ctx) spws_ctx_destroy(ctx);
  }
};

static thread_local DefaultCtx g_default;

int default_ctx(spws_ctx** out) {
  if (!g_default.ctx) {
    const char* d = getenv("SPWS_DEVICE");
    int dev = d ? atoi(d) : 0;
    int rc = spws_ctx_create(dev, nullptr, &g_default.ctx);
    if (rc) return rc;
  }
  *out = g_default.ctx;
  return SPWS_OK;
}

// Copy a host array into staging buffer `slot` of the context.
template <class T>
int stage_in(spws_ctx* ctx, int slot, const T* host, int64_t count, const T** dev) {
  int rc = ctx->stage[slot].ensure((size_t)(count > 0 ? count : 1) * sizeof(T));
  if (rc) return rc;
  if (count > 0) {
    SPWS_REQUIRE(host != nullptr, SPWS_ERR_INVALID, "null host array in taco tensor");
    SPWS_HIP(hipMemcpyAsync(ctx->stage[slot].ptr, host, (size_t)count * sizeof(T), hipMemcpyHostToDevice,
                            ctx->stream));
  }
  *dev = ctx->stage[slot].as<T>();
  return SPWS_OK;
}

template <class T>
T* host_alloc(int64_t count) {
  return static_cast<T*>(malloc((size_t)(count > 0 ? count : 1) * sizeof(T)));
}

static int check_csr(const taco_tensor_t* t, const char* name) {
  SPWS_REQUIRE(t != nullptr, SPWS_ERR_INVALID, "%s is null", name);
  SPWS_REQUIRE(t->order == 2 && t->dimensions && t->indices && t->indices[1], SPWS_ERR_INVALID,
               "%s is not a 2-level (CSR) taco tensor", name);
  SPWS_REQUIRE(t->indices[1][0] != nullptr, SPWS_ERR_INVALID, "%s has no pos array", name);
  return SPWS_OK;
}

struct StagedCsr {
  int32_t rows, cols;
  int64_t nnz;
  const int32_t* pos;
  const int32_t* crd;
  const double* vals;
};

static int stage_csr(spws_ctx* ctx, const taco_tensor_t* t, int slot0, bool with_vals, StagedCsr* out) {
  out->rows = t->dimensions[0];
  out->cols = t->dimensions[1];
  const int32_t* hp = t->indices[1][0];
  out->nnz = hp[out->rows];
  int rc = stage_in(ctx, slot0, hp, (int64_t)out->rows + 1, &out->pos);
  if (rc) return rc;
  rc = stage_in(ctx, slot0 + 1, t->indices[1][1], out->nnz, &out->crd);
  if (rc) return rc;
  out->vals = nullptr;
  if (with_vals) rc = stage_in(ctx, slot0 + 2, t->vals, out->nnz, &out->vals);
  return rc;
}

// "CC" (DCSR / DCSC) taco tensor: indices[0][0] = {0, n_outer},
// indices[0][1] = outer ids, indices[1][0] = pos, indices[1][1] = crd.
struct StagedCC {
  int32_t nouter;
  int64_t nnz;
  const int32_t *l1, *pos, *crd;
  const double* vals;
};

static int stage_cc(spws_ctx* ctx, const taco_tensor_t* t, const char* name, int slot0, StagedCC* out) {
  SPWS_REQUIRE(t && t->indices && t->indices[0] && t->indices[1] && t->indices[0][0] && t->indices[1][0],
               SPWS_ERR_INVALID, "%s is not a compressed-compressed taco tensor", name);
  out->nouter = t->indices[0][0][1] - t->indices[0][0][0];
  const int32_t* hpos = t->indices[1][0];
  out->nnz = hpos[out->nouter];
  int rc;
  if ((rc = stage_in(ctx, slot0, t->indices[0][1], out->nouter, &out->l1)) ||
      (rc = stage_in(ctx, slot0 + 1, hpos, (int64_t)out->nouter + 1, &out->pos)) ||
      (rc = stage_in(ctx, slot0 + 2, t->indices[1][1], out->nnz, &out->crd)) ||
      (rc = stage_in(ctx, slot0 + 3, t->vals, out->nnz, &out->vals)))
    return rc;
  return SPWS_OK;
}

// 3-level CSF taco tensor (levels 0, 1, 2 each with pos/crd; root pos {0, n}).
struct HostCsf {
  int32_t nroots, nfibers;
  int64_t nnz;
};

static int stage_csf(spws_ctx* ctx, const taco_tensor_t* t, const char* name, HostCsf* h, const int32_t** l1,
                     const int32_t** l2p, const int32_t** l2c, const int32_t** l3p, const int32_t** l3c,
                     const double** v) {
  SPWS_REQUIRE(t && t->order == 3 && t->indices, SPWS_ERR_INVALID, "%s is not a 3-level CSF taco tensor", name);
  for (int l = 0; l < 3; ++l)
    SPWS_REQUIRE(t->indices[l] && t->indices[l][0], SPWS_ERR_INVALID, "%s level %d has no pos array", name, l);
  h->nroots = t->indices[0][0][1] - t->indices[0][0][0];
  h->nfibers = t->indices[1][0][h->nroots];
  h->nnz = t->indices[2][0][h->nfibers];
  int rc;
  if ((rc = stage_in(ctx, 0, t->indices[0][1], h->nroots, l1)) ||
      (rc = stage_in(ctx, 1, t->indices[1][0], (int64_t)h->nroots + 1, l2p)) ||
      (rc = stage_in(ctx, 2, t->indices[1][1], h->nfibers, l2c)) ||
      (rc = stage_in(ctx, 3, t->indices[2][0], (int64_t)h->nfibers + 1, l3p)) ||
      (rc = stage_in(ctx, 4, t->indices[2][1], h->nnz, l3c)) || (rc = stage_in(ctx, 5, t->vals, h->nnz, v)))
    return rc;
  return SPWS_OK;
}

// A DCSR factor (rows x R): when every row and column is stored, *vals is
// its value array and *full is set; otherwise it is copied into `dense`
// (rows x R, zero where not stored) with `mask` holding each row's stored
// columns as bits (a missing row has mask 0), and *vals = dense.data().
static int factor_dense(const taco_tensor_t* t, const char* name, int32_t R, std::vector<double>& dense,
                        std::vector<uint64_t>& mask, bool* full, const double** vals) {
  SPWS_REQUIRE(t && t->dimensions && t->indices && t->indices[0] && t->indices[1] && t->indices[0][0] &&
                   t->indices[1][0] && (t->vals || t->indices[0][0][1] == t->indices[0][0][0]),
               SPWS_ERR_INVALID, "%s is not a compressed-compressed taco tensor", name);
  SPWS_REQUIRE(t->dimensions[1] == R, SPWS_ERR_INVALID, "%s has %d columns, not R = %d", name, t->dimensions[1], R);
  const int32_t rows_dim = t->dimensions[0];
  const int32_t nr = t->indices[0][0][1] - t->indices[0][0][0];
  const int32_t* rows = t->indices[0][1];
  const int32_t* pos = t->indices[1][0];
  const int32_t* crd = t->indices[1][1];
  SPWS_REQUIRE(nr >= 0 && nr <= rows_dim && (nr == 0 || (rows && crd)), SPWS_ERR_INVALID, "%s: malformed levels",
               name);
  bool dense_all = nr == rows_dim;
  for (int32_t r = 0; r < nr && dense_all; ++r) {
    dense_all = rows[r] == r && pos[r + 1] - pos[r] == R;
    for (int32_t q = 0; q < R && dense_all; ++q) dense_all = crd[pos[r] + q] == q;
  }
  *full = dense_all;
  if (dense_all) {
    *vals = t->vals;
    return SPWS_OK;
  }
  SPWS_REQUIRE(R <= 64, SPWS_ERR_UNSUPPORTED, "%s: a factor with rows or columns not stored needs R <= 64", name);
  dense.assign((size_t)rows_dim * R, 0.0);
  mask.assign((size_t)(rows_dim > 0 ? rows_dim : 1), 0);
  for (int32_t r = 0; r < nr; ++r) {
    const int32_t row = rows[r];
    SPWS_REQUIRE(row >= 0 && row < rows_dim && (r == 0 || rows[r - 1] < row), SPWS_ERR_INVALID,
                 "%s: stored rows must be increasing and within the dimension", name);
    for (int32_t q = pos[r]; q < pos[r + 1]; ++q) {
      const int32_t c = crd[q];
      SPWS_REQUIRE(c >= 0 && c < R && (q == pos[r] || crd[q - 1] < c), SPWS_ERR_INVALID,
                   "%s: a row's columns must be increasing and below R", name);
      dense[(size_t)row * R + c] = t->vals[q];
      mask[row] |= (uint64_t)1 << c;
    }
  }
  *vals = dense.data();
  return SPWS_OK;
}

template <class T>
static int d2h_alloc(spws_ctx* ctx, const T* dev, int64_t count, T** out) {
  T* h = host_alloc<T>(count);
  if (!h) {
    set_error("host allocation failed");
    return SPWS_ERR_OOM;
  }
  if (count > 0) {
    hipError_t e = hipMemcpyAsync(h, dev, (size_t)count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      free(h);
      return fail_hip(e, "device-to-host copy", __FILE__, __LINE__);
    }
  }
  *out = h;
  return SPWS_OK;
}

}  // namespace spws

using namespace spws;

extern "C" {

int spws_default_ctx(spws_ctx** out) {
  SPWS_REQUIRE(out != nullptr, SPWS_ERR_INVALID, "null output pointer");
  return default_ctx(out);
}

int spws_csr_csr_hash_compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap) {
  (void)w_cap;
  int rc;
  if ((rc = check_csr(A, "A")) || (rc = check_csr(B, "B"))) return rc;
  SPWS_REQUIRE(C && C->dimensions && C->indices && C->indices[1], SPWS_ERR_INVALID,
               "C must be initialised as a 2-level tensor (init_taco_tensor_DC)");
  SPWS_REQUIRE(A->dimensions[1] == B->dimensions[0], SPWS_ERR_INVALID, "inner dimensions differ (%d vs %d)",
               A->dimensions[1], B->dimensions[0]);
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx))) return rc;
  if ((rc = ctx_activate(ctx))) return rc;
  StagedCsr a, b;
  if ((rc = stage_csr(ctx, A, 0, true, &a)) || (rc = stage_csr(ctx, B, 3, true, &b))) return rc;
  int64_t nnz = 0;
  int32_t *dpos, *dcrd;
  double* dval;
  rc = spws_spgemm_rowwise(ctx, a.rows, a.cols, b.cols, a.pos, a.crd, a.vals, b.pos, b.crd, b.vals, &nnz, &dpos,
                           &dcrd, &dval);
  if (rc) return rc;
  const int32_t m = a.rows;
  int32_t* hp = host_alloc<int32_t>((int64_t)m + 1);
  int32_t* hc = host_alloc<int32_t>(nnz);
  double* hv = host_alloc<double>(nnz);
  if (!hp || !hc || !hv) {
    free(hp);
    free(hc);
    free(hv);
    set_error("host allocation of the output failed");
    return SPWS_ERR_OOM;
  }
  hipError_t e = hipMemcpyAsync(hp, dpos, ((size_t)m + 1) * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && nnz) e = hipMemcpyAsync(hc, dcrd, (size_t)nnz * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && nnz) e = hipMemcpyAsync(hv, dval, (size_t)nnz * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    free(hp);
    free(hc);
    free(hv);
    return fail_hip(e, "copy of the SpGEMM output", __FILE__, __LINE__);
  }
  C->indices[1][0] = hp;
  C->indices[1][1] = hc;
  C->vals = hv;
  C->vals_size = (int32_t)nnz;
  return SPWS_OK;
}

int spws_csr_csr_assemble(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B) {
  int rc;
  if ((rc = check_csr(A, "A")) || (rc = check_csr(B, "B"))) return rc;
  SPWS_REQUIRE(C && C->dimensions && C->indices && C->indices[1], SPWS_ERR_INVALID,
               "C must be initialised as a 2-level tensor (init_taco_tensor_DC)");
  SPWS_REQUIRE(A->dimensions[1] == B->dimensions[0], SPWS_ERR_INVALID, "inner dimensions differ");
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx))) return rc;
  if ((rc = ctx_activate(ctx))) return rc;
  StagedCsr a, b;
  if ((rc = stage_csr(ctx, A, 0, false, &a)) || (rc = stage_csr(ctx, B, 3, false, &b))) return rc;
  int64_t nnz = 0;
  int32_t* dpos;
  rc = spws_spgemm_rowwise_symbolic(ctx, a.rows, a.cols, b.cols, a.pos, a.crd, b.pos, b.crd, &nnz, &dpos);
  if (rc) return rc;
  const int32_t m = a.rows;
  int32_t* hp = host_alloc<int32_t>((int64_t)m + 1);
  int32_t* hc = host_alloc<int32_t>(nnz);
  double* hv = host_alloc<double>(nnz);
  if (!hp || !hc || !hv) {
    free(hp);
    free(hc);
    free(hv);
    set_error("host allocation of the output failed");
    return SPWS_ERR_OOM;
  }
  hipError_t e = hipMemcpyAsync(hp, dpos, ((size_t)m + 1) * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    free(hp);
    free(hc);
    free(hv);
    return fail_hip(e, "copy of C2_pos", __FILE__, __LINE__);
  }
  C->indices[1][0] = hp;
  C->indices[1][1] = hc;
  C->vals = hv;
  C->vals_size = (int32_t)nnz;
  return SPWS_OK;
}

// The taco entry points stand in for the reference's compute(): their sums
// stay reproducible run to run, so an automatic outer-product method is the
// ordered fused one there (not the atomic-add variant).
struct OrderedOuterAuto {
  spws_ctx* ctx;
  bool saved;
  explicit OrderedOuterAuto(spws_ctx* c) : ctx(c), saved(c->outer_auto_ordered) { ctx->outer_auto_ordered = true; }
  ~OrderedOuterAuto() { ctx->outer_auto_ordered = saved; }
};

int spws_csr_csr_compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B) {
  int rc;
  if ((rc = check_csr(A, "A")) || (rc = check_csr(B, "B"))) return rc;
  SPWS_REQUIRE(C && C->indices && C->indices[1] && C->indices[1][0] && C->indices[1][1] && C->vals,
               SPWS_ERR_INVALID, "compute() needs the arrays assemble() allocated");
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx))) return rc;
  if ((rc = ctx_activate(ctx))) return rc;
  StagedCsr a, b;
  if ((rc = stage_csr(ctx, A, 0, true, &a)) || (rc = stage_csr(ctx, B, 3, true, &b))) return rc;
  int64_t nnz = 0;
  int32_t *dpos, *dcrd;
  double* dval;
  rc = spws_spgemm_rowwise(ctx, a.rows, a.cols, b.cols, a.pos, a.crd, a.vals, b.pos, b.crd, b.vals, &nnz, &dpos,
                           &dcrd, &dval);
  if (rc) return rc;
  const int32_t* hp = C->indices[1][0];
  SPWS_REQUIRE(hp[a.rows] == nnz, SPWS_ERR_INVALID, "C2_pos from assemble() holds %d entries, product has %lld",
               hp[a.rows], (long long)nnz);
  SPWS_HIP(hipMemcpyAsync(C->indices[1][1], dcrd, (size_t)nnz * 4, hipMemcpyDeviceToHost, ctx->stream));
  SPWS_HIP(hipMemcpyAsync(C->vals, dval, (size_t)nnz * 8, hipMemcpyDeviceToHost, ctx->stream));
  SPWS_HIP(hipStreamSynchronize(ctx->stream));
  return SPWS_OK;
}

int spws_dcsc_dcsr_coord_compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap) {
  (void)w_cap;
  SPWS_REQUIRE(C && C->dimensions && C->indices && C->indices[1], SPWS_ERR_INVALID,
               "C must be initialised as a 2-level tensor (init_taco_tensor_DC)");
  SPWS_REQUIRE(A && B && A->dimensions && B->dimensions, SPWS_ERR_INVALID, "null operand");
  int rc;
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx)) || (rc = ctx_activate(ctx))) return rc;
  StagedCC a, b;
  if ((rc = stage_cc(ctx, A, "A", 0, &a)) || (rc = stage_cc(ctx, B, "B", 4, &b))) return rc;
  const int32_t m = C->dimensions[0], n = C->dimensions[1];
  int64_t nnz = 0;
  int64_t* dpos;
  int32_t* dcrd;
  double* dval;
  {
    OrderedOuterAuto ordered(ctx);
    rc = spws_spgemm_outer(ctx, m, A->dimensions[1], n, a.nouter, a.l1, a.pos, a.crd, a.vals, b.nouter, b.l1, b.pos,
                           b.crd, b.vals, &nnz, &dpos, &dcrd, &dval);
  }
  if (rc) return rc;
  SPWS_REQUIRE(nnz <= INT32_MAX, SPWS_ERR_UNSUPPORTED, "nnz(C) = %lld does not fit the int32 CSR of taco_tensor_t",
               (long long)nnz);
  int64_t* p64 = nullptr;
  int32_t* hc = nullptr;
  double* hv = nullptr;
  if ((rc = d2h_alloc(ctx, dpos, (int64_t)m + 1, &p64))) return rc;
  int32_t* hp = host_alloc<int32_t>((int64_t)m + 1);
  if (!hp) {
    free(p64);
    set_error("host allocation failed");
    return SPWS_ERR_OOM;
  }
  for (int32_t i = 0; i <= m; ++i) hp[i] = (int32_t)p64[i];
  free(p64);
  if ((rc = d2h_alloc(ctx, dcrd, nnz, &hc)) || (rc = d2h_alloc(ctx, dval, nnz, &hv))) {
    free(hp);
    free(hc);
    return rc;
  }
  C->indices[1][0] = hp;
  C->indices[1][1] = hc;
  C->vals = hv;
  C->vals_size = (int32_t)nnz;
  return SPWS_OK;
}

// An outer-product result (int64 row pointers on the device) as the
// malloc'd int32 CSR arrays of C.
static int outer_result_to_taco(spws_ctx* ctx, taco_tensor_t* C, int32_t rows, int64_t nnz, const int64_t* dpos,
                                const int32_t* dcrd, const double* dval) {
  int rc;
  SPWS_REQUIRE(nnz <= INT32_MAX, SPWS_ERR_UNSUPPORTED, "nnz(C) = %lld does not fit the int32 CSR of taco_tensor_t",
               (long long)nnz);
  int64_t* p64 = nullptr;
  int32_t* hc = nullptr;
  double* hv = nullptr;
  if ((rc = d2h_alloc(ctx, dpos, (int64_t)rows + 1, &p64))) return rc;
  int32_t* hp = host_alloc<int32_t>((int64_t)rows + 1);
  if (!hp) {
    free(p64);
    set_error("host allocation failed");
    return SPWS_ERR_OOM;
  }
  for (int32_t i = 0; i <= rows; ++i) hp[i] = (int32_t)p64[i];
  free(p64);
  if ((rc = d2h_alloc(ctx, dcrd, nnz, &hc)) || (rc = d2h_alloc(ctx, dval, nnz, &hv))) {
    free(hp);
    free(hc);
    return rc;
  }
  C->indices[1][0] = hp;
  C->indices[1][1] = hc;
  C->vals = hv;
  C->vals_size = (int32_t)nnz;
  return SPWS_OK;
}

int spws_csc_csr_t_hash_compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap) {
  (void)w_cap;
  SPWS_REQUIRE(C && C->dimensions && C->indices && C->indices[1], SPWS_ERR_INVALID,
               "C must be initialised as a 2-level tensor (init_taco_tensor_DC)");
  SPWS_REQUIRE(A && B && A->dimensions && B->dimensions && A->indices && B->indices && A->indices[1] &&
                   B->indices[1] && A->indices[1][0] && B->indices[1][0],
               SPWS_ERR_INVALID, "A must be CSC and B CSR taco tensors");
  // C (n x m) = (A B)^T = B^T A^T: B's rows j are the stored columns of the
  // left operand B^T, A's columns j the stored rows of the right operand A^T;
  // both outer levels are dense (every j)
  const int32_t n = C->dimensions[0], m = C->dimensions[1], k = B->dimensions[0];
  SPWS_REQUIRE(B->dimensions[1] == n && A->dimensions[0] == m && A->dimensions[1] == k, SPWS_ERR_INVALID,
               "dimensions: A %d x %d (CSC), B %d x %d, C must be %d x %d", A->dimensions[0], A->dimensions[1],
               B->dimensions[0], B->dimensions[1], B->dimensions[1], A->dimensions[0]);
  int rc;
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx)) || (rc = ctx_activate(ctx))) return rc;
  const int64_t annz = A->indices[1][0][k], bnnz = B->indices[1][0][k];
  const int32_t *apos, *acrd, *bpos, *bcrd;
  const double *aval, *bval;
  if ((rc = stage_in(ctx, 0, A->indices[1][0], (int64_t)k + 1, &apos)) ||
      (rc = stage_in(ctx, 1, A->indices[1][1], annz, &acrd)) || (rc = stage_in(ctx, 2, A->vals, annz, &aval)) ||
      (rc = stage_in(ctx, 3, B->indices[1][0], (int64_t)k + 1, &bpos)) ||
      (rc = stage_in(ctx, 4, B->indices[1][1], bnnz, &bcrd)) || (rc = stage_in(ctx, 5, B->vals, bnnz, &bval)))
    return rc;
  int64_t nnz = 0;
  int64_t* dpos;
  int32_t* dcrd;
  double* dval;
  {
    OrderedOuterAuto ordered(ctx);
    rc = spws_spgemm_outer(ctx, n, k, m, k, nullptr, bpos, bcrd, bval, k, nullptr, apos, acrd, aval, &nnz, &dpos, &dcrd,
                           &dval);
  }
  if (rc) return rc;
  return outer_result_to_taco(ctx, C, n, nnz, dpos, dcrd, dval);
}

int spws_csc_csr_nofuse_compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap) {
  (void)w_cap;
  SPWS_REQUIRE(C && C->dimensions && C->indices && C->indices[1], SPWS_ERR_INVALID,
               "C must be initialised as a 2-level tensor (init_taco_tensor_DC)");
  SPWS_REQUIRE(A && B && A->dimensions && B->dimensions && A->indices && B->indices && A->indices[1] &&
                   B->indices[1] && A->indices[1][0] && B->indices[1][0],
               SPWS_ERR_INVALID, "A must be CSC and B CSR taco tensors");
  // C (m x n) = A B: A's columns j (its stored level) are the stored
  // columns of the left operand, B's rows j those of the right; both outer
  // levels are dense (every j)
  const int32_t m = C->dimensions[0], n = C->dimensions[1], k = B->dimensions[0];
  SPWS_REQUIRE(B->dimensions[1] == n && A->dimensions[0] == m && A->dimensions[1] == k, SPWS_ERR_INVALID,
               "dimensions: A %d x %d (CSC), B %d x %d, C must be %d x %d", A->dimensions[0], A->dimensions[1],
               B->dimensions[0], B->dimensions[1], A->dimensions[0], B->dimensions[1]);
  int rc;
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx)) || (rc = ctx_activate(ctx))) return rc;
  const int64_t annz = A->indices[1][0][k], bnnz = B->indices[1][0][k];
  const int32_t *apos, *acrd, *bpos, *bcrd;
  const double *aval, *bval;
  if ((rc = stage_in(ctx, 0, A->indices[1][0], (int64_t)k + 1, &apos)) ||
      (rc = stage_in(ctx, 1, A->indices[1][1], annz, &acrd)) || (rc = stage_in(ctx, 2, A->vals, annz, &aval)) ||
      (rc = stage_in(ctx, 3, B->indices[1][0], (int64_t)k + 1, &bpos)) ||
      (rc = stage_in(ctx, 4, B->indices[1][1], bnnz, &bcrd)) || (rc = stage_in(ctx, 5, B->vals, bnnz, &bval)))
    return rc;
  int64_t nnz = 0;
  int64_t* dpos;
  int32_t* dcrd;
  double* dval;
  {
    OrderedOuterAuto ordered(ctx);
    rc = spws_spgemm_outer(ctx, m, k, n, k, nullptr, apos, acrd, aval, k, nullptr, bpos, bcrd, bval, &nnz, &dpos, &dcrd,
                           &dval);
  }
  if (rc) return rc;
  return outer_result_to_taco(ctx, C, m, nnz, dpos, dcrd, dval);
}

int spws_csr_transpose(taco_tensor_t* A, taco_tensor_t* B) {
  // A = B^T: A has B's columns as rows (A->dimensions[0] of them)
  int rc = check_csr(B, "B");
  if (rc) return rc;
  SPWS_REQUIRE(A && A->dimensions && A->indices && A->indices[1], SPWS_ERR_INVALID,
               "A must be initialised as a 2-level tensor (init_taco_tensor_DC)");
  SPWS_REQUIRE(A->dimensions[0] == B->dimensions[1] && A->dimensions[1] == B->dimensions[0], SPWS_ERR_INVALID,
               "A (%d x %d) must have B's transposed shape (%d x %d)", A->dimensions[0], A->dimensions[1],
               B->dimensions[1], B->dimensions[0]);
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx)) || (rc = ctx_activate(ctx))) return rc;
  const int32_t m = B->dimensions[0], n = A->dimensions[0];
  const int64_t nnz = B->indices[1][0][m];
  const int32_t *bpos, *bcrd;
  const double* bval;
  if ((rc = stage_in(ctx, 0, B->indices[1][0], (int64_t)m + 1, &bpos)) ||
      (rc = stage_in(ctx, 1, B->indices[1][1], nnz, &bcrd)) || (rc = stage_in(ctx, 2, B->vals, nnz, &bval)))
    return rc;
  const size_t nn = (size_t)(nnz > 0 ? nnz : 1);
  if ((rc = ctx->tr_out[0].ensure(((size_t)n + 1) * 4)) || (rc = ctx->tr_out[1].ensure(nn * 4)) ||
      (rc = ctx->tr_out[2].ensure(nn * 8)))
    return rc;
  int32_t* dpos = ctx->tr_out[0].as<int32_t>();
  int32_t* dcrd = ctx->tr_out[1].as<int32_t>();
  double* dval = ctx->tr_out[2].as<double>();
  if ((rc = spws_csr_transpose_dev(ctx, m, n, nnz, bpos, bcrd, bval, dpos, dcrd, dval))) return rc;
  int32_t *hp = nullptr, *hc = nullptr;
  double* hv = nullptr;
  if ((rc = d2h_alloc(ctx, (const int32_t*)dpos, (int64_t)n + 1, &hp))) return rc;
  if ((rc = d2h_alloc(ctx, (const int32_t*)dcrd, nnz, &hc)) || (rc = d2h_alloc(ctx, (const double*)dval, nnz, &hv))) {
    free(hp);
    free(hc);
    return rc;
  }
  A->indices[1][0] = hp;
  A->indices[1][1] = hc;
  A->vals = hv;
  A->vals_size = (int32_t)nnz;
  return SPWS_OK;
}

int spws_mttkrp_hash_compute(taco_tensor_t* A, taco_tensor_t* B, taco_tensor_t* C, taco_tensor_t* D, int32_t w_cap) {
  (void)w_cap;
  SPWS_REQUIRE(A && A->dimensions && A->indices && A->indices[1] && A->indices[2], SPWS_ERR_INVALID,
               "A must be initialised as a 2-mode COO tensor (init_taco_tensor_COO)");
  SPWS_REQUIRE(C && D && C->dimensions && D->dimensions, SPWS_ERR_INVALID, "null factor");
  const int32_t I = A->dimensions[0], R = A->dimensions[1];
  const int32_t J = C->dimensions[0], K = D->dimensions[0];
  SPWS_REQUIRE(C->dimensions[1] == R && D->dimensions[1] == R, SPWS_ERR_INVALID, "factor ranks differ from A");
  int rc;
  std::vector<double> denseC, denseD;
  std::vector<uint64_t> maskC, maskD;
  bool fullC, fullD;
  const double *hC, *hD;
  if ((rc = factor_dense(C, "C", R, denseC, maskC, &fullC, &hC)) ||
      (rc = factor_dense(D, "D", R, denseD, maskD, &fullD, &hD)))
    return rc;
  const bool masked = !fullC || !fullD;
  const uint64_t all = R == 64 ? ~(uint64_t)0 : (((uint64_t)1 << R) - 1);
  if (masked) {  // a dense factor next to a sparse one: every row stores every column
    if (fullC) maskC.assign((size_t)(J > 0 ? J : 1), all);
    if (fullD) maskD.assign((size_t)(K > 0 ? K : 1), all);
  }
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx)) || (rc = ctx_activate(ctx))) return rc;
  HostCsf h;
  const int32_t *l1, *l2p, *l2c, *l3p, *l3c;
  const double *v, *dC, *dD;
  const uint64_t *dmC = nullptr, *dmD = nullptr;
  if ((rc = stage_csf(ctx, B, "B", &h, &l1, &l2p, &l2c, &l3p, &l3c, &v)) ||
      (rc = stage_in(ctx, 6, hC, (int64_t)J * R, &dC)) || (rc = stage_in(ctx, 7, hD, (int64_t)K * R, &dD)))
    return rc;
  if (masked && ((rc = stage_in(ctx, 8, maskC.data(), (int64_t)maskC.size(), &dmC)) ||
                 (rc = stage_in(ctx, 9, maskD.data(), (int64_t)maskD.size(), &dmD))))
    return rc;
  int32_t nrows = 0;
  int32_t* drows;
  uint64_t* dmask;
  double* dM;
  rc = spws_mttkrp_masked(ctx, I, J, K, R, h.nroots, l1, l2p, l2c, h.nfibers, l3p, l3c, v, h.nnz, dC, dD, dmC, dmD,
                          &nrows, &drows, &dmask, &dM);
  if (rc) return rc;
  int32_t* rows = nullptr;
  uint64_t* rmask = nullptr;
  double* M = nullptr;
  if ((rc = d2h_alloc(ctx, drows, nrows, &rows)) || (masked && (rc = d2h_alloc(ctx, (const uint64_t*)dmask, nrows,
                                                                              &rmask))) ||
      (rc = d2h_alloc(ctx, dM, (int64_t)I * R, &M))) {
    free(rows);
    free(rmask);
    return rc;
  }
  // COO (i, r) in row order, the row's stored columns in increasing r
  int64_t n = 0;
  for (int32_t t = 0; t < nrows; ++t) n += masked ? __builtin_popcountll(rmask[t]) : R;
  int32_t* oi = host_alloc<int32_t>(n);
  int32_t* orr = host_alloc<int32_t>(n);
  double* ov = host_alloc<double>(n);
  if (!oi || !orr || !ov) {
    free(rows); free(rmask); free(M); free(oi); free(orr); free(ov);
    set_error("host allocation failed");
    return SPWS_ERR_OOM;
  }
  int64_t q = 0;
  for (int32_t t = 0; t < nrows; ++t)
    for (int32_t r = 0; r < R; ++r) {
      if (masked && !((rmask[t] >> r) & 1)) continue;
      oi[q] = rows[t];
      orr[q] = r;
      ov[q] = M[(int64_t)rows[t] * R + r];
      ++q;
    }
  free(rows);
  free(rmask);
  free(M);
  A->indices[1][0] = oi;
  A->indices[2][0] = orr;
  A->vals = ov;
  A->vals_size = (int32_t)n;
  return SPWS_OK;
}

int spws_ttm_hash_il_compute(taco_tensor_t* A, taco_tensor_t* B, taco_tensor_t* C, int32_t w_cap) {
  (void)w_cap;
  SPWS_REQUIRE(A && A->dimensions && A->indices && A->indices[1] && A->indices[2] && A->indices[3], SPWS_ERR_INVALID,
               "A must be initialised as a 3-mode COO tensor (init_taco_tensor_COO)");
  SPWS_REQUIRE(C && C->dimensions, SPWS_ERR_INVALID, "null factor");
  const int32_t I = A->dimensions[0], J = A->dimensions[1], R = A->dimensions[2];
  const int32_t K = C->dimensions[0];
  SPWS_REQUIRE(C->dimensions[1] == R, SPWS_ERR_INVALID, "factor columns differ from A's last dimension");
  int rc;
  std::vector<double> denseU;
  std::vector<uint64_t> maskU;
  bool fullU;
  const double* hU;
  if ((rc = factor_dense(C, "C", R, denseU, maskU, &fullU, &hU))) return rc;
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx)) || (rc = ctx_activate(ctx))) return rc;
  HostCsf h;
  const int32_t *l1, *l2p, *l2c, *l3p, *l3c;
  const double *v, *dU;
  const uint64_t* dmU = nullptr;
  if ((rc = stage_csf(ctx, B, "B", &h, &l1, &l2p, &l2c, &l3p, &l3c, &v)) ||
      (rc = stage_in(ctx, 6, hU, (int64_t)K * R, &dU)))
    return rc;
  if (!fullU && (rc = stage_in(ctx, 8, maskU.data(), (int64_t)maskU.size(), &dmU))) return rc;
  int64_t F = 0;
  int32_t *fi, *fj;
  uint64_t* fm;
  double* Y;
  rc = spws_ttm_masked(ctx, I, J, K, R, h.nroots, l1, l2p, l2c, h.nfibers, l3p, l3c, v, h.nnz, dU, dmU, &F, &fi, &fj,
                       &fm, &Y);
  if (rc) return rc;
  int32_t *hi = nullptr, *hj = nullptr;
  uint64_t* hm = nullptr;
  double* hy = nullptr;
  if ((rc = d2h_alloc(ctx, fi, F, &hi)) || (rc = d2h_alloc(ctx, fj, F, &hj)) ||
      (!fullU && (rc = d2h_alloc(ctx, (const uint64_t*)fm, F, &hm))) || (rc = d2h_alloc(ctx, Y, F * R, &hy))) {
    free(hi);
    free(hj);
    free(hm);
    return rc;
  }
  // COO (i, j, r) in fiber order, each fiber's stored columns in increasing r
  int64_t n = 0;
  for (int64_t f = 0; f < F; ++f) n += fullU ? R : __builtin_popcountll(hm[f]);
  SPWS_REQUIRE(n <= INT32_MAX, SPWS_ERR_UNSUPPORTED, "TTM output exceeds the int32 vals_size");
  int32_t* o1 = host_alloc<int32_t>(n);
  int32_t* o2 = host_alloc<int32_t>(n);
  int32_t* o3 = host_alloc<int32_t>(n);
  double* ov = fullU ? hy : host_alloc<double>(n);
  if (!o1 || !o2 || !o3 || !ov) {
    free(hi); free(hj); free(hm); free(o1); free(o2); free(o3);
    if (ov != hy) free(ov);
    free(hy);
    set_error("host allocation failed");
    return SPWS_ERR_OOM;
  }
  int64_t q = 0;
  for (int64_t f = 0; f < F; ++f)
    for (int32_t r = 0; r < R; ++r) {
      if (!fullU && !((hm[f] >> r) & 1)) continue;
      o1[q] = hi[f];
      o2[q] = hj[f];
      o3[q] = r;
      if (!fullU) ov[q] = hy[f * R + r];
      ++q;
    }
  free(hi);
  free(hj);
  free(hm);
  if (ov != hy) free(hy);
  A->indices[1][0] = o1;
  A->indices[2][0] = o2;
  A->indices[3][0] = o3;
  A->vals = ov;
  A->vals_size = (int32_t)n;
  return SPWS_OK;
}

}  // extern "C"
