// TACO-shaped host entry points: host taco_tensor_t in, device pipeline,
// malloc'd host arrays out (so the caller's free() works, as with the
// reference's compute()).  Each thread gets its own default context on the
// device named by SPWS_DEVICE (default 0).
#include <climits>
#include <cstdlib>

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

// A dense factor stored as DCSR: every row present, each row exactly the
// columns 0..R-1 in order.  Returns the row-major values (host pointer).
static int dense_factor(const taco_tensor_t* t, const char* name, int32_t R, const double** vals) {
  SPWS_REQUIRE(t && t->indices && t->indices[0] && t->indices[1] && t->indices[0][0] && t->indices[1][0] && t->vals,
               SPWS_ERR_INVALID, "%s is not a compressed-compressed taco tensor", name);
  const int32_t nr = t->indices[0][0][1] - t->indices[0][0][0];
  SPWS_REQUIRE(nr == t->dimensions[0], SPWS_ERR_UNSUPPORTED, "%s: only dense factor matrices are supported", name);
  const int32_t* rows = t->indices[0][1];
  const int32_t* pos = t->indices[1][0];
  const int32_t* crd = t->indices[1][1];
  for (int32_t r = 0; r < nr; ++r) {
    SPWS_REQUIRE(rows[r] == r && pos[r + 1] - pos[r] == R, SPWS_ERR_UNSUPPORTED,
                 "%s: only dense factor matrices are supported", name);
    for (int32_t q = 0; q < R; ++q)
      SPWS_REQUIRE(crd[pos[r] + q] == q, SPWS_ERR_UNSUPPORTED, "%s: only dense factor matrices are supported", name);
  }
  *vals = t->vals;
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
  rc = spws_spgemm_outer(ctx, m, A->dimensions[1], n, a.nouter, a.l1, a.pos, a.crd, a.vals, b.nouter, b.l1, b.pos,
                         b.crd, b.vals, &nnz, &dpos, &dcrd, &dval);
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
  rc = spws_spgemm_outer(ctx, n, k, m, k, nullptr, bpos, bcrd, bval, k, nullptr, apos, acrd, aval, &nnz, &dpos, &dcrd,
                         &dval);
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
  rc = spws_spgemm_outer(ctx, m, k, n, k, nullptr, apos, acrd, aval, k, nullptr, bpos, bcrd, bval, &nnz, &dpos, &dcrd,
                         &dval);
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
  const double *hC, *hD;
  if ((rc = dense_factor(C, "C", R, &hC)) || (rc = dense_factor(D, "D", R, &hD))) return rc;
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx)) || (rc = ctx_activate(ctx))) return rc;
  HostCsf h;
  const int32_t *l1, *l2p, *l2c, *l3p, *l3c;
  const double *v, *dC, *dD;
  if ((rc = stage_csf(ctx, B, "B", &h, &l1, &l2p, &l2c, &l3p, &l3c, &v)) ||
      (rc = stage_in(ctx, 6, hC, (int64_t)J * R, &dC)) || (rc = stage_in(ctx, 7, hD, (int64_t)K * R, &dD)))
    return rc;
  int32_t nrows = 0;
  int32_t* drows;
  double* dM;
  rc = spws_mttkrp(ctx, I, J, K, R, h.nroots, l1, l2p, l2c, h.nfibers, l3p, l3c, v, h.nnz, dC, dD, &nrows, &drows,
                   &dM);
  if (rc) return rc;
  int32_t* rows = nullptr;
  double* M = nullptr;
  if ((rc = d2h_alloc(ctx, drows, nrows, &rows)) || (rc = d2h_alloc(ctx, dM, (int64_t)I * R, &M))) {
    free(rows);
    return rc;
  }
  const int64_t n = (int64_t)nrows * R;
  int32_t* oi = host_alloc<int32_t>(n);
  int32_t* orr = host_alloc<int32_t>(n);
  double* ov = host_alloc<double>(n);
  if (!oi || !orr || !ov) {
    free(rows); free(M); free(oi); free(orr); free(ov);
    set_error("host allocation failed");
    return SPWS_ERR_OOM;
  }
  for (int32_t t = 0; t < nrows; ++t)
    for (int32_t r = 0; r < R; ++r) {
      const int64_t q = (int64_t)t * R + r;
      oi[q] = rows[t];
      orr[q] = r;
      ov[q] = M[(int64_t)rows[t] * R + r];
    }
  free(rows);
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
  const double* hU;
  if ((rc = dense_factor(C, "C", R, &hU))) return rc;
  spws_ctx* ctx;
  if ((rc = default_ctx(&ctx)) || (rc = ctx_activate(ctx))) return rc;
  HostCsf h;
  const int32_t *l1, *l2p, *l2c, *l3p, *l3c;
  const double *v, *dU;
  if ((rc = stage_csf(ctx, B, "B", &h, &l1, &l2p, &l2c, &l3p, &l3c, &v)) ||
      (rc = stage_in(ctx, 6, hU, (int64_t)K * R, &dU)))
    return rc;
  int64_t F = 0;
  int32_t *fi, *fj;
  double* Y;
  rc = spws_ttm(ctx, I, J, K, R, h.nroots, l1, l2p, l2c, h.nfibers, l3p, l3c, v, h.nnz, dU, &F, &fi, &fj, &Y);
  if (rc) return rc;
  SPWS_REQUIRE(F * R <= INT32_MAX, SPWS_ERR_UNSUPPORTED, "TTM output exceeds the int32 vals_size");
  int32_t *hi = nullptr, *hj = nullptr;
  double* hy = nullptr;
  if ((rc = d2h_alloc(ctx, fi, F, &hi)) || (rc = d2h_alloc(ctx, fj, F, &hj)) || (rc = d2h_alloc(ctx, Y, F * R, &hy))) {
    free(hi);
    free(hj);
    return rc;
  }
  const int64_t n = F * R;
  int32_t* o1 = host_alloc<int32_t>(n);
  int32_t* o2 = host_alloc<int32_t>(n);
  int32_t* o3 = host_alloc<int32_t>(n);
  if (!o1 || !o2 || !o3) {
    free(hi); free(hj); free(hy); free(o1); free(o2); free(o3);
    set_error("host allocation failed");
    return SPWS_ERR_OOM;
  }
  for (int64_t f = 0; f < F; ++f)
    for (int32_t r = 0; r < R; ++r) {
      o1[f * R + r] = hi[f];
      o2[f * R + r] = hj[f];
      o3[f * R + r] = r;
    }
  free(hi);
  free(hj);
  A->indices[1][0] = o1;
  A->indices[2][0] = o2;
  A->indices[3][0] = o3;
  A->vals = hy;
  A->vals_size = (int32_t)n;
  return SPWS_OK;
}

}  // extern "C"
