// TACO-shaped host entry points: host taco_tensor_t in, device pipeline,
// malloc'd host arrays out (so the caller's free() works, as with the
// reference's compute()).  Each thread gets its own default context on the
// device named by SPWS_DEVICE (default 0).
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

}  // namespace spws

using namespace spws;

extern "C" {

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

}  // extern "C"
