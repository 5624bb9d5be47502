// Sparse tensor inputs shared by the MTTKRP and TTM drivers.
#pragma once
#include "common.hpp"

namespace spws {

// A 3-level CSF on the device: nroots root coordinates l1_crd, second level
// l2_pos[nroots + 1] / l2_crd[nfibers], leaves l3_pos[nfibers + 1] /
// l3_crd[nnz], values vals[nnz]; coordinates sorted within each parent.
struct Csf3 {
  int32_t nroots;
  const int32_t* l1_crd;
  const int32_t* l2_pos;
  const int32_t* l2_crd;
  int32_t nfibers;
  const int32_t* l3_pos;
  const int32_t* l3_crd;
  const double* vals;
  int64_t nnz;
};

// TTM by roots and column windows (ttm_roots.hip) for a CSF with levels
// (i, k, j): the output fibers (i, j) in order, out_i / out_j in
// ctx->out_crd (F each), Y (F x R) in ctx->out_val.  *handled is false when
// the shapes are outside what it supports, or (unless forced) when the
// heavy roots' column windows are too sparse for it to pay; the caller then
// takes the sorting path, which recomputes everything.
int ttm_roots_run(spws_ctx* ctx, const Csf3& x, int32_t I, int32_t J, int32_t K, int32_t R, const double* U,
                  bool forced, int64_t* nfib_out, bool* handled);

}  // namespace spws
