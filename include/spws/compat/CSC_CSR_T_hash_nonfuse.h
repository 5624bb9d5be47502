/* Drop-in for benchmark/CSC_CSR_T/nonfuse/CSC_CSR_T_hash.h: the non-fused
 * transposed-output product, compute_nofuse() (:232) C_noT = A B, then
 * transpose() (:352) C = C_noT^T, and the wrapper CSC_CSR_T_hash_nonfuse()
 * (:402), on the GPU.  A is CSC, B CSR, C_noT and C CSR with int32 pos. */
#ifndef SPWS_COMPAT_CSC_CSR_T_HASH_NONFUSE_H
#define SPWS_COMPAT_CSC_CSR_T_HASH_NONFUSE_H
#include "spws/compat/timing.h"

inline int compute_nofuse(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap) {
  return spws_compat::report(spws_csc_csr_nofuse_compute(C, A, B, w_cap), "compute_nofuse");
}

inline int transpose(taco_tensor_t* A, taco_tensor_t* B) {
  return spws_compat::report(spws_csr_transpose(A, B), "transpose");
}

/* Seconds per compute_nofuse() + transpose() over `repeat` runs after
 * `warmup` runs (wall clock; the reference measures clock()).  With bench
 * set, the outputs of all but the last run are freed, as in the reference. */
inline double CSC_CSR_T_hash_nonfuse(taco_tensor_t* A, taco_tensor_t* B, taco_tensor_t* C_noT, taco_tensor_t* C,
                                     int32_t w_cap, int32_t warmup, int32_t repeat, bool bench = false,
                                     bool print = false) {
  (void)print;
  auto release = [](taco_tensor_t* t) {
    std::free(t->vals);
    std::free(t->indices[1][0]);
    std::free(t->indices[1][1]);
  };
  for (int i = 0; i < warmup; i++) {
    compute_nofuse(C_noT, A, B, w_cap);
    transpose(C, C_noT);
    if (bench) {
      release(C_noT);
      release(C);
    }
  }
  const double start = spws_compat::now_seconds();
  for (int i = 0; i < repeat; i++) {
    compute_nofuse(C_noT, A, B, w_cap);
    transpose(C, C_noT);
    if (bench && i != repeat - 1) {
      release(C_noT);
      release(C);
    }
  }
  return (spws_compat::now_seconds() - start) / (repeat > 0 ? repeat : 1);
}

#endif
