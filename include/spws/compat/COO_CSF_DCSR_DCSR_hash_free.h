/* Drop-in for benchmark/MTTKRP/COO_CSF_DCSR_DCSR_hash_free.h: MTTKRP
 * compute(A, B, C, D, w_cap) (:233) and the wrapper
 * COO_CSF_DCSR_DCSR_hash() (:369), on the GPU.  B is a 3-level CSF with the
 * output mode as leaf; C and D are dense factors stored as DCSR; A is the
 * COO output (indices[1][0] rows, indices[2][0] ranks). */
#ifndef SPWS_COMPAT_MTTKRP_HASH_H
#define SPWS_COMPAT_MTTKRP_HASH_H
#include "spws/compat/timing.h"

inline int compute(taco_tensor_t* A, taco_tensor_t* B, taco_tensor_t* C, taco_tensor_t* D,
                   int32_t w_accumulator_capacity) {
  return spws_compat::report(spws_mttkrp_hash_compute(A, B, C, D, w_accumulator_capacity), "compute");
}

inline double COO_CSF_DCSR_DCSR_hash(taco_tensor_t* A, taco_tensor_t* B, taco_tensor_t* C, taco_tensor_t* D,
                                     int32_t w_cap, int32_t warmup, int32_t repeat, bool bench = false,
                                     bool print = false) {
  (void)print;
  for (int i = 0; i < warmup; i++) {
    compute(A, B, C, D, w_cap);
    if (bench) {
      free(A->vals);
      free(A->indices[1][0]);
      free(A->indices[2][0]);
    }
  }
  const double start = spws_compat::now_seconds();
  for (int i = 0; i < repeat; i++) {
    compute(A, B, C, D, w_cap);
    if (bench && i != repeat - 1) {
      free(A->vals);
      free(A->indices[1][0]);
      free(A->indices[2][0]);
    }
  }
  return (spws_compat::now_seconds() - start) / (repeat > 0 ? repeat : 1);
}

#endif
