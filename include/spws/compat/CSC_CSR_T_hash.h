/* Drop-in for benchmark/CSC_CSR_T/CSC_CSR_T_hash.h: the transposed-output
 * compute() (:232), C = (A B)^T with the workspace keyed on (k, i), and the
 * wrapper CSC_CSR_T_hash() (:352), on the GPU.  A is CSC, B CSR, C CSR
 * (n x m) with int32 pos. */
#ifndef SPWS_COMPAT_CSC_CSR_T_HASH_H
#define SPWS_COMPAT_CSC_CSR_T_HASH_H
#include "spws/compat/timing.h"

inline int compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap) {
  return spws_compat::report(spws_csc_csr_t_hash_compute(C, A, B, w_cap), "compute");
}

/* Seconds per compute() over `repeat` runs after `warmup` runs (wall clock;
 * the reference measures clock()).  With bench set, the outputs of all but
 * the last run are freed, as in the reference. */
inline double CSC_CSR_T_hash(taco_tensor_t* A, taco_tensor_t* B, taco_tensor_t* C, int32_t w_cap, int32_t warmup,
                             int32_t repeat, bool bench = false, bool print = false) {
  (void)print;
  for (int i = 0; i < warmup; i++) {
    compute(C, A, B, w_cap);
    if (bench) {
      std::free(C->vals);
      std::free(C->indices[1][0]);
      std::free(C->indices[1][1]);
    }
  }
  const double start = spws_compat::now_seconds();
  for (int i = 0; i < repeat; i++) {
    compute(C, A, B, w_cap);
    if (bench && i != repeat - 1) {
      std::free(C->vals);
      std::free(C->indices[1][0]);
      std::free(C->indices[1][1]);
    }
  }
  return (spws_compat::now_seconds() - start) / (repeat > 0 ? repeat : 1);
}

#endif
