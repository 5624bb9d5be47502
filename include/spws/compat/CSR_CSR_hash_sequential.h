/* Drop-in for benchmark/Parallel/CSR_CSR_hash_sequential.h (and the
 * identical benchmark/CSR_CSR/CSR_CSR_hash.h): the row-wise hash SpGEMM
 * compute() (:232) and its timing wrapper CSR_CSR_hash() (:355), running on
 * the GPU through libspws.  taco_tensor_t carries FP64 values. */
#ifndef SPWS_COMPAT_CSR_CSR_HASH_H
#define SPWS_COMPAT_CSR_CSR_HASH_H
#include "spws/compat/timing.h"

inline int compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_accumulator_capacity) {
  return spws_compat::report(spws_csr_csr_hash_compute(C, A, B, w_accumulator_capacity), "compute");
}

inline double CSR_CSR_hash(taco_tensor_t* A, taco_tensor_t* B, taco_tensor_t* C, int32_t w_cap, int32_t warmup,
                           int32_t repeat, bool bench = false, bool print = false) {
  (void)print;
  for (int i = 0; i < warmup; i++) {
    compute(C, A, B, w_cap);
    if (bench) {
      free(C->vals);
      free(C->indices[1][0]);
      free(C->indices[1][1]);
    }
  }
  const double start = spws_compat::now_seconds();
  for (int i = 0; i < repeat; i++) {
    compute(C, A, B, w_cap);
    if (bench && i != repeat - 1) {
      free(C->vals);
      free(C->indices[1][0]);
      free(C->indices[1][1]);
    }
  }
  return (spws_compat::now_seconds() - start) / (repeat > 0 ? repeat : 1);
}

#endif
