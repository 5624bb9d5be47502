/* Drop-in for seperate/CSR_CSR_1.h: the TACO split assemble() (:69, the
 * symbolic phase filling C2_pos and allocating C2_crd / C_vals) and
 * compute() (:4, the numeric phase filling them), on the GPU. */
#ifndef SPWS_COMPAT_CSR_CSR_1_H
#define SPWS_COMPAT_CSR_CSR_1_H
#include "spws/compat/timing.h"

inline int assemble(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B) {
  return spws_compat::report(spws_csr_csr_assemble(C, A, B), "assemble");
}

inline int compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B) {
  return spws_compat::report(spws_csr_csr_compute(C, A, B), "compute");
}

#endif
