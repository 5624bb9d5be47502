/* Drop-in for benchmark/DCSC_DCSR/DCSC_DCSR_coord_index_chase.h: the
 * outer-product coordinate-workspace compute() (:194) and the wrapper
 * DCSC_DCSR_coord() (:312), on the GPU (expand, radix sort, segmented
 * reduce).  A is DCSC, B is DCSR, C is CSR with int32 pos. */
#ifndef SPWS_COMPAT_DCSC_DCSR_COORD_H
#define SPWS_COMPAT_DCSC_DCSR_COORD_H
#include "spws/compat/timing.h"

inline int compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap) {
  return spws_compat::report(spws_dcsc_dcsr_coord_compute(C, A, B, w_cap), "compute");
}

inline void DCSC_DCSR_coord(taco_tensor_t* A, taco_tensor_t* B, taco_tensor_t* C, int32_t w_cap,
                            bool print = false) {
  (void)print;
  compute(C, A, B, w_cap);
}

#endif
