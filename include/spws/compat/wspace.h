/* Drop-in for utils/wspace.h: the coordinate-workspace API with FP64
 * values.  init_wspace, esc_cmp / esc_cmp_rev, Enlarge and the printers are
 * the same host helpers; Sort (:206-217) and Merge (:62-204) run on the GPU
 * through libspws (radix sort of packed coordinates, segmented reduction). */
#ifndef SPWS_COMPAT_WSPACE_H
#define SPWS_COMPAT_WSPACE_H
#include <cstdlib>
#include <iostream>

#include "spws/compat/timing.h"

const int w_order = 2;

struct wspace {
  int32_t crd[2];
  double val;
};
static_assert(sizeof(wspace) == 16, "wspace must match the 16-byte device record");

inline bool init_wspace(wspace* w, int begin, int end) {
  for (int i = begin; i < end; i++) w[i].crd[0] = -1;
  return true;
}

inline int esc_cmp(const void* b, const void* a) {
  for (int i = 0; i < w_order; i++) {
    const int32_t x = ((const wspace*)b)->crd[i], y = ((const wspace*)a)->crd[i];
    if (x != y) return x < y ? -1 : 1;
  }
  return 0;
}

inline int esc_cmp_rev(const void* a, const void* b) { return esc_cmp(b, a); }

inline void print_wspace(wspace* w, const int len) {
  for (int i = 0; i < len; i++) std::cout << "(" << w[i].crd[0] << "," << w[i].crd[1] << "," << w[i].val << "),";
  std::cout << std::endl;
}

inline void print_coo(int32_t** C_COO_crd, double* C_COO_vals, const int len) {
  for (int i = 0; i < len; i++)
    std::cout << "(" << C_COO_crd[0][i] << "," << C_COO_crd[1][i] << "," << C_COO_vals[i] << "),";
  std::cout << std::endl;
}

inline int Enlarge(int32_t** C_COO_crd, double** C_COO_vals, int C_COO_capacity) {
  C_COO_capacity = C_COO_capacity * 2;
  for (int i = 0; i < w_order; i++)
    C_COO_crd[i] = (int32_t*)realloc(C_COO_crd[i], sizeof(int32_t) * C_COO_capacity);
  *C_COO_vals = (double*)realloc(*C_COO_vals, sizeof(double) * C_COO_capacity);
  return C_COO_capacity;
}

/* Merges the accumulator into the sorted COO (capacity >= COO_size +
 * accumulator_size); returns the merged size, -1 on a device error. */
inline int Merge(int32_t** COO_crd, double* COO_vals, int32_t COO_size, wspace* accumulator,
                 int32_t accumulator_size, bool rev = false) {
  (void)rev;  // the workspace is re-sorted on the device either way
  spws_ctx* ctx = nullptr;
  if (spws_compat::report(spws_default_ctx(&ctx), "Merge")) return -1;
  int64_t out = 0;
  if (spws_compat::report(spws_wspace_merge(ctx, COO_crd[0], COO_crd[1], COO_vals, COO_size, accumulator,
                                            accumulator_size, &out),
                          "Merge"))
    return -1;
  return (int)out;
}

/* Sorts size records ascending (rev = false: returns size) or descending
 * (rev = true: returns the count before the trailing empty slots). */
inline int Sort(void* array, size_t size, bool rev) {
  spws_ctx* ctx = nullptr;
  if (spws_compat::report(spws_default_ctx(&ctx), "Sort")) return -1;
  int64_t valid = 0;
  if (spws_compat::report(spws_wspace_sort(ctx, array, (int64_t)size, rev ? 1 : 0, &valid), "Sort")) return -1;
  return (int)valid;
}

#endif
