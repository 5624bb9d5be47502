// Reference-style driver over the drop-in non-fused transposed-output header
// (CSC_CSR_T/nonfuse/CSC_CSR_T_hash.h), the call site of the reference's
// memory driver unchanged, FP64.
//   drv_cscT <in> <out>: in  = dims{m,k,n}, A (CSC, m x k) pos/crd/vals, B (CSR, k x n) pos/crd/vals
//                        out = C = (A B)^T (CSR, n x m) pos/crd/vals
#include "frames.hpp"
#include "spws/compat/CSC_CSR_T_hash_nonfuse.h"

static taco_tensor_t* dc_tensor(int32_t d0, int32_t d1, int32_t outer, int32_t* pos, int32_t* crd, double* vals) {
  taco_tensor_t* t = (taco_tensor_t*)calloc(1, sizeof(taco_tensor_t));
  t->order = 2;
  t->dimensions = (int32_t*)malloc(2 * sizeof(int32_t));
  t->dimensions[0] = d0;
  t->dimensions[1] = d1;
  t->indices = (int32_t***)malloc(2 * sizeof(int32_t**));
  t->indices[0] = (int32_t**)malloc(sizeof(int32_t*));
  t->indices[1] = (int32_t**)malloc(2 * sizeof(int32_t*));
  t->indices[0][0] = (int32_t*)malloc(sizeof(int32_t));
  t->indices[0][0][0] = outer;
  t->indices[1][0] = pos;
  t->indices[1][1] = crd;
  t->vals = vals;
  return t;
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FrameReader in(argv[1]);
  FrameWriter out(argv[2]);
  int32_t* d = in.take<int32_t>();
  int32_t* ap = in.take<int32_t>();
  int32_t* ac = in.take<int32_t>();
  double* av = in.take<double>();
  int32_t* bp = in.take<int32_t>();
  int32_t* bc = in.take<int32_t>();
  double* bv = in.take<double>();
  const int32_t m = d[0], k = d[1], n = d[2];
  // A: logical m x k stored by column (outer level over k); B: k x n by row
  taco_tensor_t* A = dc_tensor(m, k, k, ap, ac, av);
  taco_tensor_t* B = dc_tensor(k, n, k, bp, bc, bv);
  taco_tensor_t* C_noT = dc_tensor(m, n, m, nullptr, nullptr, nullptr);
  taco_tensor_t* C = dc_tensor(n, m, n, nullptr, nullptr, nullptr);
  CSC_CSR_T_hash_nonfuse(A, B, C_noT, C, 1024, 0, 1, false, false);
  if (!C->indices[1][0]) return 1;
  const int32_t nnz = C->indices[1][0][n];
  out.put(C->indices[1][0], (int64_t)n + 1);
  out.put(C->indices[1][1], nnz);
  out.put(C->vals, nnz);
  return 0;
}
