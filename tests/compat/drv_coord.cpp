// Reference-style driver over the drop-in outer-product header
// (DCSC_DCSR_coord_index_chase.h), unchanged call site, FP64.
//   drv_coord <in> <out>: in  = dims{m,k,n}, A (DCSC) ids/pos/crd/vals, B (DCSR) ids/pos/crd/vals
//                         out = C pos/crd/vals (CSR)
#include "frames.hpp"
#include "spws/compat/DCSC_DCSR_coord_index_chase.h"

static taco_tensor_t* cc_tensor(int32_t d0, int32_t d1, int32_t nouter, int32_t* ids, int32_t* pos, int32_t* crd,
                                double* vals) {
  taco_tensor_t* t = (taco_tensor_t*)calloc(1, sizeof(taco_tensor_t));
  t->order = 2;
  t->dimensions = (int32_t*)malloc(2 * sizeof(int32_t));
  t->dimensions[0] = d0;
  t->dimensions[1] = d1;
  t->indices = (int32_t***)malloc(2 * sizeof(int32_t**));
  t->indices[0] = (int32_t**)malloc(2 * sizeof(int32_t*));
  t->indices[1] = (int32_t**)malloc(2 * sizeof(int32_t*));
  t->indices[0][0] = (int32_t*)malloc(2 * sizeof(int32_t));
  t->indices[0][0][0] = 0;
  t->indices[0][0][1] = nouter;
  t->indices[0][1] = ids;
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
  int64_t na, nb;
  int32_t* aid = in.take<int32_t>(&na);
  int32_t* ap = in.take<int32_t>();
  int32_t* ac = in.take<int32_t>();
  double* av = in.take<double>();
  int32_t* bid = in.take<int32_t>(&nb);
  int32_t* bp = in.take<int32_t>();
  int32_t* bc = in.take<int32_t>();
  double* bv = in.take<double>();
  // A is the DCSC of an m x k matrix (logical dimensions {m, k}, outer level
  // over k), B the DCSR of a k x n matrix
  taco_tensor_t* A = cc_tensor(d[0], d[1], (int32_t)na, aid, ap, ac, av);
  taco_tensor_t* B = cc_tensor(d[1], d[2], (int32_t)nb, bid, bp, bc, bv);
  taco_tensor_t* C = cc_tensor(d[0], d[2], 0, nullptr, nullptr, nullptr, nullptr);
  DCSC_DCSR_coord(A, B, C, 1024);
  if (!C->indices[1][0]) return 1;
  const int32_t nnz = C->indices[1][0][d[0]];
  out.put(C->indices[1][0], (int64_t)d[0] + 1);
  out.put(C->indices[1][1], nnz);
  out.put(C->vals, nnz);
  return 0;
}
