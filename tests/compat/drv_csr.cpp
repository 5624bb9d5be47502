// Reference-style driver over the drop-in headers: the row-wise hash SpGEMM
// (CSR_CSR_hash_sequential.h), the split assemble/compute (CSR_CSR_1.h) and
// the workspace Sort/Merge (utils/wspace.h), unchanged call sites, FP64.
//   drv_csr hash|split <in> <out>:   in  = dims{m,k,n}, A pos/crd/vals, B pos/crd/vals
//                                    out = C pos/crd/vals
//   drv_csr wspace <in> <out>:       in  = {rev}, records, COO crd0/crd1/vals
//                                    out = {count, merged}, sorted records, merged crd0/crd1/vals
#include <string>

#include "frames.hpp"
#include "spws/compat/CSR_CSR_1.h"
#include "spws/compat/CSR_CSR_hash_sequential.h"
#include "spws/compat/wspace.h"

static taco_tensor_t* csr_tensor(int32_t rows, int32_t cols, int32_t* pos, int32_t* crd, double* vals) {
  taco_tensor_t* t = (taco_tensor_t*)calloc(1, sizeof(taco_tensor_t));
  t->order = 2;
  t->dimensions = (int32_t*)malloc(2 * sizeof(int32_t));
  t->dimensions[0] = rows;
  t->dimensions[1] = cols;
  t->indices = (int32_t***)malloc(2 * sizeof(int32_t**));
  t->indices[0] = (int32_t**)malloc(sizeof(int32_t*));
  t->indices[1] = (int32_t**)malloc(2 * sizeof(int32_t*));
  t->indices[0][0] = (int32_t*)malloc(sizeof(int32_t));
  t->indices[0][0][0] = rows;
  t->indices[1][0] = pos;
  t->indices[1][1] = crd;
  t->vals = vals;
  return t;
}

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  const std::string mode = argv[1];
  FrameReader in(argv[2]);
  FrameWriter out(argv[3]);
  if (mode == "hash" || mode == "split") {
    int32_t* d = in.take<int32_t>();
    int32_t* ap = in.take<int32_t>();
    int32_t* ac = in.take<int32_t>();
    double* av = in.take<double>();
    int32_t* bp = in.take<int32_t>();
    int32_t* bc = in.take<int32_t>();
    double* bv = in.take<double>();
    taco_tensor_t* A = csr_tensor(d[0], d[1], ap, ac, av);
    taco_tensor_t* B = csr_tensor(d[1], d[2], bp, bc, bv);
    taco_tensor_t* C = csr_tensor(d[0], d[2], nullptr, nullptr, nullptr);
    if (mode == "hash") {
      CSR_CSR_hash(A, B, C, 1024, 1, 2, true);
    } else {
      if (assemble(C, A, B) || compute(C, A, B)) return 1;
    }
    const int32_t nnz = C->indices[1][0][d[0]];
    out.put(C->indices[1][0], (int64_t)d[0] + 1);
    out.put(C->indices[1][1], nnz);
    out.put(C->vals, nnz);
    return 0;
  }
  if (mode == "wspace") {
    int32_t* flags = in.take<int32_t>();
    int64_t nrec, ncoo;
    wspace* w = in.take<wspace>(&nrec);
    int32_t* c0 = in.take<int32_t>(&ncoo, nrec);
    int32_t* c1 = in.take<int32_t>(nullptr, nrec);
    double* cv = in.take<double>(nullptr, nrec);
    const bool rev = flags[0] != 0;
    const int count = Sort(w, (size_t)nrec, rev);
    if (count < 0) return 1;
    int32_t* crd[2] = {c0, c1};
    const int merged = Merge(crd, cv, (int32_t)ncoo, w, count, rev);
    if (merged < 0) return 1;
    const int32_t sizes[2] = {count, merged};
    out.put(sizes, 2);
    out.put(w, nrec);
    out.put(c0, merged);
    out.put(c1, merged);
    out.put(cv, merged);
    return 0;
  }
  return 2;
}
