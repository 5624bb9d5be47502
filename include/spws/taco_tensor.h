/*
 * taco_tensor_t -- the level-format container the reference's generated
 * kernels take (utils/lib.h:45-56), with FP64 values.
 *
 * Layouts used by the entry points in spws/spws.h (identical to the
 * reference builders):
 *   CSR  ("DC", DC_to_taco_tensor, utils/dataloader.h:374-425):
 *        indices[0][0] = {nrows}, indices[1][0] = pos[nrows+1],
 *        indices[1][1] = crd[nnz], vals[nnz].
 *   DCSR/DCSC ("CC", CC_to_taco_tensor, utils/dataloader.h:504-569):
 *        indices[0][0] = {0, nR}, indices[0][1] = outer ids[nR],
 *        indices[1][0] = pos[nR+1], indices[1][1] = crd[nnz], vals[nnz].
 *   CSF3 (read_tns_csf, utils/dataloader.h:636-724): per level l,
 *        indices[l][0] = pos, indices[l][1] = crd; the root pos is {0, nroots}.
 *   COO outputs (init_taco_tensor_COO, utils/dataloader.h:964-982):
 *        indices[1..order][0] = coordinate arrays, vals, vals_size = nnz.
 *
 * Deviation from the reference: `vals` is double, not float.
 */
#ifndef SPWS_TACO_TENSOR_H
#define SPWS_TACO_TENSOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum { taco_mode_dense, taco_mode_sparse } taco_mode_t;

typedef struct taco_tensor_t {
  int32_t order;          /* tensor order (number of modes) */
  int32_t* dimensions;    /* tensor dimensions */
  int32_t csize;          /* component size */
  int32_t* mode_ordering; /* mode storage ordering */
  taco_mode_t* mode_types;/* mode storage types */
  int32_t*** indices;     /* tensor index data (per mode) */
  double* vals;           /* tensor values (FP64) */
  int32_t vals_size;      /* values array size */
} taco_tensor_t;

#ifdef __cplusplus
}
#endif

#endif /* SPWS_TACO_TENSOR_H */
