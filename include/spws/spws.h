/*
 * libspws -- MI355X (gfx950) sparse-workspace kernels behind a C ABI.
 *
 * Two layers of entry points:
 *   1. TACO-shaped host entry points over taco_tensor_t (host arrays in,
 *      malloc'd host arrays out), one per reference `compute`/`assemble`.
 *      Each comment names the reference function it replaces.
 *   2. Device-resident entry points over plain device pointers and sizes,
 *      used by the batched/served path and by benchmarks.  Outputs of a
 *      device-resident call live in buffers owned by the context and stay
 *      valid until the next device-resident compute call (row-wise, outer,
 *      MTTKRP, TTM) on that context.
 *
 * Every function returns 0 (SPWS_OK) on success and a nonzero spws_status
 * otherwise; spws_last_error() describes the last failure of the calling
 * thread.  Nothing here calls exit().  Contexts are independent: different
 * host threads may use different contexts concurrently.
 */
#ifndef SPWS_SPWS_H
#define SPWS_SPWS_H

#include <stddef.h>
#include <stdint.h>

#include "spws/taco_tensor.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SPWS_OK = 0,
  SPWS_ERR_INVALID = 1,     /* bad argument or malformed input */
  SPWS_ERR_OOM = 2,         /* device or host allocation failed */
  SPWS_ERR_DEVICE = 3,      /* a HIP runtime call failed */
  SPWS_ERR_UNSUPPORTED = 4, /* input outside what the kernels handle (e.g. int32 overflow) */
} spws_status;

typedef struct spws_ctx spws_ctx;

/* Row-wise SpGEMM bin boundaries by per-row upper bound (ub = products).
 * Rows with ub <= wave_ub run one wavefront each with an LDS table of
 * 2*wave_ub slots, ub <= wave2_ub likewise; ub <= block_ub and <= block2_ub
 * run one workgroup each with an LDS table; longer rows use the dense
 * accumulator when ub * dense_ratio >= ncols, else a workgroup with an
 * HBM-resident hash table.  force_path >= 0 sends every non-empty row to
 * one path (0 wave, 1 wave2, 2 block, 3 block2, 4 dense, 5 HBM hash), which
 * tests use to cover each path on small inputs. */
typedef struct {
  int32_t wave_ub;    /* <= 64 */
  int32_t wave2_ub;   /* <= 256 */
  int32_t block_ub;   /* <= 1024 */
  int32_t block2_ub;  /* <= 4096 */
  int32_t dense_ratio;
  int32_t force_path;
} spws_rowwise_tuning;

typedef struct {
  int64_t products;         /* sum of row upper bounds (= multiply-adds) */
  int64_t nnz;              /* output nonzeros */
  int64_t rows_per_path[8]; /* rows sent to each row-wise path (index = force_path id), [6] = empty */
  int64_t chunks;           /* outer product: output-row chunks processed */
  int64_t sort_bits;        /* outer product: key bits radix-sorted per chunk (LSD method) */
  int64_t outer_method;     /* outer product: 1 LSD radix sort, 2 row/window buckets */
  int64_t window_bits;      /* outer product, bucket method: log2 of the column window */
} spws_stats;

/* ---------------------------------------------------------------- context */

const char* spws_last_error(void);
/* stream: a hipStream_t to run on, or NULL for a stream the context creates. */
int spws_ctx_create(int device, void* stream, spws_ctx** out);
int spws_ctx_destroy(spws_ctx* ctx);
int spws_ctx_set_stream(spws_ctx* ctx, void* stream);
void* spws_ctx_get_stream(spws_ctx* ctx);
int spws_ctx_get_stats(spws_ctx* ctx, spws_stats* out);
int spws_ctx_set_rowwise_tuning(spws_ctx* ctx, const spws_rowwise_tuning* t);
int spws_ctx_get_rowwise_tuning(spws_ctx* ctx, spws_rowwise_tuning* t);

/* Per-kernel timing with HIP events recorded around every launch the
 * context makes (off by default).  spws_ctx_set_profiling_only limits the
 * timing to the launches of one name (as spws_ctx_kernel_times reports it;
 * NULL or "" times every launch), so that a timed loop carries the events of
 * one kernel only.  spws_ctx_kernel_times returns the index-th kernel name
 * (alphabetical), its total milliseconds and launch count since the last
 * reset; an index past the end returns SPWS_ERR_INVALID. */
int spws_ctx_set_profiling(spws_ctx* ctx, int on);
int spws_ctx_set_profiling_only(spws_ctx* ctx, const char* name);
int spws_ctx_reset_kernel_times(spws_ctx* ctx);
int spws_ctx_kernel_times(spws_ctx* ctx, int index, char* name, size_t name_len, double* total_ms,
                          int64_t* launches);

/* Device memory helpers for callers without another HIP runtime binding. */
int spws_dev_alloc(spws_ctx* ctx, size_t bytes, void** out);
int spws_dev_free(spws_ctx* ctx, void* p);
int spws_memcpy_h2d(spws_ctx* ctx, void* dst, const void* src, size_t bytes);
int spws_memcpy_d2h(spws_ctx* ctx, void* dst, const void* src, size_t bytes);
int spws_sync(spws_ctx* ctx);
/* The calling thread's default context (device SPWS_DEVICE, default 0), the
 * one the taco entry points use; created on first use, owned by the thread. */
int spws_default_ctx(spws_ctx** out);
void spws_host_free(void* p); /* frees arrays the host entry points malloc'd */
/* Plain device-to-device copy kernel (coalesced 16-byte loads/stores) on the
 * context's stream; also the byte-count reference for hardware counters. */
int spws_device_copy(spws_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ------------------------------------------------- row-wise SpGEMM (CSR x CSR) */

/* C = A * B, A m x k, B k x n, all CSR with int32 pos/crd and FP64 values,
 * device pointers.  Output: sorted CSR, explicit zeros kept, pattern
 * {(i,k) : exists j, A(i,j) and B(j,k) stored}.  On return *C_nnz holds
 * nnz(C) and *C_pos / *C_crd / *C_vals point at context-owned device arrays.
 * Replaces the hot loop of compute() in
 * benchmark/Parallel/CSR_CSR_hash_sequential.h:232-353. */
int spws_spgemm_rowwise(spws_ctx* ctx, int32_t m, int32_t k, int32_t n,
                        const int32_t* A_pos, const int32_t* A_crd, const double* A_vals,
                        const int32_t* B_pos, const int32_t* B_crd, const double* B_vals,
                        int64_t* C_nnz, int32_t** C_pos, int32_t** C_crd, double** C_vals);

/* Symbolic phase only: exact row counts into C_pos (context-owned, m+1). */
int spws_spgemm_rowwise_symbolic(spws_ctx* ctx, int32_t m, int32_t k, int32_t n,
                                 const int32_t* A_pos, const int32_t* A_crd,
                                 const int32_t* B_pos, const int32_t* B_crd,
                                 int64_t* C_nnz, int32_t** C_pos);

/* -------------------------------------- outer-product SpGEMM (DCSC x DCSR) */

/* C = A * B with A (m x k) in DCSC and B (k x n) in DCSR, device pointers.
 * A1_crd[a_ncols] lists A's stored columns (sorted; NULL = all k columns,
 * i.e. CSC), A2_pos[a_ncols+1] / A2_crd (rows, sorted within each column) /
 * A_vals hold them; B1_crd[b_nrows] / B2_pos / B2_crd / B_vals likewise for
 * B's stored rows (NULL = CSR).  A's columns and B's rows are co-iterated on
 * j.  Output: sorted CSR with int64 row pointers (nnz may exceed 2^31),
 * context-owned.  Each entry is summed over j in increasing order.
 * Replaces compute() of benchmark/DCSC_DCSR/DCSC_DCSR_coord_index_chase.h:194. */
int spws_spgemm_outer(spws_ctx* ctx, int32_t m, int32_t k, int32_t n,
                      int32_t a_ncols, const int32_t* A1_crd, const int32_t* A2_pos, const int32_t* A2_crd,
                      const double* A_vals,
                      int32_t b_nrows, const int32_t* B1_crd, const int32_t* B2_pos, const int32_t* B2_crd,
                      const double* B_vals,
                      int64_t* C_nnz, int64_t** C_pos, int32_t** C_crd, double** C_vals);

/* Output-row chunking of the outer product: at most max_products products
 * and max_rows rows per chunk (0 = defaults: 2^30 products; rows limited by
 * the 32-bit sort key).  Small values force many chunks (used by tests). */
int spws_ctx_set_outer_chunking(spws_ctx* ctx, int64_t max_products, int32_t max_rows);

/* Sort method of the outer product's coordinate workspace:
 *   1  LSD radix sort of packed (i - r0) << kb | k keys over the whole chunk,
 *      then a segmented reduce of the runs;
 *   2  MSD by buckets: the expand scatters each product straight into its
 *      (row, column window) bucket at a position computed from the operands'
 *      structure (exact digit counts, no histogram pass over the products),
 *      then each bucket is ordered by k within the window and reduced in LDS;
 *   3  fused row windows: no product workspace; one wavefront per output row
 *      walks its column windows, gathering each window's products straight
 *      from B (symbolic pass for the exact row pointers, then the numeric
 *      pass writing every entry in place);
 *   4  as 3, the products added into the LDS window by atomic adds (no
 *      ordering of equal columns): sums reassociated within the 1e-12 bar,
 *      not reproducible bit for bit;
 *   0  automatic (4 unless the windows would be too sparse or too many,
 *      then 1).
 * window_log2 sets the column window of methods 2-4 (6..11; 0 = chosen
 * from the average products per row).  Methods 1-3 sum each entry in
 * increasing j and give identical results. */
int spws_ctx_set_outer_method(spws_ctx* ctx, int32_t method, int32_t window_log2);

/* CSR transpose on the device, the device form of transpose() in
 * benchmark/CSC_CSR_T/nonfuse/CSC_CSR_T_hash.h:352: the m x n matrix
 * (pos[m+1], crd[nnz], vals[nnz]) becomes its n x m transpose in
 * out_pos[n+1], out_crd[nnz], out_vals[nnz] (device arrays the caller
 * provides), each output row's columns in increasing order.  Asynchronous
 * on the context's stream. */
int spws_csr_transpose_dev(spws_ctx* ctx, int32_t m, int32_t n, int64_t nnz, const int32_t* pos,
                           const int32_t* crd, const double* vals, int32_t* out_pos, int32_t* out_crd,
                           double* out_vals);

/* spws_csr_transpose_dev for a CSR with int64 row pointers, the form
 * spws_spgemm_outer returns its product in (nnz still below 2^31): the
 * second half of the non-fused transposed product of
 * benchmark/CSC_CSR_T/nonfuse/CSC_CSR_T_hash.h (compute_nofuse at :232,
 * then transpose() at :352) with the product left on the device.  Same
 * output arrays and order as spws_csr_transpose_dev. */
int spws_csr64_transpose_dev(spws_ctx* ctx, int32_t m, int32_t n, int64_t nnz, const int64_t* pos,
                             const int32_t* crd, const double* vals, int32_t* out_pos, int32_t* out_crd,
                             double* out_vals);

/* spws_csr_transpose_dev without a context and without allocating anything
 * (no hipMalloc/hipFree, no host allocation): the same kernels and the same
 * result, on `stream` (a hipStream_t; NULL = the null stream) of the current
 * device, with every workspace taken from a fixed 48 MiB device array that
 * each GPU holds for the life of the process.  About 33 bytes per entry plus
 * 4 per output row must fit it (e.g. nnz 1,000,000 with n up to 3,000,000);
 * larger inputs return SPWS_ERR_UNSUPPORTED before any launch.  Calls that
 * overlap on one GPU share the array and race: serialise them (one stream,
 * or events).  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). */
int spws_csr_transpose_dev_noalloc(void* stream, int32_t m, int32_t n, int64_t nnz, const int32_t* pos,
                                   const int32_t* crd, const double* vals, int32_t* out_pos, int32_t* out_crd,
                                   double* out_vals);

/* ------------------------------------------------------ sparse tensor kernels */

/* A 3-level CSF is passed as nroots root coordinates l1_crd (the compressed
 * root level), l2_pos[nroots+1] / l2_crd[nfibers] (second level),
 * l3_pos[nfibers+1] / l3_crd[nnz] (leaves) and vals[nnz]; coordinates are
 * sorted within each parent.  Device pointers. */

/* MTTKRP, M(i,r) = sum_{j,k} X(i,j,k) Bf(j,r) Cf(k,r), for a CSF with levels
 * (j, k, i) -- the output mode is the leaf, as in the reference's B(k,l,i).
 * Bf (J x R) and Cf (K x R) are dense row-major, 1 <= R <= 64.  Output
 * (context-owned): M dense I x R row-major, and the sorted list of the
 * *nrows rows that hold a nonzero (the rows of the reference's COO output).
 * Replaces compute() of benchmark/MTTKRP/COO_CSF_DCSR_DCSR_hash_free.h:233. */
int spws_mttkrp(spws_ctx* ctx, int32_t I, int32_t J, int32_t K, int32_t R,
                int32_t nroots, const int32_t* l1_crd, const int32_t* l2_pos, const int32_t* l2_crd,
                int32_t nfibers, const int32_t* l3_pos, const int32_t* l3_crd, const double* vals, int64_t nnz,
                const double* Bf, const double* Cf, int32_t* nrows, int32_t** rows, double** M);

/* How spws_ttm groups the nonzeros of each output fiber (no counterpart in
 * the reference, whose hash_il workspace does it on the host): 0 automatic
 * (roots and column windows when the shapes allow it and the heavy roots'
 * column windows hold at least 32 nonzeros on average, else 1), 1 one stable
 * radix sort of the whole tensor on the fiber key, 2 roots and column
 * windows (SPWS_ERR_UNSUPPORTED outside its shapes: R <= 64, at most 4,096
 * windows of 1024 / R' columns with R' = R rounded up to 16, 32 or 64).
 * Both sum every fiber over k in increasing order and give identical
 * results; a U row mask always takes method 1. */
int spws_ctx_set_ttm_method(spws_ctx* ctx, int32_t method);

/* TTM, Y(i,j,r) = sum_k X(i,k,j) U(k,r), for a CSF with levels (i, k, j) --
 * the contracted mode is the middle level, as in the reference's hash_il
 * B(k,i,j).  U (K x R) dense row-major, 1 <= R <= 256.  Output
 * (context-owned): *nfib fibers (fib_i, fib_j) sorted by (i, j), each with a
 * dense length-R vector in Y (nfib x R row-major).
 * Replaces compute() of benchmark/TTM/COO_CSF_DCSR_hash_il.h:247. */
int spws_ttm(spws_ctx* ctx, int32_t I, int32_t J, int32_t K, int32_t R,
             int32_t nroots, const int32_t* l1_crd, const int32_t* l2_pos, const int32_t* l2_crd,
             int32_t nfibers, const int32_t* l3_pos, const int32_t* l3_crd, const double* vals, int64_t nnz,
             const double* U, int64_t* nfib, int32_t** fib_i, int32_t** fib_j, double** Y);

/* Hypersparse (DCSR) factors: the reference co-iterates its factors' stored
 * rows and columns with the tensor's levels (hash_free.h:277-300: k over
 * B1 and C1, l over B2 and D1, the rank over C2 and D2; hash_il.h:294-303:
 * the contracted index over B2 and C1, the rank over C2).  Here a factor is
 * passed dense (zero where not stored) with a mask per row: bit r set when
 * column r of that row is stored (a missing row has mask 0), R <= 64.  A
 * nonzero whose rows share no stored column takes no part; an output row
 * (MTTKRP) or fiber (TTM) exists when one of its nonzeros takes part, and
 * *row_mask / *fib_mask (context-owned, one uint64 per listed row or fiber)
 * give its stored output columns: the OR over its nonzeros of the shared
 * columns.  Values of columns not stored are 0.  Bmask/Cmask: both or
 * neither (NULL: dense, as spws_mttkrp); Umask NULL: as spws_ttm.  The TTM
 * mask needs the root, leaf and contracted coordinates to fit 64 bits
 * together.  Same references as spws_mttkrp / spws_ttm. */
int spws_mttkrp_masked(spws_ctx* ctx, int32_t I, int32_t J, int32_t K, int32_t R,
                       int32_t nroots, const int32_t* l1_crd, const int32_t* l2_pos, const int32_t* l2_crd,
                       int32_t nfibers, const int32_t* l3_pos, const int32_t* l3_crd, const double* vals,
                       int64_t nnz, const double* Bf, const double* Cf, const uint64_t* Bmask,
                       const uint64_t* Cmask, int32_t* nrows, int32_t** rows, uint64_t** row_mask, double** M);
int spws_ttm_masked(spws_ctx* ctx, int32_t I, int32_t J, int32_t K, int32_t R,
                    int32_t nroots, const int32_t* l1_crd, const int32_t* l2_pos, const int32_t* l2_crd,
                    int32_t nfibers, const int32_t* l3_pos, const int32_t* l3_crd, const double* vals, int64_t nnz,
                    const double* U, const uint64_t* Umask, int64_t* nfib, int32_t** fib_i, int32_t** fib_j,
                    uint64_t** fib_mask, double** Y);

/* ------------------------------------------------ workspace API (wspace.h) */

/* A workspace record is {int32_t crd[2]; double val;} (16 bytes), the
 * reference's wspace struct (utils/wspace.h:10-13) with an FP64 value; an
 * empty slot has crd[0] = -1 (init_wspace, :15-23).  Host arrays. */

/* Sort(array, size, rev) of utils/wspace.h:206-217: sorts n records by
 * (crd[0], crd[1]) in signed lexicographic order, ascending (rev = 0) or
 * descending (rev != 0).  *nvalid receives n for ascending order and, for
 * descending order, the number of records before the trailing empty slots. */
int spws_wspace_sort(spws_ctx* ctx, void* records, int64_t n, int rev, int64_t* nvalid);

/* Merge(COO_crd, COO_vals, COO_size, accumulator, accumulator_size) of
 * utils/wspace.h:62-204: merges n_acc workspace records (any order) into the
 * sorted COO (coo_crd0, coo_crd1, coo_vals; coo_size entries; capacity at
 * least coo_size + n_acc), adding the values of equal coordinates, and writes
 * the merged, sorted COO back; *out_size receives its length.  Equal
 * coordinates inside the workspace are added as well (the reference assumes
 * none). */
int spws_wspace_merge(spws_ctx* ctx, int32_t* coo_crd0, int32_t* coo_crd1, double* coo_vals, int64_t coo_size,
                      const void* records, int64_t n_acc, int64_t* out_size);

/* -------------------------------------------- batched and served SpGEMM */

/* A batch engine runs many row-wise SpGEMMs C_b = A_b * B_b (device operands,
 * int32 CSR, FP64) in one hipGraph launch: the requests are stacked into one
 * block-diagonal product, whose rows are binned together, so the per-entry
 * arithmetic is that of spws_spgemm_rowwise.  The engine's pools are sized
 * once by its capacities; a request is admitted while every pool holds it
 * (products = its upper bound of output entries = sum over A's entries of
 * the matching B row lengths, max_row_products = the largest row's).  Each
 * engine owns a stream; engines run concurrently.  No counterpart in the
 * reference, which has no batching. */
typedef struct spws_batch spws_batch;

typedef struct {
  int32_t max_requests;      /* requests per batch (<= 65535) */
  int32_t max_rows;          /* sum of A row counts */
  int32_t max_brows;         /* sum of B row counts */
  int32_t max_cols;          /* widest request (n) the engine takes, <= 2^17; 0 = any
                                width (the block bins then hash and sort instead of
                                using the LDS column bitmap) */
  int64_t max_annz;          /* sum of nnz(A_b) */
  int64_t max_bnnz;          /* sum of nnz(B_b) */
  int64_t max_products;      /* sum of upper bounds of output entries (<= 2^31 - 1) */
  int64_t max_row_products;  /* largest row upper bound of any request (<= 2^28) */
} spws_batch_caps;

int spws_batch_create(spws_ctx* ctx, const spws_batch_caps* caps, spws_batch** out);
int spws_batch_destroy(spws_batch* b);
/* Empty the batch (only while no launch is in flight). */
int spws_batch_reset(spws_batch* b);
/* Add one request; SPWS_ERR_OOM when the batch cannot hold it (full),
 * SPWS_ERR_UNSUPPORTED when a row exceeds max_row_products or the request
 * is wider than max_cols.  *slot names the
 * request in spws_batch_result.  The operand arrays must stay valid until
 * the batch completes, and must be consistent: A_pos[m] = annz,
 * B_pos[k] = bnnz, every A column < k and every B column < n.
 * products must be at least the true sum over A's entries of the matching B
 * row lengths, and max_row_products at least the largest row's sum: admission
 * sizes the pools by them.  The batch recomputes both on the device before
 * writing; when either was under-reported nothing is computed and
 * spws_batch_nnz / spws_batch_result / spws_batch_result_nnz return
 * SPWS_ERR_INVALID. */
int spws_batch_add(spws_batch* b, int32_t m, int32_t k, int32_t n, const int32_t* A_pos, const int32_t* A_crd,
                   const double* A_vals, int64_t annz, const int32_t* B_pos, const int32_t* B_crd,
                   const double* B_vals, int64_t bnnz, int64_t products, int64_t max_row_products, int32_t* slot);
/* Give request `slot` of the batch being filled an output buffer (device):
 * C_pos[m + 1], C_crd and C_vals of `capacity` entries.  The batch writes the
 * request's result (row pointers from 0) there instead of into its pools;
 * spws_batch_result_nnz then gives its nnz, or -1 when it exceeds capacity
 * (nothing copied).  C_pos = NULL removes the buffer. */
int spws_batch_set_output(spws_batch* b, int32_t slot, int32_t* C_pos, int32_t* C_crd, double* C_vals,
                          int64_t capacity);
/* Launch the batch's graph on the engine's stream (asynchronous). */
int spws_batch_launch(spws_batch* b);
/* *done = 1 once the launched batch has completed. */
int spws_batch_poll(spws_batch* b, int* done);
int spws_batch_wait(spws_batch* b);
/* Total output nonzeros of the completed batch. */
int spws_batch_nnz(spws_batch* b, int64_t* nnz);
/* Result of one request of a completed batch: its row pointers rebased to 0
 * (host array of m+1, may be NULL), nnz, and device pointers to its columns
 * and values: in the request's output buffer when it has one that holds the
 * result, else engine-owned (valid until the engine's next launch). */
int spws_batch_result(spws_batch* b, int32_t slot, int64_t* nnz, int32_t* pos_host, int32_t** crd, double** vals);
/* Output nnz of request `slot` of the completed batch (-1: it exceeded the
 * capacity given to spws_batch_set_output). */
int spws_batch_result_nnz(spws_batch* b, int32_t slot, int64_t* nnz);
void* spws_batch_stream(spws_batch* b);

/* One request of the serving loop: device operands, the upper bounds used by
 * admission, the arrival time in seconds from the start of the run, and an
 * optional device output buffer (C_pos[m + 1]; C_crd, C_vals of C_cap
 * entries; C_pos = NULL: the result is not kept). */
typedef struct {
  int32_t m, k, n, reserved;
  const int32_t* A_pos;
  const int32_t* A_crd;
  const double* A_vals;
  int64_t annz;
  const int32_t* B_pos;
  const int32_t* B_crd;
  const double* B_vals;
  int64_t bnnz;
  int64_t products;
  int64_t max_row_products;
  double arrival;
  int32_t* C_pos;
  int32_t* C_crd;
  double* C_vals;
  int64_t C_cap;
} spws_request;

typedef struct {
  int32_t mode;                /* 0: plain (one row-wise call per request); 1: continuously batched */
  int32_t max_batch_requests;  /* batched: requests per batch (0 = engine capacity) */
  int64_t max_batch_products;  /* batched: products (output upper bound) per batch, so a
                                  batch stays short and latency bounded; a single request
                                  larger than this still runs alone (0 = engine capacity) */
  int64_t large_products;      /* batched: requests with at least this many products form the
                                  large lane (0 = one lane for all) */
  int32_t large_engines;       /* batched: engines reserved for the large lane, one request
                                  each (< nengines) */
  int32_t reserved;
} spws_serve_options;

typedef struct {
  double wall_seconds;
  int64_t batches;
  int64_t fallback_requests;   /* batched mode: requests run through the plain path */
} spws_serve_stats;

/* Run the request loop in real time: requests (sorted by arrival) arrive at
 * their arrival times; finish[r] receives the time (s) the loop saw request
 * r complete, its result already in the request's output buffer; batch_of[r]
 * (optional) the batch that carried it (-1 = plain); result_nnz[r]
 * (optional) its output nnz (-1: it exceeded C_cap, nothing copied).
 * Batched mode uses the nengines engines.  No counterpart in the reference. */
int spws_serve(spws_ctx* ctx, spws_batch** engines, int32_t nengines, int32_t nreq, const spws_request* reqs,
               const spws_serve_options* opt, double* finish, int32_t* batch_of, int64_t* result_nnz,
               spws_serve_stats* stats);

/* Host TACO entry points (host taco_tensor_t in, malloc'd host arrays out). */

/* int compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap)
 * of benchmark/Parallel/CSR_CSR_hash_sequential.h:232 (and the identical
 * benchmark/CSR_CSR/CSR_CSR_hash.h).  Sets C->indices[1][0] (pos),
 * C->indices[1][1] (crd), C->vals and C->vals_size; the caller free()s them.
 * w_cap is accepted for signature compatibility; the device workspace is
 * sized per row, so it does not change the result structure. */
int spws_csr_csr_hash_compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap);

/* int assemble(C, A, B) of seperate/CSR_CSR_1.h:69-167: fills C2_pos (the
 * symbolic phase) and allocates C2_crd / C_vals for compute() to fill. */
int spws_csr_csr_assemble(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B);

/* int compute(C, A, B) of seperate/CSR_CSR_1.h:4-67: fills C_vals (and
 * C2_crd) into the arrays assemble() allocated. */
int spws_csr_csr_compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B);

/* int compute(C, A, B, w_cap) of
 * benchmark/DCSC_DCSR/DCSC_DCSR_coord_index_chase.h:194: A DCSC, B DCSR
 * ("CC" taco tensors), C CSR with int32 pos (SPWS_ERR_UNSUPPORTED when
 * nnz(C) >= 2^31).  Sets C->indices[1][0], C->indices[1][1], C->vals. */
int spws_dcsc_dcsr_coord_compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap);

/* int compute(C, A, B, w_cap) of benchmark/CSC_CSR_T/CSC_CSR_T_hash.h:232:
 * the transposed-output product C = (A B)^T with A CSC (m x k, stored by
 * column), B CSR (k x n) and C CSR (n x m) -- the workspace keyed on (k, i),
 * fused with the product.  Computed as the outer product B^T A^T (the
 * operands' arrays swapped), each entry summed over j in increasing order.
 * Sets C->indices[1][0], C->indices[1][1], C->vals (SPWS_ERR_UNSUPPORTED
 * when nnz(C) >= 2^31). */
int spws_csc_csr_t_hash_compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap);

/* int compute_nofuse(C, A, B, w_cap) of
 * benchmark/CSC_CSR_T/nonfuse/CSC_CSR_T_hash.h:232: the same product
 * without the transposition, C = A B with A CSC (m x k), B CSR (k x n) and
 * C CSR (m x n), each entry summed over j in increasing order.  Sets
 * C->indices[1][0], C->indices[1][1], C->vals (SPWS_ERR_UNSUPPORTED when
 * nnz(C) >= 2^31). */
int spws_csc_csr_nofuse_compute(taco_tensor_t* C, taco_tensor_t* A, taco_tensor_t* B, int32_t w_cap);

/* int transpose(A, B) of benchmark/CSC_CSR_T/nonfuse/CSC_CSR_T_hash.h:352:
 * A = B^T for a CSR B with B->dimensions[0] rows.  A, initialised as a
 * 2-level tensor whose A->dimensions[0] covers B's columns, receives
 * malloc'd pos/crd/vals, each row's columns in increasing order. */
int spws_csr_transpose(taco_tensor_t* A, taco_tensor_t* B);

/* int compute(A, B, C, D, w_cap) of
 * benchmark/MTTKRP/COO_CSF_DCSR_DCSR_hash_free.h:233: B a 3-level CSF with
 * the output mode as leaf, C and D dense factors stored as DCSR.  Sets the
 * COO output A->indices[1][0] (row), A->indices[2][0] (rank), A->vals and
 * A->vals_size, sorted by (row, rank). */
int spws_mttkrp_hash_compute(taco_tensor_t* A, taco_tensor_t* B, taco_tensor_t* C, taco_tensor_t* D,
                             int32_t w_cap);

/* int compute(A, B, C, w_cap) of benchmark/TTM/COO_CSF_DCSR_hash_il.h:247: B
 * a 3-level CSF with the contracted mode in the middle, C a dense factor
 * stored as DCSR.  Sets the COO output A->indices[1][0], [2][0], [3][0],
 * A->vals and A->vals_size, sorted by (root, leaf, column). */
int spws_ttm_hash_il_compute(taco_tensor_t* A, taco_tensor_t* B, taco_tensor_t* C, int32_t w_cap);

#ifdef __cplusplus
}
#endif

#endif /* SPWS_SPWS_H */
