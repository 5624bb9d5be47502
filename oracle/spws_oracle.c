/*
 * spws_oracle.c -- CPU restatement of the SparseWS sparse-workspace algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
 * oracle/libspws_oracle.so and loaded by tests/, by __graft_entry__.smoke()
 * and by the cpu_baseline leg of bench.py.  The product library
 * (sparsews_amd/libspws.so) never links, loads or calls it.
 *
 * Values are FP64 (the reference stores float, utils/lib.h:54; the FP64
 * build is the documented deviation).  Every product a*b is rounded before
 * it is added (no FMA contraction: build with -ffp-contract=off), which is
 * what the reference's x86-64 build without -march does.
 *
 * Functions and the reference code each one restates:
 *   orc_spgemm_hash        benchmark/Parallel/CSR_CSR_hash_sequential.h:232-353
 *                          (TryInsert_hash :151-190, flush/Merge_hash :55-123,
 *                           refresh_wspace growth :202-230)
 *   orc_spgemm_hash_parallel benchmark/Parallel/CSR_CSR_hash_parallel.h:235-408 (CAP
 *                          workers + a merger; the racy flush handshake serialised)
 *   orc_spgemm_gustavson   benchmark/CSR_CSR/CSR_CSR_taco_parallel.h:4-113
 *   orc_spgemm_row_counts  seperate/CSR_CSR_1.h:69-108 (C2_nnz per row, int64 total)
 *   orc_spgemm_symbolic    seperate/CSR_CSR_1.h:69-167 (assemble: C2_nnz -> C2_pos)
 *   orc_spgemm_row_ub      evaluate/CSR_CSR_1.h:153-171 (C2_nnz_upper_bound)
 *   orc_spgemm_refsum      FP64 sorted-summation oracle (no reference counterpart):
 *                          long-double sum of the products in j order, plus sum|terms|
 *   orc_outer_coord        benchmark/DCSC_DCSR/DCSC_DCSR_coord_index_chase.h:194-311
 *                          (TryInsert_coord_t :66-79, Sort_t :36-47, Merge_coord_t :81-192)
 *   orc_mttkrp_hash        benchmark/MTTKRP/COO_CSF_DCSR_DCSR_hash_free.h:233-366
 *   orc_ttm_hash_il        benchmark/TTM/COO_CSF_DCSR_hash_il.h:247-377
 *   orc_wspace_sort        utils/wspace.h:206-217 (Sort; esc_cmp :24-31, esc_cmp_rev :32-39)
 *   orc_wspace_merge       utils/wspace.h:62-204 (Merge, both directions)
 *   orc_csr_transpose      benchmark/CSC_CSR_T/nonfuse/CSC_CSR_T_hash.h:352-400 (transpose)
 */
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* Bounded hash workspace with the reference's flush schedule.                */
/* ------------------------------------------------------------------------- */
/*
 * The reference keeps a chained hash table of `table_size = w_cap` buckets
 * and a flat buffer of capacity `buffer_capacity` (2*w_cap initially,
 * CSR_CSR_hash_sequential.h:192-200).  An insert first checks
 * `numel == buffer_capacity` (:152); if the buffer is full it gathers the
 * chains, sorts them (:153-155), the caller merges them into the sorted
 * "all-array" with add-on-equal (:294, Merge_hash :55-123), clears the table,
 * grows the capacity (x2 below 4194304, else x1.5 truncated, :221-225) and
 * re-inserts the product (:297).  Otherwise the product is added to the
 * existing entry (:170-174) or appended (:185-187).
 *
 * The epoch boundaries therefore depend only on the number of distinct keys
 * per epoch, never on the hash function, so this restatement uses open
 * addressing over a packed 64-bit key; the sums it produces are the
 * reference's: left-to-right within an epoch, then all_old + epoch_partial.
 */
typedef struct {
  uint64_t key;
  double val;
} orc_kv;

typedef struct {
  uint64_t* tkeys;   /* open-addressing keys, UINT64_MAX = empty */
  int64_t* tslot;    /* buffer index for the key */
  uint64_t tmask;
  orc_kv* buf;       /* insertion-ordered entries of this epoch */
  int64_t bufcap;    /* allocated entries of buf (storage only) */
  int64_t numel;
  int64_t cap;       /* buffer_capacity: the flush threshold */
  int grow_policy;   /* 0: row-wise/MTTKRP (x2 < 4194304, else x1.5)
                        1: TTM hash_il (never flushes: cap = INT32_MAX) */
  /* sorted all-array */
  uint64_t* akey;
  double* aval;
  int64_t asize;
} orc_hws;

static uint64_t orc_mix(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

/* Storage grows independently of the flush threshold; only `cap` decides
 * when an epoch ends. */
static void hws_table_reset(orc_hws* w, uint64_t ts) {
  free(w->tkeys);
  free(w->tslot);
  w->tkeys = (uint64_t*)malloc(ts * sizeof(uint64_t));
  memset(w->tkeys, 0xff, ts * sizeof(uint64_t));
  w->tslot = (int64_t*)malloc(ts * sizeof(int64_t));
  w->tmask = ts - 1;
}

static void hws_rehash(orc_hws* w) {
  uint64_t ts = (w->tmask + 1) * 2;
  hws_table_reset(w, ts);
  for (int64_t s = 0; s < w->numel; s++) {
    uint64_t h = orc_mix(w->buf[s].key) & w->tmask;
    while (w->tkeys[h] != UINT64_MAX) h = (h + 1) & w->tmask;
    w->tkeys[h] = w->buf[s].key;
    w->tslot[h] = s;
  }
}

static void hws_init(orc_hws* w, int32_t w_cap, int grow_policy) {
  memset(w, 0, sizeof(*w));
  w->grow_policy = grow_policy;
  w->cap = grow_policy == 1 ? (int64_t)INT32_MAX : 2 * (int64_t)w_cap;
  hws_table_reset(w, 1024);
  w->bufcap = 512;
  w->buf = (orc_kv*)malloc((size_t)w->bufcap * sizeof(orc_kv));
  w->akey = (uint64_t*)malloc(sizeof(uint64_t));
  w->aval = (double*)malloc(sizeof(double));
}

static void hws_free(orc_hws* w) {
  free(w->tkeys);
  free(w->tslot);
  free(w->buf);
  free(w->akey);
  free(w->aval);
}

static int kv_cmp(const void* a, const void* b) {
  uint64_t x = ((const orc_kv*)a)->key, y = ((const orc_kv*)b)->key;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* Sort the epoch and merge it into the all-array (Merge_hash :55-123). */
static void hws_flush(orc_hws* w) {
  if (w->numel == 0) return;
  qsort(w->buf, (size_t)w->numel, sizeof(orc_kv), kv_cmp);
  int64_t need = w->asize + w->numel;
  uint64_t* nk = (uint64_t*)malloc((size_t)need * sizeof(uint64_t));
  double* nv = (double*)malloc((size_t)need * sizeof(double));
  int64_t a = 0, c = 0, t = 0;
  while (a < w->numel && c < w->asize) {
    uint64_t ka = w->buf[a].key, kc = w->akey[c];
    if (ka == kc) {
      nk[t] = ka;
      nv[t] = w->buf[a].val + w->aval[c];
      a++; c++; t++;
    } else if (ka < kc) {
      nk[t] = ka; nv[t] = w->buf[a].val; a++; t++;
    } else {
      nk[t] = kc; nv[t] = w->aval[c]; c++; t++;
    }
  }
  while (a < w->numel) { nk[t] = w->buf[a].key; nv[t] = w->buf[a].val; a++; t++; }
  while (c < w->asize) { nk[t] = w->akey[c]; nv[t] = w->aval[c]; c++; t++; }
  free(w->akey);
  free(w->aval);
  w->akey = nk;
  w->aval = nv;
  w->asize = t;
}

/* refresh_wspace (:202-230): clear the table and grow the capacity. */
static void hws_refresh(orc_hws* w) {
  memset(w->tkeys, 0xff, (w->tmask + 1) * sizeof(uint64_t));
  w->numel = 0;
  if (w->grow_policy == 0) {
    if (w->cap < 4194304) w->cap *= 2;
    else w->cap = (int64_t)(int32_t)((double)w->cap * 1.5);
  }
}

static void hws_insert_nocheck(orc_hws* w, uint64_t key, double v) {
  uint64_t h = orc_mix(key) & w->tmask;
  for (;;) {
    uint64_t k = w->tkeys[h];
    if (k == key) {
      w->buf[w->tslot[h]].val += v;
      return;
    }
    if (k == UINT64_MAX) {
      if (w->numel == w->bufcap) {
        w->bufcap *= 2;
        w->buf = (orc_kv*)realloc(w->buf, (size_t)w->bufcap * sizeof(orc_kv));
      }
      w->tkeys[h] = key;
      w->tslot[h] = w->numel;
      w->buf[w->numel].key = key;
      w->buf[w->numel].val = v;
      w->numel++;
      if ((uint64_t)w->numel * 2 > w->tmask + 1) hws_rehash(w);
      return;
    }
    h = (h + 1) & w->tmask;
  }
}

/* TryInsert_hash + the caller's insertFail protocol (:280-298). */
static void hws_insert(orc_hws* w, uint64_t key, double v) {
  if (w->numel == w->cap) {
    hws_flush(w);
    hws_refresh(w);
  }
  hws_insert_nocheck(w, key, v);
}

/* ------------------------------------------------------------------------- */
/* Row-wise SpGEMM                                                            */
/* ------------------------------------------------------------------------- */

/* C = A*B with the bounded hash workspace; returns nnz(C).
 * Cp[m+1] is caller-provided; Ci and Cx are malloc'd (free with orc_free). */
ORC_API int64_t orc_spgemm_hash(int32_t m, int32_t n,
                                const int32_t* Ap, const int32_t* Ai, const double* Ax,
                                const int32_t* Bp, const int32_t* Bi, const double* Bx,
                                int32_t w_cap, int32_t* Cp, int32_t** Ci, double** Cx) {
  orc_hws w;
  hws_init(&w, w_cap, 0);
  (void)n;
  for (int32_t i = 0; i < m; i++) {
    for (int32_t p = Ap[i]; p < Ap[i + 1]; p++) {
      int32_t j = Ai[p];
      double a = Ax[p];
      for (int32_t q = Bp[j]; q < Bp[j + 1]; q++) {
        uint64_t key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)Bi[q];
        double prod = a * Bx[q];
        hws_insert(&w, key, prod);
      }
    }
  }
  hws_flush(&w);
  int64_t nnz = w.asize;
  int32_t* ci = (int32_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
  double* cx = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
  memset(Cp, 0, (size_t)(m + 1) * sizeof(int32_t));
  for (int64_t t = 0; t < nnz; t++) {
    int32_t r = (int32_t)(w.akey[t] >> 32);
    Cp[r + 1]++;
    ci[t] = (int32_t)(uint32_t)(w.akey[t] & 0xffffffffu);
    cx[t] = w.aval[t];
  }
  for (int32_t r = 0; r < m; r++) Cp[r + 1] += Cp[r];
  hws_free(&w);
  *Ci = ci;
  *Cx = cx;
  return nnz;
}

/* CAP-parallel hash SpGEMM (benchmark/Parallel/CSR_CSR_hash_parallel.h:
 * 235-408): `cap` worker threads, worker t inserting the rows i = t, t + cap,
 * ... into its own bounded hash workspace (same flush schedule as the
 * sequential path, :280-298), and one merger that merges each worker's
 * sorted epochs into the global sorted array with add-on-equal (Merge_hash
 * :55-123).  The reference hands a full buffer to the merger through a spin
 * on a plain int (racy: it can hang or merge a buffer being refilled); here
 * a worker's mid-run epochs go into its own sorted array, and the merger
 * merges the workers' arrays in worker order after they finish.  Workers own
 * disjoint rows, so equal keys only ever meet within one worker, and every
 * value is that worker's epoch partials added in epoch order, as in the
 * reference.  *flushes receives the number of mid-run epochs (0 whenever
 * every worker's entries fit 2 w_cap, e.g. the 10,000 x 10,000 8-per-row
 * workload with cap >= 5 and w_cap 2^16).  Returns nnz(C). */
ORC_API int64_t orc_spgemm_hash_parallel(int32_t m, int32_t n, const int32_t* Ap, const int32_t* Ai,
                                         const double* Ax, const int32_t* Bp, const int32_t* Bi,
                                         const double* Bx, int32_t w_cap, int32_t cap, int32_t* Cp, int32_t** Ci,
                                         double** Cx, int64_t* flushes) {
  (void)n;
  if (cap < 1) cap = 1;
  orc_hws* ws = (orc_hws*)calloc((size_t)cap, sizeof(orc_hws));
  int64_t* nfl = (int64_t*)calloc((size_t)cap, sizeof(int64_t));
#pragma omp parallel num_threads(cap)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
#else
    const int t = 0, nt = 1;
#endif
    /* a worker per thread; with fewer threads than cap, a thread runs several */
    for (int wkr = t; wkr < cap; wkr += nt) {
      orc_hws* w = &ws[wkr];
      hws_init(w, w_cap, 0);
      for (int32_t i = wkr; i < m; i += cap) {
        for (int32_t p = Ap[i]; p < Ap[i + 1]; p++) {
          int32_t j = Ai[p];
          double a = Ax[p];
          for (int32_t q = Bp[j]; q < Bp[j + 1]; q++) {
            uint64_t key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)Bi[q];
            if (w->numel == w->cap) nfl[wkr]++;
            hws_insert(w, key, a * Bx[q]);
          }
        }
      }
      hws_flush(w);
    }
  }
  /* the merger: each worker's sorted array into the global one, in order */
  uint64_t* gk = (uint64_t*)malloc(sizeof(uint64_t));
  double* gv = (double*)malloc(sizeof(double));
  int64_t gs = 0, tot_fl = 0;
  for (int wkr = 0; wkr < cap; wkr++) {
    orc_hws* w = &ws[wkr];
    tot_fl += nfl[wkr];
    int64_t need = gs + w->asize;
    uint64_t* nk = (uint64_t*)malloc((size_t)(need > 0 ? need : 1) * sizeof(uint64_t));
    double* nv = (double*)malloc((size_t)(need > 0 ? need : 1) * sizeof(double));
    int64_t a = 0, c = 0, o = 0;
    while (a < w->asize && c < gs) {
      if (w->akey[a] == gk[c]) {
        nk[o] = gk[c];
        nv[o] = w->aval[a] + gv[c];
        a++; c++; o++;
      } else if (w->akey[a] < gk[c]) {
        nk[o] = w->akey[a]; nv[o] = w->aval[a]; a++; o++;
      } else {
        nk[o] = gk[c]; nv[o] = gv[c]; c++; o++;
      }
    }
    while (a < w->asize) { nk[o] = w->akey[a]; nv[o] = w->aval[a]; a++; o++; }
    while (c < gs) { nk[o] = gk[c]; nv[o] = gv[c]; c++; o++; }
    free(gk);
    free(gv);
    gk = nk;
    gv = nv;
    gs = o;
    hws_free(w);
  }
  free(ws);
  free(nfl);
  int32_t* ci = (int32_t*)malloc((size_t)(gs > 0 ? gs : 1) * sizeof(int32_t));
  double* cx = (double*)malloc((size_t)(gs > 0 ? gs : 1) * sizeof(double));
  memset(Cp, 0, (size_t)(m + 1) * sizeof(int32_t));
  for (int64_t t = 0; t < gs; t++) {
    Cp[(int32_t)(gk[t] >> 32) + 1]++;
    ci[t] = (int32_t)(uint32_t)(gk[t] & 0xffffffffu);
    cx[t] = gv[t];
  }
  for (int32_t r = 0; r < m; r++) Cp[r + 1] += Cp[r];
  free(gk);
  free(gv);
  *Ci = ci;
  *Cx = cx;
  if (flushes) *flushes = tot_fl;
  return gs;
}

static int int_cmp(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* The per-row counts of the symbolic phase, C2_nnz[i] = |U_{j in A_i}
 * cols(B_j)| (seperate/CSR_CSR_1.h:69-108: a dense flag array and an index
 * list per row), into cnt[m]; returns their int64 total, so products with
 * more than 2^31 - 1 entries are counted too. */
ORC_API int64_t orc_spgemm_row_counts(int32_t m, int32_t n, const int32_t* Ap, const int32_t* Ai,
                                      const int32_t* Bp, const int32_t* Bi, int32_t* cnt, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int64_t total = 0;
#pragma omp parallel reduction(+ : total)
  {
    unsigned char* seen = (unsigned char*)calloc((size_t)n + 1, 1);
    int32_t* list = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
#pragma omp for schedule(dynamic, 64)
    for (int32_t i = 0; i < m; i++) {
      int32_t L = 0;
      for (int32_t p = Ap[i]; p < Ap[i + 1]; p++) {
        int32_t j = Ai[p];
        for (int32_t q = Bp[j]; q < Bp[j + 1]; q++) {
          int32_t k = Bi[q];
          if (!seen[k]) { seen[k] = 1; list[L++] = k; }
        }
      }
      for (int32_t t = 0; t < L; t++) seen[list[t]] = 0;
      cnt[i] = L;
      total += L;
    }
    free(seen);
    free(list);
  }
  return total;
}

/* Symbolic phase (assemble): Cp[i+1]-Cp[i] = |U_{j in A_i} cols(B_j)|. */
ORC_API int64_t orc_spgemm_symbolic(int32_t m, int32_t n, const int32_t* Ap, const int32_t* Ai,
                                    const int32_t* Bp, const int32_t* Bi, int32_t* Cp, int nthreads) {
  int32_t* cnt = (int32_t*)calloc((size_t)m + 1, sizeof(int32_t));
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    unsigned char* seen = (unsigned char*)calloc((size_t)n + 1, 1);
    int32_t* list = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
#pragma omp for schedule(dynamic, 64)
    for (int32_t i = 0; i < m; i++) {
      int32_t L = 0;
      for (int32_t p = Ap[i]; p < Ap[i + 1]; p++) {
        int32_t j = Ai[p];
        for (int32_t q = Bp[j]; q < Bp[j + 1]; q++) {
          int32_t k = Bi[q];
          if (!seen[k]) { seen[k] = 1; list[L++] = k; }
        }
      }
      for (int32_t t = 0; t < L; t++) seen[list[t]] = 0;
      cnt[i] = L;
    }
    free(seen);
    free(list);
  }
  int64_t run = 0;
  Cp[0] = 0;
  for (int32_t i = 0; i < m; i++) {
    run += cnt[i];
    if (run > INT32_MAX) { /* the CSR output uses int32 row pointers */
      free(cnt);
      return -1;
    }
    Cp[i + 1] = (int32_t)run;
  }
  free(cnt);
  return run;
}

/* OpenMP Gustavson with a dense workspace and a per-row qsort of the index
 * list.  Cp must hold the symbolic result; Ci/Cx are caller-provided. */
ORC_API void orc_spgemm_gustavson_numeric(int32_t m, int32_t n,
                                          const int32_t* Ap, const int32_t* Ai, const double* Ax,
                                          const int32_t* Bp, const int32_t* Bi, const double* Bx,
                                          const int32_t* Cp, int32_t* Ci, double* Cx, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    double* wv = (double*)malloc(((size_t)n + 1) * sizeof(double));
    unsigned char* seen = (unsigned char*)calloc((size_t)n + 1, 1);
    int32_t* list = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
#pragma omp for schedule(dynamic, 64)
    for (int32_t i = 0; i < m; i++) {
      int32_t L = 0;
      for (int32_t p = Ap[i]; p < Ap[i + 1]; p++) {
        int32_t j = Ai[p];
        double a = Ax[p];
        for (int32_t q = Bp[j]; q < Bp[j + 1]; q++) {
          int32_t k = Bi[q];
          double prod = a * Bx[q];
          if (!seen[k]) {
            wv[k] = prod;
            seen[k] = 1;
            list[L++] = k;
          } else {
            wv[k] = wv[k] + prod;
          }
        }
      }
      qsort(list, (size_t)L, sizeof(int32_t), int_cmp);
      int32_t base = Cp[i];
      for (int32_t t = 0; t < L; t++) {
        int32_t k = list[t];
        Ci[base + t] = k;
        Cx[base + t] = wv[k];
        seen[k] = 0;
      }
    }
    free(wv);
    free(seen);
    free(list);
  }
}

/* Symbolic + numeric in one call (CSR_CSR_taco_parallel.h compute). */
ORC_API int64_t orc_spgemm_gustavson(int32_t m, int32_t n,
                                     const int32_t* Ap, const int32_t* Ai, const double* Ax,
                                     const int32_t* Bp, const int32_t* Bi, const double* Bx,
                                     int32_t* Cp, int32_t** Ci, double** Cx, int nthreads) {
  int64_t nnz = orc_spgemm_symbolic(m, n, Ap, Ai, Bp, Bi, Cp, nthreads);
  if (nnz < 0) {
    *Ci = NULL;
    *Cx = NULL;
    return -1;
  }
  int32_t* ci = (int32_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
  double* cx = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
  orc_spgemm_gustavson_numeric(m, n, Ap, Ai, Ax, Bp, Bi, Bx, Cp, ci, cx, nthreads);
  *Ci = ci;
  *Cx = cx;
  return nnz;
}

/* Per-row upper bound sum_{j in A_i} len(B_j). */
ORC_API void orc_spgemm_row_ub(int32_t m, const int32_t* Ap, const int32_t* Ai, const int32_t* Bp,
                               int64_t* ub) {
  for (int32_t i = 0; i < m; i++) {
    int64_t s = 0;
    for (int32_t p = Ap[i]; p < Ap[i + 1]; p++) {
      int32_t j = Ai[p];
      s += Bp[j + 1] - Bp[j];
    }
    ub[i] = s;
  }
}

/* Sorted-summation oracle: for each entry of the pattern (Cp, Ci), the
 * long-double sum of its products in j order, and sum |products|. */
ORC_API void orc_spgemm_refsum(int32_t m, int32_t n,
                               const int32_t* Ap, const int32_t* Ai, const double* Ax,
                               const int32_t* Bp, const int32_t* Bi, const double* Bx,
                               const int32_t* Cp, const int32_t* Ci, double* exact, double* abssum,
                               int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    long double* acc = (long double*)calloc((size_t)n + 1, sizeof(long double));
    long double* aab = (long double*)calloc((size_t)n + 1, sizeof(long double));
#pragma omp for schedule(dynamic, 64)
    for (int32_t i = 0; i < m; i++) {
      for (int32_t p = Ap[i]; p < Ap[i + 1]; p++) {
        int32_t j = Ai[p];
        for (int32_t q = Bp[j]; q < Bp[j + 1]; q++) {
          long double prod = (long double)Ax[p] * (long double)Bx[q];
          acc[Bi[q]] += prod;
          aab[Bi[q]] += prod < 0 ? -prod : prod;
        }
      }
      for (int32_t t = Cp[i]; t < Cp[i + 1]; t++) {
        int32_t k = Ci[t];
        exact[t] = (double)acc[k];
        abssum[t] = (double)aab[k];
        acc[k] = 0;
        aab[k] = 0;
      }
    }
    free(acc);
    free(aab);
  }
}

/* ------------------------------------------------------------------------- */
/* Outer-product SpGEMM with a coordinate workspace                           */
/* ------------------------------------------------------------------------- */
/*
 * C(i,k) = sum_j A(i,j) B(j,k) with A in DCSC (A1_pos/A1_crd list the
 * non-empty columns, A2_pos/A2_crd the rows of each) and B in DCSR, co-iterated
 * on j with a sorted merge (TACO_MIN, :228-264).  Each product (i,k,a*b) is
 * appended to a workspace of capacity w_cap (TryInsert_coord_t :66-79).  When
 * it is full the workspace is sorted by (i,k) (Sort_t :36-47) and merged with
 * reduction into the sorted all-array (Merge_coord_t :81-192).
 *
 * Deviation: the reference sorts with qsort, whose order among equal keys is
 * unspecified; this restatement sorts stably, so duplicates inside one
 * workspace are summed in append order (j, then i, then k).
 * Output row pointers are int64 (the 200k x 200k case exceeds int32).
 */
typedef struct {
  uint64_t key;
  double val;
  int64_t seq;
} orc_kvs;

static int kvs_cmp(const void* a, const void* b) {
  const orc_kvs* x = (const orc_kvs*)a;
  const orc_kvs* y = (const orc_kvs*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->seq < y->seq ? -1 : (x->seq > y->seq ? 1 : 0);
}

typedef struct {
  orc_kvs* ws;
  int64_t wsize, wcap;
  uint64_t* akey;
  double* aval;
  int64_t asize;
} orc_cws;

static void cws_flush(orc_cws* c) {
  if (c->wsize == 0) return;
  qsort(c->ws, (size_t)c->wsize, sizeof(orc_kvs), kvs_cmp);
  int64_t need = c->asize + c->wsize;
  uint64_t* nk = (uint64_t*)malloc((size_t)need * sizeof(uint64_t));
  double* nv = (double*)malloc((size_t)need * sizeof(double));
  int64_t a = 0, p = 0, t = -1;
  /* Merge with reduction: an element equal to the current target is added
   * into it (Merge_coord_t :125-177); equal (acc, content) pairs add
   * acc + content first (:129-134). */
  while (a < c->wsize || p < c->asize) {
    uint64_t key;
    double v;
    if (a < c->wsize && p < c->asize && c->ws[a].key == c->akey[p]) {
      key = c->ws[a].key;
      v = c->ws[a].val + c->aval[p];
      a++; p++;
    } else if (p >= c->asize || (a < c->wsize && c->ws[a].key < c->akey[p])) {
      key = c->ws[a].key;
      v = c->ws[a].val;
      a++;
    } else {
      key = c->akey[p];
      v = c->aval[p];
      p++;
    }
    if (t >= 0 && nk[t] == key) {
      nv[t] += v;
    } else {
      t++;
      nk[t] = key;
      nv[t] = v;
    }
  }
  free(c->akey);
  free(c->aval);
  c->akey = nk;
  c->aval = nv;
  c->asize = t + 1;
  c->wsize = 0;
}

ORC_API int64_t orc_outer_coord(int32_t m, int32_t n,
                                int32_t a_ncols, const int32_t* A1_crd, const int32_t* A2_pos,
                                const int32_t* A2_crd, const double* Ax,
                                int32_t b_nrows, const int32_t* B1_crd, const int32_t* B2_pos,
                                const int32_t* B2_crd, const double* Bx,
                                int32_t w_cap, int64_t* Cp, int32_t** Ci, double** Cx) {
  orc_cws c;
  memset(&c, 0, sizeof(c));
  c.wcap = w_cap > 0 ? w_cap : 1;
  c.ws = (orc_kvs*)malloc((size_t)c.wcap * sizeof(orc_kvs));
  c.akey = (uint64_t*)malloc(sizeof(uint64_t));
  c.aval = (double*)malloc(sizeof(double));
  int64_t seq = 0;
  (void)n;
  int32_t ja = 0, jb = 0;
  while (ja < a_ncols && jb < b_nrows) {
    int32_t j0a = A1_crd[ja], j0b = B1_crd[jb];
    int32_t j = j0a < j0b ? j0a : j0b;
    if (j0a == j && j0b == j) {
      for (int32_t pa = A2_pos[ja]; pa < A2_pos[ja + 1]; pa++) {
        int32_t i = A2_crd[pa];
        double a = Ax[pa];
        for (int32_t pb = B2_pos[jb]; pb < B2_pos[jb + 1]; pb++) {
          if (c.wsize == c.wcap) cws_flush(&c);
          orc_kvs* e = &c.ws[c.wsize++];
          e->key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)B2_crd[pb];
          e->val = a * Bx[pb];
          e->seq = seq++;
        }
      }
    }
    ja += (j0a == j);
    jb += (j0b == j);
  }
  cws_flush(&c);
  int64_t nnz = c.asize;
  int32_t* ci = (int32_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
  double* cx = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
  memset(Cp, 0, ((size_t)m + 1) * sizeof(int64_t));
  for (int64_t t = 0; t < nnz; t++) {
    Cp[(c.akey[t] >> 32) + 1]++;
    ci[t] = (int32_t)(c.akey[t] & 0xffffffffu);
    cx[t] = c.aval[t];
  }
  for (int32_t r = 0; r < m; r++) Cp[r + 1] += Cp[r];
  free(c.ws);
  free(c.akey);
  free(c.aval);
  *Ci = ci;
  *Cx = cx;
  return nnz;
}

/* Count of products sum_j |A(:,j)| * |B(j,:)| over matched j. */
ORC_API int64_t orc_outer_products(int32_t a_ncols, const int32_t* A1_crd, const int32_t* A2_pos,
                                   int32_t b_nrows, const int32_t* B1_crd, const int32_t* B2_pos) {
  int64_t s = 0;
  int32_t ja = 0, jb = 0;
  while (ja < a_ncols && jb < b_nrows) {
    int32_t j0a = A1_crd[ja], j0b = B1_crd[jb];
    int32_t j = j0a < j0b ? j0a : j0b;
    if (j0a == j && j0b == j)
      s += (int64_t)(A2_pos[ja + 1] - A2_pos[ja]) * (B2_pos[jb + 1] - B2_pos[jb]);
    ja += (j0a == j);
    jb += (j0b == j);
  }
  return s;
}

/* ------------------------------------------------------------------------- */
/* MTTKRP                                                                     */
/* ------------------------------------------------------------------------- */
/*
 * A(i,r) = sum_{k,l} B(k,l,i) C(k,r) D(l,r).  B is a 3-level CSF with levels
 * (k, l, i) (B1 compressed: B1_pos[0..1], B1_crd; B2_pos/B2_crd; B3_pos/B3_crd),
 * C and D are DCSR matrices (C1_pos[0..1], C1_crd, C2_pos, C2_crd, C_vals).
 * Co-iteration order (:276-332): k over B1 and C1, l over B2[k] and D1,
 * i over B3[l], r over C2[k] and D2[l].  The product is (B*C)*D (:296).
 * Workspace: the same bounded hash as the row-wise SpGEMM, keyed on (i, r).
 * Output: COO sorted by (i, r).
 */
ORC_API int64_t orc_mttkrp_hash(int32_t I, int32_t R,
                                int32_t nk, const int32_t* B1_crd, const int32_t* B2_pos,
                                const int32_t* B2_crd, const int32_t* B3_pos, const int32_t* B3_crd,
                                const double* Bx,
                                int32_t c_nrows, const int32_t* C1_crd, const int32_t* C2_pos,
                                const int32_t* C2_crd, const double* Cx,
                                int32_t d_nrows, const int32_t* D1_crd, const int32_t* D2_pos,
                                const int32_t* D2_crd, const double* Dx,
                                int32_t w_cap, int32_t** Oi, int32_t** Or, double** Ov) {
  orc_hws w;
  hws_init(&w, w_cap, 0);
  (void)I;
  (void)R;
  int32_t kb = 0, kc = 0;
  while (kb < nk && kc < c_nrows) {
    int32_t k0b = B1_crd[kb], k0c = C1_crd[kc];
    int32_t k = k0b < k0c ? k0b : k0c;
    if (k0b == k && k0c == k) {
      int32_t lb = B2_pos[kb], lb_end = B2_pos[kb + 1];
      int32_t ld = 0;
      while (lb < lb_end && ld < d_nrows) {
        int32_t l0b = B2_crd[lb], l0d = D1_crd[ld];
        int32_t l = l0b < l0d ? l0b : l0d;
        if (l0b == l && l0d == l) {
          for (int32_t ib = B3_pos[lb]; ib < B3_pos[lb + 1]; ib++) {
            int32_t i = B3_crd[ib];
            int32_t jc = C2_pos[kc], jc_end = C2_pos[kc + 1];
            int32_t jd = D2_pos[ld], jd_end = D2_pos[ld + 1];
            while (jc < jc_end && jd < jd_end) {
              int32_t j0c = C2_crd[jc], j0d = D2_crd[jd];
              int32_t r = j0c < j0d ? j0c : j0d;
              if (j0c == r && j0d == r) {
                double v = Bx[ib] * Cx[jc] * Dx[jd];
                uint64_t key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)r;
                hws_insert(&w, key, v);
              }
              jc += (j0c == r);
              jd += (j0d == r);
            }
          }
        }
        lb += (l0b == l);
        ld += (l0d == l);
      }
    }
    kb += (k0b == k);
    kc += (k0c == k);
  }
  hws_flush(&w);
  int64_t nnz = w.asize;
  int32_t* oi = (int32_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
  int32_t* orr = (int32_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
  double* ov = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
  for (int64_t t = 0; t < nnz; t++) {
    oi[t] = (int32_t)(w.akey[t] >> 32);
    orr[t] = (int32_t)(w.akey[t] & 0xffffffffu);
    ov[t] = w.aval[t];
  }
  hws_free(&w);
  *Oi = oi;
  *Or = orr;
  *Ov = ov;
  return nnz;
}

/* ------------------------------------------------------------------------- */
/* TTM (hash_il)                                                              */
/* ------------------------------------------------------------------------- */
/*
 * A(k,j,l) = sum_i B(k,i,j) C(i,l).  B is a 3-level CSF with levels
 * (k, i, j): B1 is the list of roots k (B1_crd, nk entries), B2_pos/B2_crd the
 * contracted coordinates i under each root, B3_pos/B3_crd the leaves j.
 * C is a DCSR matrix (C1_crd rows, C2_pos/C2_crd/C_vals).  Loops (:291-331):
 * k over B1, i over B2[k] co-iterated with C1, j over B3[i], l over C2[i].
 * Workspace keyed on (k, j, l); it never flushes because the buffer capacity
 * is max(INT32_MAX, 2*w_cap) (:206).  Output: COO sorted by (k, j, l).
 * Key packing: ((k * J) + j) * L + l in 64 bits (lexicographic order).
 */
ORC_API int64_t orc_ttm_hash_il(int32_t K, int32_t J, int32_t L,
                                int32_t nk, const int32_t* B1_crd, const int32_t* B2_pos,
                                const int32_t* B2_crd, const int32_t* B3_pos, const int32_t* B3_crd,
                                const double* Bx,
                                int32_t c_nrows, const int32_t* C1_crd, const int32_t* C2_pos,
                                const int32_t* C2_crd, const double* Cx,
                                int32_t** Ok, int32_t** Oj, int32_t** Ol, double** Ov) {
  orc_hws w;
  hws_init(&w, 1, 1);
  (void)K;
  for (int32_t kb = 0; kb < nk; kb++) {
    int32_t k = B1_crd[kb];
    int32_t ib = B2_pos[kb], ib_end = B2_pos[kb + 1];
    int32_t ic = 0;
    while (ib < ib_end && ic < c_nrows) {
      int32_t i0b = B2_crd[ib], i0c = C1_crd[ic];
      int32_t i = i0b < i0c ? i0b : i0c;
      if (i0b == i && i0c == i) {
        for (int32_t jb = B3_pos[ib]; jb < B3_pos[ib + 1]; jb++) {
          int32_t j = B3_crd[jb];
          for (int32_t lc = C2_pos[ic]; lc < C2_pos[ic + 1]; lc++) {
            int32_t l = C2_crd[lc];
            uint64_t key = ((uint64_t)k * (uint64_t)J + (uint64_t)j) * (uint64_t)L + (uint64_t)l;
            hws_insert(&w, key, Bx[jb] * Cx[lc]);
          }
        }
      }
      ib += (i0b == i);
      ic += (i0c == i);
    }
  }
  hws_flush(&w);
  int64_t nnz = w.asize;
  int32_t* ok = (int32_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
  int32_t* oj = (int32_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
  int32_t* ol = (int32_t*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
  double* ov = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
  for (int64_t t = 0; t < nnz; t++) {
    uint64_t key = w.akey[t];
    ol[t] = (int32_t)(key % (uint64_t)L);
    key /= (uint64_t)L;
    oj[t] = (int32_t)(key % (uint64_t)J);
    ok[t] = (int32_t)(key / (uint64_t)J);
    ov[t] = w.aval[t];
  }
  hws_free(&w);
  *Ok = ok;
  *Oj = oj;
  *Ol = ol;
  *Ov = ov;
  return nnz;
}

/* ------------------------------------------------------------------------- */
/* Coordinate workspace records: Sort and Merge of utils/wspace.h.            */

typedef struct {
  int32_t crd[2];
  double val;
} orc_wsrec;

/* esc_cmp (utils/wspace.h:24-31) compares with int subtraction; coordinates
 * are >= -1 there, so a three-way compare gives the same sign. */
static int ws_cmp_up(const void* pa, const void* pb) {
  const orc_wsrec* a = (const orc_wsrec*)pa;
  const orc_wsrec* b = (const orc_wsrec*)pb;
  for (int i = 0; i < 2; i++)
    if (a->crd[i] != b->crd[i]) return a->crd[i] < b->crd[i] ? -1 : 1;
  return 0;
}

static int ws_cmp_down(const void* pa, const void* pb) { return ws_cmp_up(pb, pa); }

/* Sort (utils/wspace.h:206-217): ascending returns size; descending returns
 * one past the last record whose crd[0] is not -1 (0 if none). */
ORC_API int64_t orc_wspace_sort(orc_wsrec* w, int64_t size, int rev) {
  if (rev) {
    qsort(w, (size_t)size, sizeof(orc_wsrec), ws_cmp_down);
    for (int64_t i = size - 1; i >= 0; i--)
      if (w[i].crd[0] != -1) return i + 1;
    /* the reference falls through to the ascending sort when every slot is
     * empty; an all-empty workspace stays all-empty either way */
  }
  qsort(w, (size_t)size, sizeof(orc_wsrec), ws_cmp_up);
  return size;
}

/* Merge (utils/wspace.h:62-204): two-pointer merge of a sorted COO
 * (c0, c1, v; coo_size entries, capacity >= coo_size + acc_size) with the
 * accumulator (ascending, or descending when rev), adding the value of an
 * equal coordinate as acc.val + coo.val.  Returns the merged size. */
ORC_API int64_t orc_wspace_merge(int32_t* c0, int32_t* c1, double* v, int64_t coo_size, const orc_wsrec* acc,
                                 int64_t acc_size, int rev) {
  const int64_t n = coo_size + acc_size;
  int32_t* t0 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* t1 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  double* tv = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  int64_t a = 0, c = 0, t = 0;
  while (a < acc_size && c < coo_size) {
    const orc_wsrec* r = &acc[rev ? acc_size - 1 - a : a];
    orc_wsrec q;
    q.crd[0] = c0[c];
    q.crd[1] = c1[c];
    const int cmp = ws_cmp_up(r, &q);
    if (cmp == 0) {
      t0[t] = r->crd[0];
      t1[t] = r->crd[1];
      tv[t] = r->val + v[c];
      a++, c++, t++;
    } else if (cmp < 0) {
      t0[t] = r->crd[0];
      t1[t] = r->crd[1];
      tv[t] = r->val;
      a++, t++;
    } else {
      t0[t] = c0[c];
      t1[t] = c1[c];
      tv[t] = v[c];
      c++, t++;
    }
  }
  for (; a < acc_size; a++, t++) {
    const orc_wsrec* r = &acc[rev ? acc_size - 1 - a : a];
    t0[t] = r->crd[0];
    t1[t] = r->crd[1];
    tv[t] = r->val;
  }
  for (; c < coo_size; c++, t++) {
    t0[t] = c0[c];
    t1[t] = c1[c];
    tv[t] = v[c];
  }
  memcpy(c0, t0, sizeof(int32_t) * (size_t)t);
  memcpy(c1, t1, sizeof(int32_t) * (size_t)t);
  memcpy(v, tv, sizeof(double) * (size_t)t);
  free(t0);
  free(t1);
  free(tv);
  return t;
}

ORC_API void orc_free(void* p) { free(p); }

/* transpose(A, B) of benchmark/CSC_CSR_T/nonfuse/CSC_CSR_T_hash.h:352-400,
 * At = A^T: the entries of each column of A are counted (:366-371), the
 * counts prefixed into At's row pointers (:372-376), and A's rows scattered
 * in row order through one cursor per column (:380-388), so each row of At
 * lists its columns (A's rows) in increasing order. */
ORC_API void orc_csr_transpose(int32_t m, int32_t n, const int32_t* Ap, const int32_t* Ai, const double* Ax,
                               int32_t* Atp, int32_t* Ati, double* Atx) {
  int32_t* cur = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
  for (int32_t i = 0; i < m; i++)
    for (int32_t p = Ap[i]; p < Ap[i + 1]; p++) cur[Ai[p] + 1]++;
  for (int32_t c = 0; c < n; c++) cur[c + 1] += cur[c];
  memcpy(Atp, cur, ((size_t)n + 1) * sizeof(int32_t));
  for (int32_t i = 0; i < m; i++)
    for (int32_t p = Ap[i]; p < Ap[i + 1]; p++) {
      const int32_t q = cur[Ai[p]]++;
      Ati[q] = i;
      Atx[q] = Ax[p];
    }
  free(cur);
}
