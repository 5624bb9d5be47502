/*
 * gen_native.c -- seeded synthetic inputs that are too large for numpy:
 * spwsgen_coo3 (3-order COO tensors with an exact count of distinct
 * nonzeros and Zipf or uniform per-mode indices) and spwsgen_csr (CSR rows
 * of given lengths with distinct uniform columns).
 *
 * spwsgen_coo3:
 *
 * Definition of the tensor (independent of the thread count):
 *   - mode m draws rank r in [1, d_m] with P(r) proportional to r^-s_m
 *     (inverse CDF), mapped to an index by a seeded permutation of [0, d_m);
 *     s_m <= 0 draws the index uniformly;
 *   - draw t uses the uniforms U(seed, stream, 3t + m), where U is splitmix64
 *     of a counter (no sequential RNG state);
 *   - the tensor's coordinates are the first nnz distinct triples in draw
 *     order, sorted lexicographically; values are U[lo, hi) indexed by the
 *     sorted position.
 * Draws are processed in batches; a concurrent open-addressing set keeps each
 * distinct key with the smallest draw index that produced it, so the batch
 * that crosses nnz keeps exactly the earliest-drawn new keys.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GEN_API __attribute__((visibility("default")))

static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

static inline double u01(uint64_t base, uint64_t ctr) {
  return (double)(splitmix64(base ^ splitmix64(ctr)) >> 11) * (1.0 / 9007199254740992.0);
}

static int64_t upper_bound_d(const double* cdf, int64_t n, double u) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (cdf[mid] <= u) lo = mid + 1; else hi = mid;
  }
  return lo < n ? lo : n - 1;
}

typedef struct {
  int64_t d;
  double s;
  double* cdf;
  int64_t* perm;
} mode_t_;

static void mode_init(mode_t_* m, int64_t d, double s, uint64_t base, uint64_t which) {
  m->d = d;
  m->s = s;
  m->cdf = NULL;
  m->perm = (int64_t*)malloc((size_t)d * sizeof(int64_t));
  for (int64_t i = 0; i < d; i++) m->perm[i] = i;
  /* Fisher-Yates with counter uniforms */
  for (int64_t i = d - 1; i > 0; i--) {
    int64_t j = (int64_t)(u01(base ^ (0x1234567ULL + which), (uint64_t)i) * (double)(i + 1));
    if (j > i) j = i;
    int64_t t = m->perm[i];
    m->perm[i] = m->perm[j];
    m->perm[j] = t;
  }
  if (s > 0) {
    m->cdf = (double*)malloc((size_t)d * sizeof(double));
    double acc = 0;
    for (int64_t r = 0; r < d; r++) {
      acc += pow((double)(r + 1), -s);
      m->cdf[r] = acc;
    }
    for (int64_t r = 0; r < d; r++) m->cdf[r] /= acc;
  }
}

static inline int64_t mode_draw(const mode_t_* m, double u) {
  if (m->s > 0) return m->perm[upper_bound_d(m->cdf, m->d, u)];
  int64_t i = (int64_t)(u * (double)m->d);
  return i < m->d ? i : m->d - 1;
}

static void radix_sort_u64(uint64_t* a, uint64_t* tmp, int64_t n) {
  for (int shift = 0; shift < 64; shift += 16) {
    int64_t* cnt = (int64_t*)calloc(65537, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) cnt[((a[i] >> shift) & 0xffff) + 1]++;
    for (int b = 0; b < 65536; b++) cnt[b + 1] += cnt[b];
    for (int64_t i = 0; i < n; i++) tmp[cnt[(a[i] >> shift) & 0xffff]++] = a[i];
    memcpy(a, tmp, (size_t)n * sizeof(uint64_t));
    free(cnt);
  }
}

static int cmp_pair(const void* x, const void* y) {
  const uint64_t* a = (const uint64_t*)x;
  const uint64_t* b = (const uint64_t*)y;
  return a[0] < b[0] ? -1 : (a[0] > b[0] ? 1 : 0);
}

/* Returns nnz on success, -1 when the space holds fewer than nnz triples or
 * the draw budget runs out. out_idx: nnz x 3 int32 (row-major), out_vals: nnz. */
GEN_API int64_t spwsgen_coo3(int64_t d0, int64_t d1, int64_t d2, double s0, double s1, double s2, int64_t nnz,
                             uint64_t seed, uint64_t stream, double lo, double hi, int32_t* out_idx,
                             double* out_vals, int nthreads) {
  if (nnz <= 0) return 0;
  if ((double)d0 * (double)d1 * (double)d2 < (double)nnz) return -1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const uint64_t base = splitmix64(seed * 0x100000001b3ULL ^ splitmix64(stream + 0x51ed27ULL));
  mode_t_ md[3];
  mode_init(&md[0], d0, s0, base, 0);
  mode_init(&md[1], d1, s1, base, 1);
  mode_init(&md[2], d2, s2, base, 2);
  const uint64_t draw_base = splitmix64(base ^ 0xd1b54a32d192ed03ULL);

  const uint64_t batch = (uint64_t)(nnz > 1000000 ? nnz : 1000000);
  /* at most nnz kept keys plus one batch of new ones live in the set */
  uint64_t cap = 1024;
  while (cap < 2 * ((uint64_t)nnz + batch)) cap <<= 1;
  const uint64_t mask = cap - 1;
  const uint64_t EMPTY = ~0ULL;
  uint64_t* tkey = (uint64_t*)malloc(cap * sizeof(uint64_t));
  uint64_t* tdraw = (uint64_t*)malloc(cap * sizeof(uint64_t));
  memset(tkey, 0xff, cap * sizeof(uint64_t));
  memset(tdraw, 0xff, cap * sizeof(uint64_t));
  int64_t have = 0;
  uint64_t t0 = 0;
  const uint64_t max_draws = (uint64_t)nnz * 4000ULL + 100000000ULL;
  int64_t result = -1;
  while (t0 < max_draws) {
    int64_t added = 0;
#pragma omp parallel for schedule(static) reduction(+ : added)
    for (int64_t q = 0; q < (int64_t)batch; q++) {
      const uint64_t t = t0 + (uint64_t)q;
      const uint64_t i = (uint64_t)mode_draw(&md[0], u01(draw_base, 3 * t));
      const uint64_t j = (uint64_t)mode_draw(&md[1], u01(draw_base, 3 * t + 1));
      const uint64_t k = (uint64_t)mode_draw(&md[2], u01(draw_base, 3 * t + 2));
      const uint64_t key = (i * (uint64_t)d1 + j) * (uint64_t)d2 + k;
      uint64_t h = splitmix64(key) & mask;
      for (;;) {
        uint64_t cur = __atomic_load_n(&tkey[h], __ATOMIC_RELAXED);
        if (cur == EMPTY) {
          uint64_t expect = EMPTY;
          if (__atomic_compare_exchange_n(&tkey[h], &expect, key, 0, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED)) {
            added++;
            cur = key;
          } else {
            cur = expect;
          }
        }
        if (cur == key) {
          uint64_t old = __atomic_load_n(&tdraw[h], __ATOMIC_RELAXED);
          while (t < old && !__atomic_compare_exchange_n(&tdraw[h], &old, t, 1, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED)) {
          }
          break;
        }
        h = (h + 1) & mask;
      }
    }
    if (have + added >= nnz) {
      /* keep the earliest-drawn new keys of this batch */
      const int64_t deficit = nnz - have;
      int64_t nnew = 0;
      for (uint64_t s = 0; s < cap; s++)
        if (tkey[s] != EMPTY && tdraw[s] >= t0) nnew++;
      uint64_t* cand = (uint64_t*)malloc((size_t)(nnew > 0 ? nnew : 1) * 2 * sizeof(uint64_t));
      int64_t c = 0;
      for (uint64_t s = 0; s < cap; s++)
        if (tkey[s] != EMPTY && tdraw[s] >= t0) {
          cand[2 * c] = tdraw[s];
          cand[2 * c + 1] = s;
          c++;
        }
      qsort(cand, (size_t)nnew, 2 * sizeof(uint64_t), cmp_pair);
      for (int64_t q = deficit; q < nnew; q++) tkey[cand[2 * q + 1]] = EMPTY;
      free(cand);
      have = nnz;
      result = nnz;
      break;
    }
    have += added;
    t0 += batch;
  }
  if (result == nnz) {
    uint64_t* keys = (uint64_t*)malloc((size_t)nnz * sizeof(uint64_t));
    uint64_t* tmp = (uint64_t*)malloc((size_t)nnz * sizeof(uint64_t));
    int64_t c = 0;
    for (uint64_t s = 0; s < cap; s++)
      if (tkey[s] != EMPTY) keys[c++] = tkey[s];
    radix_sort_u64(keys, tmp, nnz);
    const uint64_t vbase = splitmix64(base ^ 0x8cb92ba72f3d8dd7ULL);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nnz; q++) {
      const uint64_t key = keys[q];
      out_idx[3 * q + 2] = (int32_t)(key % (uint64_t)d2);
      out_idx[3 * q + 1] = (int32_t)((key / (uint64_t)d2) % (uint64_t)d1);
      out_idx[3 * q + 0] = (int32_t)(key / ((uint64_t)d1 * (uint64_t)d2));
      out_vals[q] = lo + (hi - lo) * u01(vbase, (uint64_t)q);
    }
    free(keys);
    free(tmp);
  }
  free(tkey);
  free(tdraw);
  for (int m = 0; m < 3; m++) {
    free(md[m].cdf);
    free(md[m].perm);
  }
  return result;
}

/* Row r gets lengths[r] distinct columns uniform in [0, n), sorted, and
 * values U[lo, hi) without exact zeros.  Draw t of row r uses
 * U(seed, stream, r, t); long rows (4 * len > n) are sampled by a partial
 * Fisher-Yates shuffle of [0, n) driven by the same uniforms.  pos[nrows+1]
 * must be the exclusive prefix of lengths (computed by the caller). */
static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

GEN_API int spwsgen_csr(int64_t nrows, int64_t n, const int32_t* lengths, const int64_t* pos, uint64_t seed,
                        uint64_t stream, double lo, double hi, int32_t* crd, double* vals, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  for (int64_t r = 0; r < nrows; r++)
    if (lengths[r] < 0 || lengths[r] > n) return -1;
  const uint64_t base = splitmix64(seed * 0x100000001b3ULL ^ splitmix64(stream + 0x6a09e667ULL));
  int bad = 0;
#pragma omp parallel reduction(| : bad)
  {
    int32_t* scratch = NULL;
    int64_t scap = 0;
#pragma omp for schedule(dynamic, 64)
    for (int64_t r = 0; r < nrows; r++) {
      const int64_t L = lengths[r];
      int32_t* out = crd + pos[r];
      const uint64_t rb = splitmix64(base ^ splitmix64((uint64_t)r + 0x3c6ef372ULL));
      uint64_t t = 0;
      if (L * 4 > n) {
        if (scap < n) {
          free(scratch);
          scratch = (int32_t*)malloc((size_t)n * sizeof(int32_t));
          scap = n;
        }
        for (int64_t i = 0; i < n; i++) scratch[i] = (int32_t)i;
        for (int64_t i = 0; i < L; i++) {
          int64_t j = i + (int64_t)(u01(rb, t++) * (double)(n - i));
          if (j >= n) j = n - 1;
          int32_t tmp = scratch[i];
          scratch[i] = scratch[j];
          scratch[j] = tmp;
        }
        memcpy(out, scratch, (size_t)L * sizeof(int32_t));
        qsort(out, (size_t)L, sizeof(int32_t), cmp_i32);
      } else {
        int64_t have = 0;
        while (have < L) {
          for (int64_t i = have; i < L; i++) {
            int64_t c = (int64_t)(u01(rb, t++) * (double)n);
            out[i] = (int32_t)(c < n ? c : n - 1);
          }
          qsort(out, (size_t)L, sizeof(int32_t), cmp_i32);
          int64_t w = 0;
          for (int64_t i = 0; i < L; i++)
            if (w == 0 || out[i] != out[w - 1]) out[w++] = out[i];
          have = w;
        }
      }
      const uint64_t vb = splitmix64(rb ^ 0x510e527fade682d1ULL);
      for (int64_t i = 0; i < L; i++) {
        double v = 0.0;
        uint64_t k = 0;
        while (v == 0.0) v = lo + (hi - lo) * u01(vb, (uint64_t)i * 1024 + k++);
        vals[pos[r] + i] = v;
      }
    }
    free(scratch);
  }
  return bad ? -1 : 0;
}

/* Upper bound of the output entries of A*B (sum over A's entries of the
 * matching B row lengths) and of its largest row: the numbers a batch's
 * memory-based admission is decided on. */
GEN_API void spwsgen_row_products(int64_t nrows, const int32_t* A_pos, const int32_t* A_crd, const int32_t* B_pos,
                                  int64_t* total, int64_t* max_row, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int64_t t = 0, mx = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : t) reduction(max : mx)
  for (int64_t r = 0; r < nrows; r++) {
    int64_t s = 0;
    for (int32_t p = A_pos[r]; p < A_pos[r + 1]; p++) s += B_pos[A_crd[p] + 1] - B_pos[A_crd[p]];
    t += s;
    if (s > mx) mx = s;
  }
  *total = t;
  *max_row = mx;
}
