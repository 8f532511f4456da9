/*
 * lightne_oracle.c -- CPU restatement of LightNE's stage-1 sparsifier and of the
 * sparse products used by stages 2 and 3.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * entry point and the cpu_baseline leg of bench.py load it.  The product
 * library (lightne_amd/csrc) never links or calls it.
 *
 * Every function cites the reference lines it restates (paths relative to the
 * LightNE source tree).  Parity status of each function is recorded in the
 * header comment of that function:
 *   - "pinned by" names the independent check that would fail on a plausible
 *     mistake (golden vectors from the reference-header harness, closed forms,
 *     brute force on tiny inputs).
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off -shared).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORC_EMPTY UINT64_MAX

/* pbbslib/utilities.h:154-159 -- splitmix64 finalizer used as hash64_2.
 * pinned by: Python big-int restatement + golden values from the
 * reference-header harness (tests/test_oracle.py). */
uint64_t orc_hash64_2(uint64_t x) {
  x = (x ^ (x >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
  x = (x ^ (x >> 27)) * UINT64_C(0x94d049bb133111eb);
  x = x ^ (x >> 31);
  return x;
}

/* pbbslib/random.h:38-51.  state s: rand() = H(s), next() = fork(0) = H(H(s)),
 * fork(i) = H(H(i + s)).  Only these three operations are used by the path. */
static inline uint64_t r_rand(uint64_t s) { return orc_hash64_2(s); }
static inline uint64_t r_next(uint64_t s) { return orc_hash64_2(orc_hash64_2(s)); }
static inline uint64_t r_fork(uint64_t s, uint64_t i) { return orc_hash64_2(orc_hash64_2(i + s)); }

/* PathEmbed.h:25-30 map_uint64_t: (double)u / 2^64 (round-to-nearest conversion). */
static inline double map_u64(uint64_t u) { return (double)u / 18446744073709551616.0; }

/* PathEmbed.h:89,99,108-111 sampling scalars.
 * out_d = {logn, M, one_over_lambda, fractional_bits}
 * out_u = {exp_rv_rounded, table_size}
 * Evaluation order follows the reference expression trees:
 *   table_size = max((size_t)(walk_len*n * logn * mem_ratio), tablesz)
 *   M          = walk_len*n * logn * sample_ratio          (size_t*size_t, then double)
 * pinned by: the reference-header harness prints the same scalars (golden). */
void orc_sample_scalars(uint64_t n, uint64_t m, uint64_t walk_len, float sample_ratio,
                        float mem_ratio, uint64_t tablesz, double* out_d, uint64_t* out_u) {
  double logn = log((double)n);
  uint64_t tn = walk_len * n;
  uint64_t ts = (uint64_t)((double)tn * logn * (double)mem_ratio);
  if (tablesz > ts) ts = tablesz;
  double M = (double)tn * logn * (double)sample_ratio;
  double inv_lambda = M / (double)m;
  uint64_t e = (uint64_t)inv_lambda;
  double frac = inv_lambda - (double)e;
  out_d[0] = logn; out_d[1] = M; out_d[2] = inv_lambda; out_d[3] = frac;
  out_u[0] = e; out_u[1] = ts;
}

/* ---------------------------------------------------------------------------
 * Concurrent open-addressing table with linear probing (par_table.h:60-109):
 * slot = hash64_2(key) & mask, empty key = SIZE_MAX.  The oracle keeps the
 * count of samples per key (u64) and the summed weight (double) separately so
 * that the integer count is checkable bit-for-bit in both sampling modes.
 * ------------------------------------------------------------------------- */
typedef struct {
  uint64_t cap, mask;
  uint64_t* keys;
  uint64_t* cnt;
  double* wgh;
  int overflow;
} orc_table;

static void tab_init(orc_table* t, uint64_t cap) {
  t->cap = cap; t->mask = cap - 1; t->overflow = 0;
  t->keys = (uint64_t*)malloc(cap * sizeof(uint64_t));
  t->cnt = (uint64_t*)calloc(cap, sizeof(uint64_t));
  t->wgh = (double*)calloc(cap, sizeof(double));
  for (uint64_t i = 0; i < cap; ++i) t->keys[i] = ORC_EMPTY;
}
static void tab_free(orc_table* t) { free(t->keys); free(t->cnt); free(t->wgh); }

static inline void atomic_add_double(double* p, double v) {
  uint64_t* ip = (uint64_t*)p;
  uint64_t old = __atomic_load_n(ip, __ATOMIC_RELAXED);
  for (;;) {
    double d; memcpy(&d, &old, 8); d += v;
    uint64_t nw; memcpy(&nw, &d, 8);
    if (__atomic_compare_exchange_n(ip, &old, nw, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return;
  }
}

/* par_table.h:86-109 insert_add (key CAS into an empty slot, else add). */
static inline void tab_insert(orc_table* t, uint64_t key, double w) {
  uint64_t h = orc_hash64_2(key) & t->mask;
  for (uint64_t probes = 0; probes <= t->cap; ++probes) {
    uint64_t cur = __atomic_load_n(&t->keys[h], __ATOMIC_RELAXED);
    if (cur == ORC_EMPTY) {
      uint64_t expect = ORC_EMPTY;
      if (__atomic_compare_exchange_n(&t->keys[h], &expect, key, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
        cur = key;
      else
        cur = expect;
    }
    if (cur == key) {
      __atomic_fetch_add(&t->cnt[h], 1, __ATOMIC_RELAXED);
      atomic_add_double(&t->wgh[h], w);
      return;
    }
    h = (h + 1) & t->mask;
  }
  t->overflow = 1;
}

/* PathEmbed.h:37-49 random_walk: walk_len steps, i = rand % deg(u) (u64 % u64),
 * u = get_ith_out_neighbor(u, i), state advances.  ith = 0: neighbour i in
 * input order (ligra/vertex.h:296-300).  ith = 1: a graph read from the
 * compressed format, whose lookup (ligra/encodings/byte_pd_amortized.h:653-688)
 * decodes i - start - 1 differences after block start = 64 (i / 64), i.e.
 * returns neighbour i - 1 unless i is a multiple of 64.
 * pinned by: the harness's dump of the reference's get_ith_out_neighbor on
 * compressed files (every index of three graphs). */
static inline uint32_t walk(const uint64_t* off, const uint32_t* adj, uint32_t u,
                            uint64_t len, uint64_t* s, uint64_t* steps, int ith) {
  for (uint64_t i = 0; i < len; ++i) {
    uint64_t d = off[u + 1] - off[u];
    uint64_t idx = r_rand(*s) % d;
    if (ith && (idx % 64) != 0) idx -= 1;
    u = adj[off[u] + idx];
    *s = r_next(*s);
  }
  *steps += len;
  return u;
}

/* PathEmbed.h:51-58 path_sample with walk length r. */
static inline void path_sample(const uint64_t* off, const uint32_t* adj, uint32_t u, uint32_t v,
                               uint64_t r, uint64_t* s, uint32_t* x, uint32_t* z, uint64_t* steps, int ith) {
  uint32_t k = (uint32_t)(r_rand(*s) % r + 1); *s = r_next(*s);
  *x = walk(off, adj, u, (uint64_t)k - 1, s, steps, ith);
  *z = walk(off, adj, v, r - k, s, steps, ith);
}

typedef struct {
  uint64_t nkeys;      /* distinct keys */
  uint64_t* keys;      /* sorted ascending */
  uint64_t* cnt;       /* samples per key */
  double* wgh;         /* summed weight per key */
  uint64_t samples;    /* emitted samples */
  uint64_t attempts;   /* per-edge loop trips (= sum of exp_rv) */
  uint64_t steps;      /* total random-walk steps */
  int overflow;
} orc_sample_result;

/* LSD radix sort of (key, idx) by key over the low `bits` bits. */
static void radix_sort_u64_idx(uint64_t* keys, uint64_t* idx, uint64_t n, int bits) {
  uint64_t* kbuf = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
  uint64_t* ibuf = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
  uint64_t *ka = keys, *ia = idx, *kb = kbuf, *ib = ibuf;
  for (int shift = 0; shift < bits; shift += 8) {
    uint64_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for (uint64_t i = 0; i < n; ++i) cnt[((ka[i] >> shift) & 255) + 1]++;
    for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t p = cnt[(ka[i] >> shift) & 255]++;
      kb[p] = ka[i]; ib[p] = ia[i];
    }
    uint64_t* t = ka; ka = kb; kb = t;
    t = ia; ia = ib; ib = t;
  }
  if (ka != keys) {
    memcpy(keys, ka, n * sizeof(uint64_t));
    memcpy(idx, ia, n * sizeof(uint64_t));
  }
  free(kbuf); free(ibuf);
}

static int bits_for(uint64_t n) { int b = 0; while (((uint64_t)1 << b) < n) ++b; return b; }

/* PathEmbed.h:128-172 per-directed-edge sampler, both -sample modes, followed
 * by the table dump (PathEmbed.h:175-188 / par_table.h:157-161) sorted by key.
 *   seed      : pbbs::random state handed to generate_trunc_log_matrix_v2
 *   cap_log2  : table capacity 2^cap_log2 (must exceed the distinct keys)
 *   nthreads  : OpenMP threads for the edge loop (1 = serial, deterministic weights)
 *   u_begin, u_end, u_stride : sample only the edges of source vertices
 *               u_begin, u_begin + u_stride, ... < u_end (u_end = 0: n)
 *   ith       : neighbour lookup of the walks (0 plain, 1 compressed graph)
 * pinned by: golden key->count tables from the reference-header harness;
 * T=1 closed form; zero-state fixed point; sum(cnt) == samples invariant. */
orc_sample_result* orc_sample(uint64_t n, const uint64_t* off, const uint32_t* adj, uint64_t seed,
                              uint64_t walk_len, int sample, uint64_t exp_rv, double frac,
                              double logn, int cap_log2, int nthreads, uint64_t u_begin, uint64_t u_end,
                              uint64_t u_stride, int ith) {
  orc_table t;
  tab_init(&t, (uint64_t)1 << cap_log2);
  uint64_t samples = 0, attempts = 0, steps = 0;
  if (nthreads < 1) nthreads = 1;
  if (u_end == 0 || u_end > n) u_end = n;
  if (u_stride == 0) u_stride = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(+ : samples, attempts, steps)
  for (uint64_t u = u_begin; u < u_end; u += u_stride) {
    for (uint64_t j = off[u]; j < off[u + 1]; ++j) {
      uint32_t v = adj[j];
      uint64_t uv = (u << 32) | (uint64_t)v;
      uint64_t s = r_fork(seed, uv);
      uint64_t e = exp_rv;
      if (frac > 0.0000000001) {
        e += (map_u64(r_rand(s)) < frac) ? 1 : 0;
        s = r_next(s);
      }
      attempts += e;
      if (!sample) {
        for (uint64_t k = 0; k < e; ++k) {
          uint64_t r = (uint32_t)(r_rand(s) % walk_len + 1);
          s = r_next(s);
          uint32_t x, z;
          path_sample(off, adj, (uint32_t)u, v, r, &s, &x, &z, &steps, ith);
          uint64_t key = ((uint64_t)(x < z ? x : z) << 32) + (uint64_t)(x < z ? z : x);
          tab_insert(&t, key, 1.0);
          samples++;
        }
      } else {
        double du = (double)(off[u + 1] - off[u]);
        double dv = (double)(off[v + 1] - off[v]);
        double p = (1.0 / du + 1.0 / dv) * logn;
        if (p > 1.0) p = 1.0;
        double ipw = 1.0 / p;
        for (uint64_t k = 0; k < e; ++k) {
          double q = map_u64(r_rand(s));
          s = r_next(s);
          if (q < p) {
            uint64_t r = (uint32_t)(r_rand(s) % walk_len + 1);
            s = r_next(s);
            uint32_t x, z;
            path_sample(off, adj, (uint32_t)u, v, r, &s, &x, &z, &steps, ith);
            uint64_t key = ((uint64_t)(x < z ? x : z) << 32) + (uint64_t)(x < z ? z : x);
            tab_insert(&t, key, ipw);
            samples++;
          }
        }
      }
    }
  }
  orc_sample_result* res = (orc_sample_result*)calloc(1, sizeof(orc_sample_result));
  res->overflow = t.overflow;
  res->samples = samples; res->attempts = attempts; res->steps = steps;
  uint64_t nk = 0;
  for (uint64_t i = 0; i < t.cap; ++i) nk += (t.keys[i] != ORC_EMPTY);
  uint64_t* keys = (uint64_t*)malloc((nk ? nk : 1) * sizeof(uint64_t));
  uint64_t* idx = (uint64_t*)malloc((nk ? nk : 1) * sizeof(uint64_t));
  uint64_t j = 0;
  for (uint64_t i = 0; i < t.cap; ++i)
    if (t.keys[i] != ORC_EMPTY) { keys[j] = t.keys[i]; idx[j] = i; ++j; }
  int nb = 32 + bits_for(n);
  radix_sort_u64_idx(keys, idx, nk, nb > 64 ? 64 : nb);
  res->nkeys = nk;
  res->keys = keys;
  res->cnt = (uint64_t*)malloc((nk ? nk : 1) * sizeof(uint64_t));
  res->wgh = (double*)malloc((nk ? nk : 1) * sizeof(double));
  for (uint64_t i = 0; i < nk; ++i) { res->cnt[i] = t.cnt[idx[i]]; res->wgh[i] = t.wgh[idx[i]]; }
  free(idx);
  tab_free(&t);
  return res;
}

uint64_t orc_sample_nkeys(const orc_sample_result* r) { return r->nkeys; }
void orc_sample_stats(const orc_sample_result* r, uint64_t* out) {
  out[0] = r->samples; out[1] = r->attempts; out[2] = r->steps; out[3] = (uint64_t)r->overflow;
}
void orc_sample_copy(const orc_sample_result* r, uint64_t* keys, uint64_t* cnt, double* wgh) {
  memcpy(keys, r->keys, r->nkeys * sizeof(uint64_t));
  memcpy(cnt, r->cnt, r->nkeys * sizeof(uint64_t));
  memcpy(wgh, r->wgh, r->nkeys * sizeof(double));
}
void orc_sample_free(orc_sample_result* r) {
  if (!r) return;
  free(r->keys); free(r->cnt); free(r->wgh); free(r);
}

/* ---------------------------------------------------------------------------
 * PathEmbed.h:175-311: total weight, trunc-log and CSR assembly.
 *   M_exact = sum of table values                         (:175-188)
 *   factor  = vol * m / M_exact / 2 / negative, vol = m   (:199-200)
 *   per key: w = (u != v) ? val : 2 val; adj = w*factor/d_u/d_v;
 *            keep log(adj) iff log(adj) > 0               (:203-218)
 *   CSR: row u gets (v, log), row v gets (u, log) unless u == v or upper
 *   (:278-311); values cast to float.  Rows are emitted sorted by column
 *   (the reference's within-row order is unspecified).
 * keys must be sorted ascending (as orc_sample returns them); a sequential
 * scatter in key order then leaves every row sorted: row r receives its
 * entries (c, r), c < r, in increasing c before its entries (r, c), c >= r.
 * pinned by: harness golden trunc-log values; scipy/numpy restatement of the
 * same formula over the golden table is NOT used as a pin; the closed form
 * E[adj] (NetMF matrix) is checked statistically on tiny graphs.
 * ------------------------------------------------------------------------- */
typedef struct {
  uint64_t n, nnz, kept_upper;
  double M_exact, factor;
  int64_t* row_ptr;
  int32_t* col;
  float* val;
} orc_csr;

orc_csr* orc_trunc_log_csr(uint64_t n, const uint64_t* off, const uint64_t* keys, const double* vals,
                           uint64_t nkeys, uint64_t negative, int upper, double M_exact_given) {
  uint64_t m = off[n];
  double M_exact = 0.0;
  /* the sum order of :187 (pbbs blocked reduce) only matters for -sample 1;
   * a caller may pass the total it wants the formula evaluated with */
  if (M_exact_given > 0.0)
    M_exact = M_exact_given;
  else
    for (uint64_t i = 0; i < nkeys; ++i) M_exact += vals[i];
  double vol = (double)m;
  double factor = vol * (double)m / M_exact / 2 / (double)negative;
  orc_csr* c = (orc_csr*)calloc(1, sizeof(orc_csr));
  c->n = n; c->M_exact = M_exact; c->factor = factor;
  float* lw = (float*)malloc((nkeys ? nkeys : 1) * sizeof(float));
  unsigned char* keep = (unsigned char*)malloc(nkeys ? nkeys : 1);
  int64_t* cnt = (int64_t*)calloc(n + 1, sizeof(int64_t));
  uint64_t kept = 0;
  for (uint64_t i = 0; i < nkeys; ++i) {
    uint32_t u = (uint32_t)(keys[i] >> 32), v = (uint32_t)keys[i];
    double du = (double)(off[u + 1] - off[u]), dv = (double)(off[v + 1] - off[v]);
    double w = (u != v) ? vals[i] : vals[i] * 2;
    double adj = w * factor / du / dv;
    double l = log(adj);
    keep[i] = l > 0.0;
    lw[i] = (float)l;
    if (keep[i]) {
      kept++;
      cnt[u]++;
      if (!upper && u != v) cnt[v]++;
    }
  }
  c->kept_upper = kept;
  c->row_ptr = (int64_t*)malloc((n + 1) * sizeof(int64_t));
  c->row_ptr[0] = 0;
  for (uint64_t i = 0; i < n; ++i) c->row_ptr[i + 1] = c->row_ptr[i] + cnt[i];
  c->nnz = (uint64_t)c->row_ptr[n];
  c->col = (int32_t*)malloc((c->nnz ? c->nnz : 1) * sizeof(int32_t));
  c->val = (float*)malloc((c->nnz ? c->nnz : 1) * sizeof(float));
  memset(cnt, 0, (n + 1) * sizeof(int64_t));
  for (uint64_t i = 0; i < nkeys; ++i) {
    if (!keep[i]) continue;
    uint32_t u = (uint32_t)(keys[i] >> 32), v = (uint32_t)keys[i];
    int64_t p = c->row_ptr[u] + cnt[u]++;
    c->col[p] = (int32_t)v; c->val[p] = lw[i];
    if (!upper && u != v) {
      p = c->row_ptr[v] + cnt[v]++;
      c->col[p] = (int32_t)u; c->val[p] = lw[i];
    }
  }
  free(cnt); free(lw); free(keep);
  return c;
}

void orc_csr_info(const orc_csr* c, uint64_t* out_u, double* out_d) {
  out_u[0] = c->n; out_u[1] = c->nnz; out_u[2] = c->kept_upper;
  out_d[0] = c->M_exact; out_d[1] = c->factor;
}
void orc_csr_copy(const orc_csr* c, int64_t* row_ptr, int32_t* col, float* val) {
  memcpy(row_ptr, c->row_ptr, (c->n + 1) * sizeof(int64_t));
  memcpy(col, c->col, c->nnz * sizeof(int32_t));
  memcpy(val, c->val, c->nnz * sizeof(float));
}
void orc_csr_free(orc_csr* c) {
  if (!c) return;
  free(c->row_ptr); free(c->col); free(c->val); free(c);
}

/* ---------------------------------------------------------------------------
 * Sparse x dense products (the MKL mkl_sparse_?_mm calls of
 * mklredsvd.cc:235,271 and SpectralPropagation.h:101,116,153,168), row-major
 * X (n x k), Y = alpha*A*X + beta*Y, rows summed in CSR order.  Columns are
 * processed in blocks of ORC_KB so any k works (the reference's -rank 4096).
 * pinned by: scipy.sparse CSR products on the same inputs, alpha/beta forms and
 * k > ORC_KB included (tests/test_oracle.py::test_spmm_oracle_vs_scipy,
 * test_spmm_f32_oracle_vs_scipy).
 * ------------------------------------------------------------------------- */
#define ORC_KB 256

void orc_spmm_f64(uint64_t n, const int64_t* rp, const int32_t* col, const float* val, const double* X,
                  uint64_t k, double alpha, double beta, double* Y, int nthreads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (uint64_t i = 0; i < n; ++i) {
    double acc[ORC_KB];
    for (uint64_t c0 = 0; c0 < k; c0 += ORC_KB) {
      uint64_t kb = k - c0 < ORC_KB ? k - c0 : ORC_KB;
      for (uint64_t c = 0; c < kb; ++c) acc[c] = 0.0;
      for (int64_t j = rp[i]; j < rp[i + 1]; ++j) {
        double a = (double)val[j];
        const double* x = X + (uint64_t)col[j] * k + c0;
        for (uint64_t c = 0; c < kb; ++c) acc[c] += a * x[c];
      }
      double* y = Y + i * k + c0;
      for (uint64_t c = 0; c < kb; ++c) y[c] = alpha * acc[c] + (beta != 0.0 ? beta * y[c] : 0.0);
    }
  }
}

void orc_spmm_f32(uint64_t n, const int64_t* rp, const int32_t* col, const float* val, const float* X,
                  uint64_t k, float alpha, float beta, float* Y, int nthreads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (uint64_t i = 0; i < n; ++i) {
    float acc[ORC_KB];
    for (uint64_t c0 = 0; c0 < k; c0 += ORC_KB) {
      uint64_t kb = k - c0 < ORC_KB ? k - c0 : ORC_KB;
      for (uint64_t c = 0; c < kb; ++c) acc[c] = 0.0f;
      for (int64_t j = rp[i]; j < rp[i + 1]; ++j) {
        float a = val[j];
        const float* x = X + (uint64_t)col[j] * k + c0;
        for (uint64_t c = 0; c < kb; ++c) acc[c] += a * x[c];
      }
      float* y = Y + i * k + c0;
      for (uint64_t c = 0; c < kb; ++c) y[c] = alpha * acc[c] + (beta != 0.0f ? beta * y[c] : 0.0f);
    }
  }
}

/* SpectralPropagation.h:51-75 propagation operator applied implicitly:
 * (M x)_i = (1 - mu - 1/(d_i+1)) x_i - 1/(d_i+1) * sum_{j in N(i)} x_j,
 * with the two coefficients formed as the reference forms them (double
 * expression, stored as FP).  Graph CSR (off, adj) without self loops.
 * Y = alpha * M X + beta * Y.  Double precision throughout.
 * pinned by: the explicit scipy M of oracle.py:prop_matrix (itself pinned by
 * M 1 = -mu 1 and the closed-form spectra of a cycle and a star), on random
 * graphs with alpha/beta forms (tests/test_oracle.py::test_prop_apply_vs_explicit_matrix). */
void orc_prop_apply_f64(uint64_t n, const uint64_t* off, const uint32_t* adj, double mu, const double* X,
                        uint64_t k, double alpha, double beta, double* Y, int nthreads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (uint64_t i = 0; i < n; ++i) {
    double acc[ORC_KB];
    double deg = (double)(off[i + 1] - off[i] + 1);
    double dg = 1.0 - mu - 1.0 / deg, of = -1.0 / deg;
    for (uint64_t c0 = 0; c0 < k; c0 += ORC_KB) {
      uint64_t kb = k - c0 < ORC_KB ? k - c0 : ORC_KB;
      const double* xi = X + i * k + c0;
      for (uint64_t c = 0; c < kb; ++c) acc[c] = dg * xi[c];
      for (uint64_t j = off[i]; j < off[i + 1]; ++j) {
        const double* x = X + (uint64_t)adj[j] * k + c0;
        for (uint64_t c = 0; c < kb; ++c) acc[c] += of * x[c];
      }
      double* y = Y + i * k + c0;
      for (uint64_t c = 0; c < kb; ++c) y[c] = alpha * acc[c] + (beta != 0.0 ? beta * y[c] : 0.0);
    }
  }
}

/* Same operator in FP32 with coefficients rounded to float as the reference
 * stores them (SpectralPropagation.h:62,66,72).  CPU baseline kernel.
 * pinned by: the FP64 form above at FP32 tolerance, and its coefficients equal
 * to (float)(1.0 - mu - 1.0/deg) and (float)(-1.0/deg) exactly on unit vectors
 * (tests/test_oracle.py::test_prop_apply_vs_explicit_matrix). */
void orc_prop_apply_f32(uint64_t n, const uint64_t* off, const uint32_t* adj, float mu, const float* X,
                        uint64_t k, float alpha, float beta, float* Y, int nthreads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (uint64_t i = 0; i < n; ++i) {
    float acc[ORC_KB];
    float deg = (float)(off[i + 1] - off[i] + 1);
    float dg = (float)(1.0 - (double)mu - 1.0 / (double)deg), of = (float)(-1.0 / (double)deg);
    for (uint64_t c0 = 0; c0 < k; c0 += ORC_KB) {
      uint64_t kb = k - c0 < ORC_KB ? k - c0 : ORC_KB;
      const float* xi = X + i * k + c0;
      for (uint64_t c = 0; c < kb; ++c) acc[c] = dg * xi[c];
      for (uint64_t j = off[i]; j < off[i + 1]; ++j) {
        const float* x = X + (uint64_t)adj[j] * k + c0;
        for (uint64_t c = 0; c < kb; ++c) acc[c] += of * x[c];
      }
      float* y = Y + i * k + c0;
      for (uint64_t c = 0; c < kb; ++c) y[c] = alpha * acc[c] + (beta != 0.0f ? beta * y[c] : 0.0f);
    }
  }
}

/* SpectralPropagation.h:210-221 neighbour aggregation z_i = y_i + sum_{j in N(i)} y_j.
 * pinned by: scipy (A + I) @ Y on random graphs (tests/test_oracle.py::test_aggregate_vs_scipy). */
void orc_aggregate_f64(uint64_t n, const uint64_t* off, const uint32_t* adj, const double* Yin, uint64_t k,
                       double* Z, int nthreads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (uint64_t i = 0; i < n; ++i) {
    double* z = Z + i * k;
    const double* yi = Yin + i * k;
    for (uint64_t c = 0; c < k; ++c) z[c] = yi[c];
    for (uint64_t j = off[i]; j < off[i + 1]; ++j) {
      const double* y = Yin + (uint64_t)adj[j] * k;
      for (uint64_t c = 0; c < k; ++c) z[c] += y[c];
    }
  }
}

/* ---------------------------------------------------------------------------
 * Counter-based Gaussian sketch shared by the CPU and GPU paths (replaces MKL
 * VSL SFMT19937 + Box-Muller2 of mklredsvd.cc:141 / util.cc:21-38, whose
 * stream cannot be reproduced).  Philox4x32-10 (Salmon et al., SC'11) keyed by
 * (seed_lo, seed_hi), counter (pair index lo, pair index hi, stream, 0); the
 * first two 32-bit words give u1 = (w0 + 0.5) / 2^32 and u2 = (w1 + 0.5) / 2^32
 * and Box-Muller gives the pair (r cos t, r sin t) for elements 2p and 2p+1.
 * Computed in double, rounded to float, then to the nearest BF16-representable
 * value (the sketch is drawn at BF16 precision).
 * pinned by: Philox4x32-10 known-answer vectors from the Random123
 * distribution (counter 0 / key 0 and counter = key = all-ones); moments. */
void orc_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* nearest BF16-representable float, ties to even */
static float round_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return f;
}

void orc_gaussian(uint64_t seed, uint32_t stream, uint64_t count, float* out) {
  const double two_pi = 6.283185307179586476925286766559;
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma omp parallel for schedule(static)
  for (uint64_t p = 0; p < (count + 1) / 2; ++p) {
    uint32_t ctr[4] = {(uint32_t)p, (uint32_t)(p >> 32), stream, 0u}, w[4];
    orc_philox4x32(ctr, key, w);
    double u1 = ((double)w[0] + 0.5) * 2.3283064365386963e-10;
    double u2 = ((double)w[1] + 0.5) * 2.3283064365386963e-10;
    double r = sqrt(-2.0 * log(u1));
    double t = two_pi * u2;
    out[2 * p] = round_bf16((float)(r * cos(t)));
    if (2 * p + 1 < count) out[2 * p + 1] = round_bf16((float)(r * sin(t)));
  }
}

/* Very sparse random projection matrix (mklredsvd.cc:148-189), dense n x k
 * FP64: entry (i, c) is present when the first Philox word of counter
 * (i, c, stream 2) is below thresh = floor(2^32 / s), with value -sqrt(s) when
 * the second word is odd, else +sqrt(s); s (float) = s_param if >= 1, else
 * sqrt(n).  Same distribution as the reference's binomial row counts plus a
 * Fisher-Yates column choice. */
void orc_sparse_projection(uint64_t n, uint64_t k, uint64_t seed, double s_param, double* out) {
  const float s = s_param >= 1.0 ? (float)s_param : (float)sqrt((double)n);
  const double density = 1.0 / (double)s;
  const uint32_t thresh = density >= 1.0 ? 0xffffffffu : (uint32_t)(density * 4294967296.0);
  const double v = (double)sqrtf(s);
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t c = 0; c < k; ++c) {
      uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)c, 2u}, w[4];
      orc_philox4x32(ctr, key, w);
      out[i * k + c] = (w[0] < thresh) ? ((w[1] & 1u) ? -v : v) : 0.0;
    }
}

int orc_max_threads(void) { return omp_get_max_threads(); }
