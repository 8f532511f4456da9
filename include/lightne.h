/*
 * lightne.h -- C-ABI of the MI355X LightNE embedding pipeline.
 *
 * The reference exposes its three stages as in-process C++ templates; each
 * entry point below replaces one of them and keeps its argument meaning:
 *
 *   lightne_sparsify / lightne_sparsify_counts
 *       path_embed::generate_trunc_log_matrix_v2  (LightNE/PathEmbed.h:81-84)
 *   lightne_rsvd
 *       mkl_redsvd::MKLRedSVD<FP>::MKLRedSVD + run (LightNE/mklredsvd/mklredsvd.h:8-26,
 *                                                   mklredsvd.cc:54-81,93-403)
 *   lightne_u_sigma_root
 *       path_embed::compute_u_sigma_root          (LightNE/PathEmbed.h:371-408)
 *   lightne_propagate / lightne_chebyshev_expansion
 *       spectral_propagation::chebyshev_expansion (LightNE/SpectralPropagation.h:45-229)
 *   lightne_embed_dense
 *       spectral_propagation::get_embedding_dense (LightNE/SpectralPropagation.h:16-43)
 *   lightne_netsmf
 *       path_embed::NetSMF                        (LightNE/PathEmbed.h:494-563)
 *   lightne_graph_read_adj
 *       gbbs_io::read_unweighted_symmetric_graph  (ligra/graph_io.h:185-252)
 *   lightne_graph_decode_compressed
 *       gbbs_io::read_compressed_symmetric_graph  (ligra/graph_io.h:344-377)
 *       + bytepd_amortized::decode                (ligra/encodings/byte_pd_amortized.h:330-386)
 *
 * Conventions (differences from the reference are deliberate):
 *   - every call returns an int status (LIGHTNE_OK = 0); nothing calls exit()
 *     (the reference exits with status 0 on a full table, par_table.h:104-107,
 *     and on LAPACK failure, mklredsvd/util.cc:59-62);
 *   - array arguments are DEVICE pointers unless the name ends in _host;
 *   - outputs whose size is data dependent (the sampled table, the CSR) are
 *     returned as library-owned handles with _info/_copy/_free accessors;
 *   - one context per HIP stream; calls on one context are serialised by the
 *     caller; contexts are independent;
 *   - propagation does not modify its input (the reference's in-place
 *     `a -= conv`, SpectralPropagation.h:201-208, is not observable because the
 *     stage-2 embedding is saved before stage 3 runs, LightNE.cc:108-118).
 */
#ifndef LIGHTNE_H
#define LIGHTNE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define LIGHTNE_OK 0
#define LIGHTNE_ERR_BAD_ARG 1
#define LIGHTNE_ERR_OOM 2
#define LIGHTNE_ERR_TABLE_FULL 3
#define LIGHTNE_ERR_HIP 4
#define LIGHTNE_ERR_BAD_GRAPH 5
#define LIGHTNE_ERR_NUMERIC 6
#define LIGHTNE_ERR_IO 7
#define LIGHTNE_ERR_NO_DEVICE 8

/* panel precision */
#define LIGHTNE_FP32 0
#define LIGHTNE_BF16 1
#define LIGHTNE_FP64 2 /* randomized SVD only: FP64 panels, products and Grams (accuracy reference) */
#define LIGHTNE_BF16_POWER 3 /* as LIGHTNE_BF16, and the randomized SVD's power iterations also gather BF16 */
/* randomized SVD only: FP64 panels, products and Grams in the FP32 form's
 * schedule (thin SVD of B = A Q from its Gram) with two n x rank FP64 panels
 * instead of four; its outputs (U_out of lightne_rsvd, emb_out of
 * lightne_netsmf) may be host memory.  The accuracy reference where the
 * LIGHTNE_FP64 form does not fit. */
#define LIGHTNE_FP64_LEAN 4

typedef struct lightne_ctx lightne_ctx;
typedef struct lightne_table lightne_table;
typedef struct lightne_csr lightne_csr;

/* ---- context --------------------------------------------------------------
 * No reference counterpart (the reference is a single process-wide
 * computation); `stream` is a hipStream_t (NULL = the default stream). */
int lightne_ctx_create(int device, void* stream, lightne_ctx** out);
int lightne_ctx_destroy(lightne_ctx* ctx);
int lightne_ctx_synchronize(lightne_ctx* ctx);
/* Free the device memory the context keeps between calls (its workspace and
 * named scratch buffers; the next call allocates what it needs again).  No
 * reference counterpart: the reference frees its buffers at the end of each
 * stage. */
int lightne_ctx_release(lightne_ctx* ctx);
/* Text of the last error on this context ("" if none). */
const char* lightne_last_error(const lightne_ctx* ctx);
/* Text of the last error raised outside any context (create, graph IO). */
const char* lightne_global_error(void);
/* Library version string. */
const char* lightne_version(void);

/* ---- graph ----------------------------------------------------------------
 * Symmetric CSR in device memory; neighbours in input (file) order, which the
 * sampler's i-th-neighbour draw depends on (ligra/vertex.h:296-300).
 * offsets[n] == m == number of directed entries (GA.m). */
typedef struct {
  int64_t n;
  int64_t m;
  const int64_t* offsets; /* n+1 */
  const uint32_t* edges;  /* m   */
  int32_t ith_rule;       /* LIGHTNE_ITH_*: how a walk's draw i maps to a neighbour */
  int32_t reserved;       /* 0 */
} lightne_graph;

/* The walk (PathEmbed.h:37-49) takes get_ith_out_neighbor(u, i), whose result
 * depends on how the reference stored the graph:
 *   PLAIN  - neighbour i in input order (ligra/vertex.h:296-300): text and
 *            binary CSR inputs;
 *   BYTEPD - graphs read from the compressed format (-c): the byte-pd-amortized
 *            lookup (ligra/encodings/byte_pd_amortized.h:653-688) returns
 *            neighbour i - 1 unless i is a multiple of 64.
 * Only the sampler reads it; every other stage sees the same CSR. */
#define LIGHTNE_ITH_PLAIN 0
#define LIGHTNE_ITH_BYTEPD 1

/* ---- stage 1: sparsifier (PathEmbed.h:81-369) ------------------------------ */
typedef struct {
  uint64_t walk_len;        /* -walklen (T)                                  */
  uint64_t negative;        /* -negative (b)                                 */
  int32_t sample;           /* -sample: 0 = plain, 1 = eff.-resistance       */
  int32_t upper;            /* -upper: store the upper triangle only         */
  float sample_ratio;       /* -sample_ratio (kept as float, LightNE.cc:42)  */
  float mem_ratio;          /* -mem_ratio                                    */
  uint64_t tablesz;         /* -tablesz                                      */
  uint64_t seed;            /* pbbs::random state (reference: time(0))      */
  uint64_t max_table_bytes; /* cap on the hash table (0 = no cap)            */
  int32_t table_layout;     /* 0 auto (8-byte slots when n <= 2^26), 1 16-byte
                               slots, 2 8-byte slots                        */
  int32_t csr_parts;        /* 0 auto: the CSR is built by row ranges only
                               when one global sort does not fit; k >= 2:
                               at least k ranges                           */
  int32_t value_cache;      /* pipeline only: 0 auto (the count pass caches
                               each distinct key's CSR value where memory
                               allows), 1 always (the workspace grows for
                               it), -1 never; the CSR is the same either way */
  int32_t wide_offsets;     /* 0: the walks read each neighbour with its row
                               (entries built per call) when the graph has
                               fewer than 2^32 directed edges, else the
                               graph's 64-bit offsets; 1: neighbours and the
                               64-bit offsets always; 2: neighbours and 32-bit
                               row pointers where they fit (the same samples
                               every way)                                    */
} lightne_sparsify_params;

typedef struct {
  /* host scalars, PathEmbed.h:89,99,108-111 */
  double logn, M, one_over_lambda, fractional_bits;
  uint64_t exp_rv_rounded, table_size;
  /* sampler */
  uint64_t table_capacity; /* slots allocated on the device                */
  uint64_t attempts;       /* per-edge loop trips, sum of exp_rv           */
  uint64_t samples;        /* emitted samples                              */
  uint64_t walk_steps;     /* total random-walk steps                      */
  uint64_t distinct_keys;  /* U                                            */
  /* trunc-log / CSR, PathEmbed.h:175-311 */
  double M_exact, factor;
  uint64_t kept_upper;     /* keys with log(adj) > 0                       */
  uint64_t nnz;            /* CSR entries                                  */
  /* device time in seconds (HIP events on the context stream) */
  double t_sample, t_table, t_csr;
  uint32_t table_layout;   /* 1 = 16-byte slots, 2 = 8-byte slots          */
  uint32_t csr_parts;      /* row ranges the CSR was built in (1 = one sort) */
  uint32_t value_cache;    /* 1: the CSR values came from the count pass's
                              cache in the workspace, not a second read of
                              degrees; 2: from the cache in lightne_netsmf's
                              output buffer (which stage 2 writes later),
                              covering all keys or the first occupied slots
                              (the rest read the degrees); 0: no cache     */
  uint32_t sampler_runs;   /* 1, or 2 when the workspace grew after the count
                              pass and the sampler ran again                */
  double t_canon;          /* device seconds of the pass that sorts every
                              table cluster by key                         */
  uint32_t row_order;      /* order of each CSR row's entries: 1 fixed by
                              the key set (canonical table order), 2 by
                              column (a table cluster was too long to
                              reorder); the same on every run either way   */
  uint32_t walk_rows;      /* 1: the walks read each neighbour with its row
                              (one 16-byte entry per directed edge, for the
                              sampler's run), 0: neighbours and row pointers
                              apart (64-bit offsets, or no memory for the
                              entries); the same samples either way        */
} lightne_sparsify_stats;

/* Host-only evaluation of the sampling scalars (PathEmbed.h:89,99,108-117). */
int lightne_sample_scalars(int64_t n, int64_t m, const lightne_sparsify_params* p, lightne_sparsify_stats* out);

/* Sampler + table only: the key->(count, weight) multiset before trunc-log.
 * Keys are (min(x,z) << 32) + max(x,z) (PathEmbed.h:122-126); the handle
 * returns them sorted ascending. */
int lightne_sparsify_counts(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* p,
                            lightne_table** out, lightne_sparsify_stats* stats);
int lightne_table_info(const lightne_table* t, uint64_t* nkeys);
/* keys: u64[nkeys], counts: u32[nkeys], weights: f64[nkeys] (may be NULL);
 * destination pointers may be host or device. */
int lightne_table_copy(lightne_ctx* ctx, const lightne_table* t, uint64_t* keys, uint32_t* counts, double* weights);
int lightne_table_free(lightne_table* t);

/* Full stage 1: sampler, table, trunc-log, CSR (each row's entries in an
 * order the key set alone fixes: the same on every run).
 * (generate_trunc_log_matrix_v2 with rows_start != NULL, PathEmbed.h:224-316.) */
int lightne_sparsify(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* p,
                     lightne_csr** out, lightne_sparsify_stats* stats);

/* CSR handle: int64 row_ptr[n+1], int32 col[nnz], float val[nnz] on the device. */
int lightne_csr_create(lightne_ctx* ctx, int64_t n, int64_t nnz, const int64_t* row_ptr, const int32_t* col,
                       const float* val, lightne_csr** out);
int lightne_csr_info(const lightne_csr* c, int64_t* n, int64_t* nnz);
int lightne_csr_device_ptrs(const lightne_csr* c, const int64_t** row_ptr, const int32_t** col, const float** val);
int lightne_csr_copy(lightne_ctx* ctx, const lightne_csr* c, int64_t* row_ptr, int32_t* col, float* val);
/* the first `count` entries of col / val (count <= nnz; e.g. row_ptr[r] for
 * the entries of rows [0, r)), host or device destinations */
int lightne_csr_copy_prefix(lightne_ctx* ctx, const lightne_csr* c, int64_t count, int32_t* col, float* val);
int lightne_csr_free(lightne_csr* c);

/* ---- stage 2: randomized SVD (mklredsvd.cc:93-403) -------------------------- */
typedef struct {
  int64_t rank;             /* -rank: sketch width k (d + oversampling), any k
                               in [1, n] below 32768 in the FP32 and BF16
                               forms; the FP64 forms take multiples of 16 up
                               to 256, sparse_project any k up to 256        */
  int32_t power_iters;      /* q >= 1; 1 = the reference schedule             */
  int32_t precision;        /* LIGHTNE_FP32 / BF16 / BF16_POWER / FP64 / FP64_LEAN */
  int32_t random_project_only; /* -random_project_only (mklredsvd.cc:254-258) */
  int32_t upper;            /* A holds only the upper triangle (symmetric)    */
  uint64_t seed;            /* sketch seed (reference: VSL seed 777)          */
  int32_t sparse_project;   /* -sparse_project: very sparse random projection
                               instead of the Gaussian sketch (mklredsvd.cc:148-226) */
  float sparse_project_s;   /* -sparse_project_s: 1/s density; < 1 selects sqrt(n) */
} lightne_rsvd_params;

typedef struct {
  double t_sketch, t_spmm, t_qr, t_gemm, t_small_svd, t_total; /* seconds */
  int32_t spmm_calls, qr_calls;
  double t_spmm_sketch;     /* the first product Y = A Omega (part of t_spmm) */
} lightne_rsvd_stats;

/* U_out: n x rank row-major (device).  S_host: rank singular values,
 * descending (NULL allowed when random_project_only, which returns Y = A*Omega
 * in U_out as the reference returns it in matU with S = NULL). */
int lightne_rsvd(lightne_ctx* ctx, const lightne_csr* A, const lightne_rsvd_params* p, float* U_out,
                 float* S_host, lightne_rsvd_stats* stats);

/* emb[:, j] = sqrt(S_j) * U[:, j] for j < dim (U has row stride rank), then the
 * row L2 normalisation (norm > 1e-5) when normalize != 0.  S_host NULL means
 * S = 1 (PathEmbed.h:371-408). */
int lightne_u_sigma_root(lightne_ctx* ctx, const float* U, const float* S_host, int64_t n, int64_t rank,
                         int64_t dim, int32_t normalize, float* emb_out);

/* ---- stage 3: spectral propagation (SpectralPropagation.h:45-229) ----------- */
typedef struct {
  int64_t order;            /* -order K                                      */
  float mu;                 /* reference hard-codes 0.2 (LightNE.cc:114)     */
  float theta;              /* reference hard-codes 0.5 (LightNE.cc:114)     */
  int32_t precision;        /* LIGHTNE_FP32 / LIGHTNE_BF16 iterates          */
} lightne_prop_params;

typedef struct {
  double t_cheb, t_aggregate, t_svd, t_total;
  int32_t spmm_calls;
} lightne_prop_stats;

/* y = a - sum_{i<K} c_i T_i a and z = (A + I) y (SpectralPropagation.h:94-221).
 * a: n x dim (device, not modified), dim in [1, 1024].  y_out / z_out may be
 * NULL. */
int lightne_propagate(lightne_ctx* ctx, const lightne_graph* g, const float* a, int64_t dim,
                      const lightne_prop_params* p, float* y_out, float* z_out, lightne_prop_stats* stats);

/* get_embedding_dense: thin SVD of mat (n x rank), emb = rownorm(U[:, :dim] sqrt(S)).
 * Any 1 <= dim <= rank <= n with rank < 32768.  S_host (rank values,
 * descending) may be NULL. */
int lightne_embed_dense(lightne_ctx* ctx, const float* mat, int64_t n, int64_t rank, int64_t dim,
                        float* emb_out, float* S_host);

/* chebyshev_expansion(a, GA, rank=dim, dim, order, mu, s): propagation, then
 * get_embedding_dense.  emb_out: n x dim; it may be a itself (a is read in
 * full before emb_out is written), which saves one n x dim buffer. */
int lightne_chebyshev_expansion(lightne_ctx* ctx, const lightne_graph* g, const float* a, int64_t dim,
                                const lightne_prop_params* p, float* emb_out, lightne_prop_stats* stats);

/* Adjoint of lightne_propagate's y (z = (A+I)y when with_aggregate != 0):
 * given grad_out = dL/dy (or dL/dz), returns dL/da (n x dim, device) and
 * dL/dmu, dL/dtheta (host doubles).  The forward Chebyshev iterates are
 * recomputed from checkpoints taken every `checkpoint_every` steps.
 * No reference counterpart: the reference has no gradients. */
int lightne_propagate_backward(lightne_ctx* ctx, const lightne_graph* g, const float* a, int64_t dim,
                               const lightne_prop_params* p, int32_t with_aggregate, const float* grad_out,
                               float* grad_a, double* grad_mu_host, double* grad_theta_host,
                               int32_t checkpoint_every);

/* ---- whole NetSMF stage 1 + 2 (PathEmbed.h:494-563) -------------------------- */
typedef struct {
  double t_sparsify, t_svd, t_total;
  lightne_sparsify_stats sparsify;
  lightne_rsvd_stats rsvd;
} lightne_netsmf_stats;

int lightne_netsmf(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* sp,
                   const lightne_rsvd_params* rp, int64_t dim, int32_t normalize, float* emb_out,
                   float* S_host, lightne_netsmf_stats* stats);

/* ---- the ne_zhang_et_al method (PathEmbed.h:410-463, LightNE.cc:106) -------
 * Matrix on the graph's own pattern: A_uv = log(1/d_u) - log(neg_v / sum neg),
 * neg_j = (sum_{i in N(j)} 1/d_i)^0.75, then the randomized SVD (rp) and
 * emb = rownorm(U[:, :dim] sqrt(S)) (normalisation always on, as :459).
 * n must be below 2^31 (the neighbour ids are the CSR columns). */
int lightne_ne_zhang(lightne_ctx* ctx, const lightne_graph* g, const lightne_rsvd_params* rp, int64_t dim,
                     float* emb_out, float* S_host, lightne_rsvd_stats* stats);
/* the matrix alone (PathEmbed.h:412-444), as a CSR handle: row_ptr = the graph's
 * offsets, columns in adjacency order */
int lightne_ne_zhang_matrix(lightne_ctx* ctx, const lightne_graph* g, lightne_csr** out);

/* ---- dense panel kernels (building blocks of stages 2 and 3) ---------------
 * No reference counterpart as separate calls: they are the pieces that replace
 * the reference's cblas_?gemm (mklredsvd.cc:297,330,382) and the
 * geqrf/orgqr orthonormalisation (mklredsvd/util.cc:41-79); exported so each
 * can be checked on its own.  Panels are n x k row-major FP32 on the device,
 * k and k2 multiples of 16 up to 256. */
/* G_host (k1 x k2, FP64) = X^T Y; symmetric != 0 requires X == Y, k1 == k2 */
int lightne_dense_gram(lightne_ctx* ctx, const float* X, const float* Y, int64_t n, int64_t k1, int64_t k2,
                       int32_t symmetric, double* G_host);
/* As lightne_dense_gram, but allocating no device memory: no counterpart in
 * the reference; for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Its
 * temporaries (tile-pair list, per-block partial sums, the summed Gram) live in
 * a 48 MiB device array of the library, one per device, held until the process
 * ends; k1 and k2 must be multiples of 16 in [16, 256] (partial sums of up to
 * 94 blocks at 256 x 256; widths up to 144 use the same blocks, and so give the
 * same bits, as lightne_dense_gram).  Calls that overlap on one device share
 * the array and race: the caller serialises them. */
int lightne_dense_gram_noalloc(lightne_ctx* ctx, const float* X, const float* Y, int64_t n, int64_t k1, int64_t k2,
                               int32_t symmetric, double* G_host);
/* Out (n x k2) = X (n x k1) W_host (k1 x k2, used as FP32); optional row L2
 * normalisation; Out may alias X when k1 == k2 */
int lightne_dense_panel_gemm(lightne_ctx* ctx, const float* X, int64_t n, int64_t k1, const double* W_host,
                             int64_t k2, float* Out, int32_t normalize_rows);
/* CholeskyQR2 with FP64 Gram/Cholesky: Y <- Q in place, R_host (k x k) with
 * Y_in = Q R; passes_out (may be NULL) receives the number of passes */
int lightne_dense_cholqr(lightne_ctx* ctx, float* Y, int64_t n, int64_t k, double* R_host, int32_t* passes_out);

/* ---- compressed graph input (-c) -------------------------------------------
 * Decodes a file in Ligra's compressed format (byte-pd-amortized encoding,
 * ligra/encodings/byte_pd_amortized.h; layout of
 * gbbs_io::read_compressed_symmetric_graph, ligra/graph_io.h:344-377) into the
 * symmetric CSR.  `image` is the whole file copied to device memory (8-byte
 * aligned), as the reference reads or maps it; n and m are the file's first
 * two int64 words (checked).  offsets_out (n+1) and edges_out (m) are device
 * buffers; neighbours come out in decode order.  A malformed file returns
 * LIGHTNE_ERR_BAD_GRAPH with the reason; nothing is read outside the image.
 * Sample from the result with ith_rule = LIGHTNE_ITH_BYTEPD to reproduce the
 * reference's walks on the compressed input. */
int lightne_graph_decode_compressed(lightne_ctx* ctx, const void* image, uint64_t image_bytes, int64_t n, int64_t m,
                                    int64_t* offsets_out, uint32_t* edges_out);

/* ---- host graph utilities (no device work) ---------------------------------- */
/* Ligra AdjacencyGraph text reader/writer (ligra/graph_io.h:185-231).
 * *offsets_host has n+1 entries; free both with lightne_host_free. */
int lightne_graph_read_adj(const char* path, int64_t* n, int64_t* m, int64_t** offsets_host,
                           uint32_t** edges_host);
int lightne_graph_write_adj(const char* path, int64_t n, const int64_t* offsets_host, const uint32_t* edges_host);
/* Binary CSR cache (no reference counterpart: a load path that skips the
 * AdjacencyGraph text parse of ligra/graph_io.h:185-231).  Layout: 8 bytes
 * "LNECSR01", int64 n, int64 m, int64 offsets[n+1], uint32 edges[m]. */
int lightne_graph_read_bin(const char* path, int64_t* n, int64_t* m, int64_t** offsets_host, uint32_t** edges_host);
int lightne_graph_write_bin(const char* path, int64_t n, const int64_t* offsets_host, const uint32_t* edges_host);
/* Seeded synthetic graphs: undirected, no self loops, no duplicates, each
 * neighbour list sorted.  Chung-Lu with power-law expected degrees
 * w_i ~ (i + i0)^(-1/(beta-1)) scaled to mean `mean_deg` and capped at
 * `max_deg`; R-MAT with quadrant probabilities (a, b, c, 1-a-b-c). */
int lightne_gen_chung_lu(int64_t n, double beta, double mean_deg, double max_deg, uint64_t seed, int64_t* m,
                         int64_t** offsets_host, uint32_t** edges_host);
int lightne_gen_rmat(int32_t scale, double edge_factor, double a, double b, double c, uint64_t seed,
                     int64_t* n, int64_t* m, int64_t** offsets_host, uint32_t** edges_host);
void lightne_host_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* LIGHTNE_H */
