// Stage 2: randomized SVD of the sparsifier (LightNE/mklredsvd/mklredsvd.cc:93-403)
// and the U sqrt(Sigma) embedding (LightNE/PathEmbed.h:371-408), plus the CSR
// handle API, the NetSMF driver (PathEmbed.h:494-563) and ne_zhang_et_al's
// stage 2 (PathEmbed.h:445-460).
//
// Reference schedule (q = power_iters; q = 1 is the reference's):
//   Omega ~ N(0,1) n x k            (mklredsvd.cc:229-231, own counter-based RNG;
//                                    or the very sparse projection, :148-226)
//   Y = A Omega                      (:233-250)
//   Q = orth(Y)                      (:262)   CholeskyQR2, FP64 Gram + Cholesky
//   repeat q-1 times: Q = orth(A Q)
//   B = A Q                          (:268-287)
//   P ~ N(0,1) k x k; Z = orth(B P); C = Z^T B; C = U_C S V^T; matU = Z U_C
//                                    (:289-400)
// The FP32 and BF16 forms replace the last line by the thin SVD of B from its
// FP64 Gram, B^T B = V S^2 V^T, U = B V S^-1 (range(Z) = range(B) for any
// invertible P, so U and S are the same); the FP64 form keeps it as written.
// The sketch Omega is drawn BF16-representable and stored as BF16 in every
// non-FP64 form: Y = A Omega gathered from it with FP32 accumulation is the
// product with the FP32 values, bit for bit, at half the gathered bytes.  The
// power iterations gather the FP32 Q in the FP32 and BF16 forms (which are
// therefore the same in stage 2): a BF16 rounding of Q (2^-9) is amplified by
// 1/(sigma_d/sigma_{d+1} - 1) in the top-d subspace, ~5000 on an R-MAT
// scale-24 sparsifier.  The BF16_POWER form also gathers the power-iteration
// panels as BF16 (all but the last product B = A Q, the FP32 refinement pass).
// Grams are FP64 in every form.
#include "lightne_internal.h"
#include "spmm_epi.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace {

int csr_plan(lightne_ctx* ctx, lightne_csr* A) {
  if (A->plan) return LIGHTNE_OK;
  auto* P = new SpmmPlan();
  int rc = spmm_plan_build(ctx, A->row_ptr, A->n, *P);
  if (rc != LIGHTNE_OK) {
    spmm_plan_free(*P);
    delete P;
    return rc;
  }
  A->plan = P;
  return LIGHTNE_OK;
}

// Y(f32) = A X, X in FP32 or BF16; `partial` (nparts x k floats, or NULL for
// a context buffer) holds the partial sums of rows split over several chunks
int csr_mm(lightne_ctx* ctx, lightne_csr* A, const void* X, int xt, int64_t k, float* Y, float* partial = nullptr) {
  LN_CHECK(csr_plan(ctx, A));
  SpmmArgs a;
  a.partial = partial;
  a.n = A->n;
  a.k = k;
  a.rp = A->row_ptr;
  a.col = A->col;
  a.val = A->val;
  a.impl_mode = IMPL_NONE;
  a.X = X;
  a.out = Y;
  a.alpha = 1.f;
  return spmm_run(ctx, *A->plan, a, xt, EPI_STORE_F32, false);
}

// Y(double) = A X(double), the FP64 form
int csr_mm64(lightne_ctx* ctx, lightne_csr* A, const double* X, int64_t k, double* Y) {
  LN_CHECK(csr_plan(ctx, A));
  SpmmArgs a;
  a.n = A->n;
  a.k = k;
  a.rp = A->row_ptr;
  a.col = A->col;
  a.val = A->val;
  a.X = X;
  a.out = Y;
  return spmm_run64(ctx, *A->plan, a);
}

// Y <- Y W for a square host W (in place for FP32 panels; FP64 panels go
// through tmp)
int apply_small(lightne_ctx* ctx, float* Y, float*, int64_t n, int64_t k, const double* W) {
  return panel_times_small(ctx, Y, n, k, k, W, k, Y, 0, nullptr, 1);  // W = R^-1, upper triangular
}
int apply_small(lightne_ctx* ctx, double* Y, double* tmp, int64_t n, int64_t k, const double* W) {
  LN_CHECK(panel_times_small64(ctx, Y, n, k, W, k, tmp));
  LN_HIP(ctx, hipMemcpyAsync(Y, tmp, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  return LIGHTNE_OK;
}

}  // namespace

template <typename T>
int cholqr2_t(lightne_ctx* ctx, T* Y, T* tmp, int64_t n, int64_t k, double* R_out_host, int* passes_out) {
  const int kk = (int)k;
  std::vector<double> G(k * k), Gs(k * k), Rs(k * k), R(k * k), Rinv(k * k), Rtot(k * k, 0.0), Rnext(k * k);
  for (int i = 0; i < kk; ++i) Rtot[i * k + i] = 1.0;
  int passes = 0;
  for (int pass = 0; pass < 4; ++pass) {
    LN_CHECK(gram_f64(ctx, Y, Y, n, k, k, 1, G.data()));
    // deviation from orthonormality of the current panel
    double dev = 0.0;
    for (int i = 0; i < kk; ++i)
      for (int j = 0; j < kk; ++j) dev = std::max(dev, std::fabs(G[i * k + j] - (i == j ? 1.0 : 0.0)));
    if (pass >= 2 && dev < 1e-5) break;
    // diagonal equilibration: Gs = D G D, D = diag(G)^-1/2
    std::vector<double> d(k);
    for (int i = 0; i < kk; ++i) d[i] = G[i * k + i] > 0.0 ? 1.0 / std::sqrt(G[i * k + i]) : 1.0;
    for (int i = 0; i < kk; ++i)
      for (int j = 0; j < kk; ++j) Gs[i * k + j] = d[i] * G[i * k + j] * d[j];
    if (!chol_upper(Gs.data(), kk, Rs.data())) {
      // shifted CholeskyQR (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa 2020)
      double fro = 0.0;
      for (double v : Gs) fro += v * v;
      fro = std::sqrt(fro);
      const double u = std::ldexp(1.0, -53);
      double shift = 11.0 * ((double)n * k + (double)k * (k + 1)) * u * fro;
      bool ok = false;
      for (int attempt = 0; attempt < 8 && !ok; ++attempt, shift *= 100.0) {
        std::vector<double> Gt(Gs);
        for (int i = 0; i < kk; ++i) Gt[i * k + i] += shift;
        ok = chol_upper(Gt.data(), kk, Rs.data());
      }
      if (!ok) return ctx_fail(ctx, LIGHTNE_ERR_NUMERIC, "CholeskyQR: Gram matrix is not positive definite");
    }
    // R = Rs D^-1
    for (int i = 0; i < kk; ++i)
      for (int j = 0; j < kk; ++j) R[i * k + j] = Rs[i * k + j] / d[j];
    tri_upper_inverse(R.data(), kk, Rinv.data());
    LN_CHECK(apply_small(ctx, Y, tmp, n, k, Rinv.data()));
    if (R_out_host) {  // accumulated R only when the caller wants it
      matmul_host(R.data(), Rtot.data(), kk, kk, kk, Rnext.data());
      Rtot.swap(Rnext);
    }
    ++passes;
    if (pass >= 1 && dev < 1e-1) break;  // input was already close to orthonormal
  }
  if (R_out_host) std::copy(Rtot.begin(), Rtot.end(), R_out_host);
  if (passes_out) *passes_out = passes;
  return LIGHTNE_OK;
}

int cholqr2(lightne_ctx* ctx, float* Y, int64_t n, int64_t k, double* R_out_host, int* passes_out) {
  return cholqr2_t<float>(ctx, Y, nullptr, n, k, R_out_host, passes_out);
}

extern "C" {

int lightne_csr_create(lightne_ctx* ctx, int64_t n, int64_t nnz, const int64_t* row_ptr, const int32_t* col,
                       const float* val, lightne_csr** out) {
  if (!ctx || !out || n <= 0 || nnz < 0 || !row_ptr || (nnz > 0 && (!col || !val)))
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_csr_create: bad argument");
  CallScope scope(ctx);
  *out = nullptr;
  LN_HIP(ctx, hipSetDevice(ctx->device));
  DevMem a, b, c;
  LN_CHECK(dev_alloc(ctx, a, (n + 1) * 8, "csr row_ptr"));
  LN_CHECK(dev_alloc(ctx, b, std::max<int64_t>(nnz, 1) * 4, "csr col"));
  LN_CHECK(dev_alloc(ctx, c, std::max<int64_t>(nnz, 1) * 4, "csr val"));
  LN_HIP(ctx, hipMemcpyAsync(a.p, row_ptr, (n + 1) * 8, hipMemcpyDefault, ctx->stream));
  if (nnz) {
    LN_HIP(ctx, hipMemcpyAsync(b.p, col, nnz * 4, hipMemcpyDefault, ctx->stream));
    LN_HIP(ctx, hipMemcpyAsync(c.p, val, nnz * 4, hipMemcpyDefault, ctx->stream));
  }
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  auto* h = new lightne_csr();
  h->n = n;
  h->nnz = nnz;
  h->row_ptr = (int64_t*)a.release();
  h->col = (int32_t*)b.release();
  h->val = (float*)c.release();
  *out = h;
  return LIGHTNE_OK;
}

int lightne_csr_info(const lightne_csr* c, int64_t* n, int64_t* nnz) {
  if (!c) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_csr_info: NULL handle");
  if (n) *n = c->n;
  if (nnz) *nnz = c->nnz;
  return LIGHTNE_OK;
}

int lightne_csr_device_ptrs(const lightne_csr* c, const int64_t** row_ptr, const int32_t** col, const float** val) {
  if (!c) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_csr_device_ptrs: NULL handle");
  if (row_ptr) *row_ptr = c->row_ptr;
  if (col) *col = c->col;
  if (val) *val = c->val;
  return LIGHTNE_OK;
}

int lightne_csr_copy(lightne_ctx* ctx, const lightne_csr* c, int64_t* row_ptr, int32_t* col, float* val) {
  if (!ctx || !c) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_csr_copy: NULL argument");
  if (row_ptr) LN_HIP(ctx, hipMemcpyAsync(row_ptr, c->row_ptr, (c->n + 1) * 8, hipMemcpyDefault, ctx->stream));
  if (c->nnz) {
    if (col) LN_HIP(ctx, hipMemcpyAsync(col, c->col, c->nnz * 4, hipMemcpyDefault, ctx->stream));
    if (val) LN_HIP(ctx, hipMemcpyAsync(val, c->val, c->nnz * 4, hipMemcpyDefault, ctx->stream));
  }
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LIGHTNE_OK;
}

int lightne_csr_copy_prefix(lightne_ctx* ctx, const lightne_csr* c, int64_t count, int32_t* col, float* val) {
  if (!ctx || !c) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_csr_copy_prefix: NULL argument");
  if (count < 0 || count > c->nnz)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_csr_copy_prefix: count %lld outside [0, %lld]",
                    (long long)count, (long long)c->nnz);
  if (count) {
    if (col) LN_HIP(ctx, hipMemcpyAsync(col, c->col, count * 4, hipMemcpyDefault, ctx->stream));
    if (val) LN_HIP(ctx, hipMemcpyAsync(val, c->val, count * 4, hipMemcpyDefault, ctx->stream));
  }
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LIGHTNE_OK;
}

int lightne_csr_free(lightne_csr* c) {
  if (!c) return LIGHTNE_OK;
  if (c->plan) {
    spmm_plan_free(*c->plan);
    delete c->plan;
  }
  if (c->owns) {
    if (c->row_ptr) (void)hipFree(c->row_ptr);
    if (c->col) (void)hipFree(c->col);
    if (c->val) (void)hipFree(c->val);
  }
  delete c;
  return LIGHTNE_OK;
}

// FP64 form of the schedule above: every panel double, the same sketch (its
// values are BF16-representable, hence exact in double), FP64 products and
// Grams; U is returned rounded to FP32.  The accuracy reference for the FP32
// and BF16 forms at sizes the host oracle cannot reach.
static int rsvd64(lightne_ctx* ctx, lightne_csr* A, const lightne_rsvd_params* p, int q, float* U_out, float* S_host,
                  lightne_rsvd_stats* stats) {
  const int64_t n = A->n, k = p->rank;
  hipStream_t s = ctx->stream;
  const size_t pd = (size_t)n * k * sizeof(double);
  double *P0 = nullptr, *P1 = nullptr, *P2 = nullptr, *T = nullptr;  // Omega then Z; Y / Q; B; scratch
  {
    WsCarve wc;
    wc.add((void**)&P0, pd);
    wc.add((void**)&P1, pd);
    wc.add((void**)&P2, pd);
    wc.add((void**)&T, pd);
    LN_CHECK(wc.commit(ctx));
  }
  lightne_rsvd_stats st{};
  hipEvent_t* e = ctx->ev;
  auto lap = [&](hipEvent_t a, hipEvent_t b) { return ev_seconds(a, b); };
  double t_spmm = 0, t_qr = 0;
  LN_HIP(ctx, hipEventRecord(e[0], s));
  if (!p->sparse_project) {
    LN_CHECK(gaussian_fill(ctx, p->seed, 0u, n * k, (float*)T));
    LN_CHECK(f32_to_f64(ctx, (const float*)T, P0, n * k));
  }
  LN_HIP(ctx, hipEventRecord(e[1], s));
  LN_CHECK(csr_plan(ctx, A));
  LN_HIP(ctx, hipEventRecord(e[2], s));
  if (p->sparse_project)
    LN_CHECK(sparse_project_sketch(ctx, A->row_ptr, A->col, A->val, n, k, p->seed, p->sparse_project_s, P1, true));
  else
    LN_CHECK(csr_mm64(ctx, A, P0, k, P1));
  LN_HIP(ctx, hipEventRecord(e[3], s));
  LN_HIP(ctx, hipEventSynchronize(e[3]));
  st.t_sketch = lap(e[0], e[1]);
  t_spmm += lap(e[2], e[3]);
  st.t_spmm_sketch = lap(e[2], e[3]);
  st.spmm_calls = 1;
  if (p->random_project_only) {
    LN_CHECK(f64_to_f32(ctx, P1, U_out, n * k));
    LN_HIP(ctx, hipStreamSynchronize(s));
    st.t_spmm = t_spmm;
    st.t_total = lap(e[0], e[3]);
    if (stats) *stats = st;
    return LIGHTNE_OK;
  }
  int passes = 0;
  for (int it = 0; it < q; ++it) {
    LN_HIP(ctx, hipEventRecord(e[4], s));
    LN_CHECK(cholqr2_t<double>(ctx, P1, T, n, k, nullptr, &passes));
    st.qr_calls++;
    LN_HIP(ctx, hipEventRecord(e[5], s));
    LN_CHECK(csr_mm64(ctx, A, P1, k, P2));
    st.spmm_calls++;
    LN_HIP(ctx, hipEventRecord(e[6], s));
    LN_HIP(ctx, hipEventSynchronize(e[6]));
    t_qr += lap(e[4], e[5]);
    t_spmm += lap(e[5], e[6]);
    if (it + 1 < q) std::swap(P1, P2);
  }
  std::vector<float> Pf(k * k);
  gaussian_fill_host(p->seed, 1u, k * k, Pf.data());
  std::vector<double> Pd(Pf.begin(), Pf.end());
  LN_CHECK(panel_times_small64(ctx, P2, n, k, Pd.data(), k, P0));  // Z = B P
  LN_CHECK(cholqr2_t<double>(ctx, P0, T, n, k, nullptr, &passes));
  st.qr_calls++;
  std::vector<double> C(k * k), Uc(k * k), Sc(k);
  LN_CHECK(gram_f64(ctx, P0, P2, n, k, k, 0, C.data()));  // C = Z^T B
  auto t0 = std::chrono::steady_clock::now();
  if (!svd_left_via_eig(C.data(), (int)k, Uc.data(), Sc.data()))
    return ctx_fail(ctx, LIGHTNE_ERR_NUMERIC, "lightne_rsvd: small SVD did not converge");
  st.t_small_svd = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  LN_CHECK(panel_times_small64(ctx, P0, n, k, Uc.data(), k, T));  // matU = Z U_C
  LN_CHECK(f64_to_f32(ctx, T, U_out, n * k));
  LN_HIP(ctx, hipEventRecord(e[12], s));
  LN_HIP(ctx, hipEventSynchronize(e[12]));
  if (S_host)
    for (int64_t j = 0; j < k; ++j) S_host[j] = (float)Sc[j];
  st.t_spmm = t_spmm;
  st.t_qr = t_qr;
  st.t_total = lap(e[0], e[12]);
  if (stats) *stats = st;
  return LIGHTNE_OK;
}

// Lean FP64 form: the FP32 form's schedule (thin SVD of B = A Q from its Gram)
// with FP64 panels, products and Grams in two n x k FP64 panels -- half the
// explicit form's four, so that it fits beside a sparsifier of 1.5e10 nonzeros
// at n = 2^26.  The triangular applies Y R^-1 run over row blocks through a
// small staging buffer instead of a third panel.  On return `out` (host or
// device, row-major n x cols) holds B V[:, :cols] S^-pw rounded to FP32, row
// normalised when asked, or Y = A Omega with random_project_only.
static int rsvd64_lean(lightne_ctx* ctx, lightne_csr* A, const lightne_rsvd_params* p, int q, int64_t cols,
                       double pw, int normalize, float* out, std::vector<double>& S, lightne_rsvd_stats* stp) {
  const int64_t n = A->n, k = p->rank;
  lightne_rsvd_stats& st = *stp;
  st = lightne_rsvd_stats{};
  hipStream_t s = ctx->stream;
  const size_t pd = (size_t)n * k * sizeof(double);
  const int64_t blk_rows = std::min<int64_t>(n, (int64_t)1 << 20);
  // its own work split with 8192-nnz chunks: FP64 partial sums of the rows
  // split over several chunks take nparts x k x 8 bytes (29.7 GB with the
  // default chunks on a 1.5e10-nonzero sparsifier, 2 GB with these)
  struct Plan64 {
    SpmmPlan P;
    ~Plan64() { spmm_plan_free(P); }
  } plan;
  LN_CHECK(spmm_plan_build(ctx, A->row_ptr, n, plan.P, 8192));
  auto mm64 = [&](const double* X, double* Y) -> int {
    SpmmArgs a;
    a.n = n;
    a.k = k;
    a.rp = A->row_ptr;
    a.col = A->col;
    a.val = A->val;
    a.X = X;
    a.out = Y;
    return spmm_run64(ctx, plan.P, a);
  };
  double *P0 = nullptr, *P1 = nullptr, *T = nullptr;  // Omega, then B; Y / Q; row-block staging
  {
    WsCarve wc;
    wc.add((void**)&P0, pd);
    wc.add((void**)&P1, pd);
    wc.add((void**)&T, (size_t)blk_rows * k * sizeof(double));
    LN_CHECK(wc.commit(ctx));
  }
  hipEvent_t* e = ctx->ev;
  auto lap = [&](hipEvent_t a, hipEvent_t b) { return ev_seconds(a, b); };
  // Y <- Y W for a square host W, one row block at a time through T
  auto apply_rows = [&](double* Y, const double* W) -> int {
    for (int64_t r0 = 0; r0 < n; r0 += blk_rows) {
      const int64_t rows = std::min(blk_rows, n - r0);
      LN_CHECK(panel_times_small64(ctx, Y + r0 * k, rows, k, W, k, T));
      LN_HIP(ctx, hipMemcpyAsync(Y + r0 * k, T, (size_t)rows * k * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    return LIGHTNE_OK;
  };
  double t_spmm = 0, t_qr = 0;
  LN_HIP(ctx, hipEventRecord(e[0], s));
  if (!p->sparse_project) {
    // the sketch's FP32 values (BF16-representable, exact in double) drawn
    // into the second half of P1's bytes, widened into P0
    float* tmp32 = reinterpret_cast<float*>(P1) + (size_t)n * k;
    LN_CHECK(gaussian_fill(ctx, p->seed, 0u, n * k, tmp32));
    LN_CHECK(f32_to_f64(ctx, tmp32, P0, n * k));
  }
  LN_HIP(ctx, hipEventRecord(e[1], s));
  LN_HIP(ctx, hipEventRecord(e[2], s));
  if (p->sparse_project)
    LN_CHECK(sparse_project_sketch(ctx, A->row_ptr, A->col, A->val, n, k, p->seed, p->sparse_project_s, P1, true));
  else
    LN_CHECK(mm64(P0, P1));
  LN_HIP(ctx, hipEventRecord(e[3], s));
  LN_HIP(ctx, hipEventSynchronize(e[3]));
  st.t_sketch = lap(e[0], e[1]);
  t_spmm += lap(e[2], e[3]);
  st.t_spmm_sketch = lap(e[2], e[3]);
  st.spmm_calls = 1;
  float* stage = reinterpret_cast<float*>(P0);  // FP32 result staging (P0's bytes hold 2 n k floats)
  if (p->random_project_only) {
    LN_CHECK(f64_to_f32(ctx, P1, stage, n * k));
    LN_HIP(ctx, hipMemcpyAsync(out, stage, (size_t)n * k * sizeof(float), hipMemcpyDefault, s));
    LN_HIP(ctx, hipStreamSynchronize(s));
    S.clear();
    st.t_spmm = t_spmm;
    st.t_total = lap(e[0], e[3]);
    return LIGHTNE_OK;
  }
  std::vector<double> G(k * k), Gs(k * k), Rs(k * k), R(k * k), Rinv(k * k), w(k), V(k * k, 0.0);
  for (int it = 0; it < q; ++it) {
    LN_HIP(ctx, hipEventRecord(e[4], s));
    // CholeskyQR2 in FP64 (the same passes as cholqr2_t)
    for (int pass = 0; pass < 4; ++pass) {
      LN_CHECK(gram_f64(ctx, P1, P1, n, k, k, 1, G.data()));
      double dev = 0.0;
      for (int64_t i = 0; i < k; ++i)
        for (int64_t j = 0; j < k; ++j) dev = std::max(dev, std::fabs(G[i * k + j] - (i == j ? 1.0 : 0.0)));
      if (pass >= 2 && dev < 1e-5) break;
      std::vector<double> d(k);
      for (int64_t i = 0; i < k; ++i) d[i] = G[i * k + i] > 0.0 ? 1.0 / std::sqrt(G[i * k + i]) : 1.0;
      for (int64_t i = 0; i < k; ++i)
        for (int64_t j = 0; j < k; ++j) Gs[i * k + j] = d[i] * G[i * k + j] * d[j];
      if (!chol_upper(Gs.data(), (int)k, Rs.data()))
        return ctx_fail(ctx, LIGHTNE_ERR_NUMERIC, "FP64 CholeskyQR: Gram matrix is not positive definite");
      for (int64_t i = 0; i < k; ++i)
        for (int64_t j = 0; j < k; ++j) R[i * k + j] = Rs[i * k + j] / d[j];
      tri_upper_inverse(R.data(), (int)k, Rinv.data());
      LN_CHECK(apply_rows(P1, Rinv.data()));
      if (pass >= 1 && dev < 1e-1) break;
    }
    st.qr_calls++;
    LN_HIP(ctx, hipEventRecord(e[5], s));
    LN_CHECK(mm64(P1, P0));  // A Q over the dead sketch
    st.spmm_calls++;
    LN_HIP(ctx, hipEventRecord(e[6], s));
    LN_HIP(ctx, hipEventSynchronize(e[6]));
    t_qr += lap(e[4], e[5]);
    t_spmm += lap(e[5], e[6]);
    if (it + 1 < q) std::swap(P0, P1);
  }
  // B = P0: B^T B = V diag(S^2) V^T, out = B V[:, :cols] S^-pw
  LN_HIP(ctx, hipEventRecord(e[7], s));
  LN_CHECK(gram_f64(ctx, P0, P0, n, k, k, 1, G.data()));
  auto t0 = std::chrono::steady_clock::now();
  if (!sym_eig(G.data(), (int)k, w.data(), V.data()))
    return ctx_fail(ctx, LIGHTNE_ERR_NUMERIC, "lightne_rsvd: eigensolver did not converge");
  st.t_small_svd = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  S.resize(k);
  for (int64_t j = 0; j < k; ++j) S[j] = std::sqrt(std::max(w[j], 0.0));
  // W padded to k columns (k_panel64 needs widths in multiples of 16)
  const int64_t cw = (cols + 15) / 16 * 16;
  std::vector<double> W(k * cw, 0.0);
  for (int64_t j = 0; j < cols; ++j) {
    if (!(S[j] > 1e-12 * S[0])) continue;
    const double f = std::pow(S[j], -pw);
    for (int64_t i = 0; i < k; ++i) W[i * cw + j] = V[i * k + j] * f;
  }
  // P1 (Q) is dead: its bytes stage the FP32 rows
  float* res = reinterpret_cast<float*>(P1);
  float* res2 = res + (size_t)n * cols;
  for (int64_t r0 = 0; r0 < n; r0 += blk_rows) {
    const int64_t rows = std::min(blk_rows, n - r0);
    LN_CHECK(panel_times_small64(ctx, P0 + r0 * k, rows, k, W.data(), cw, T));
    if (cw == cols) {
      LN_CHECK(f64_to_f32(ctx, T, res + r0 * cols, rows * cols));
    } else {
      LN_CHECK(f64_to_f32(ctx, T, res2, rows * cw));
      LN_HIP(ctx, hipMemcpy2DAsync(res + r0 * cols, cols * sizeof(float), res2, cw * sizeof(float),
                                   cols * sizeof(float), rows, hipMemcpyDeviceToDevice, s));
    }
  }
  if (normalize) {
    LN_CHECK(u_sigma_root_dev(ctx, res, nullptr, n, cols, cols, 1, res2));
    res = res2;
  }
  LN_HIP(ctx, hipMemcpyAsync(out, res, (size_t)n * cols * sizeof(float), hipMemcpyDefault, s));
  LN_HIP(ctx, hipEventRecord(e[8], s));
  LN_HIP(ctx, hipEventSynchronize(e[8]));
  st.t_spmm = t_spmm;
  st.t_qr = t_qr;
  st.t_gemm = lap(e[7], e[8]) - st.t_small_svd;
  st.t_total = lap(e[0], e[8]);
  return LIGHTNE_OK;
}

// FP32 / BF16 schedule up to B = A Q and the eigendecomposition of its FP64
// Gram.  On return *B_out is B (n x k, in the workspace) with S (descending)
// and V (k x k, row-major) such that B^T B = V diag(S^2) V^T; with
// random_project_only *B_out is Y = A Omega and S, V are empty.
static int rsvd_core(lightne_ctx* ctx, lightne_csr* A, const lightne_rsvd_params* p, int q, const float** B_out,
                     std::vector<double>& S, std::vector<double>& V, lightne_rsvd_stats* stp) {
  const int64_t n = A->n, k = p->rank;
  lightne_rsvd_stats& st = *stp;
  st = lightne_rsvd_stats{};
  hipStream_t s = ctx->stream;
  const bool bf_power = p->precision == LIGHTNE_BF16_POWER;
  const size_t pf = (size_t)n * k * sizeof(float);
  // P1: Y, then Q.  P2: the sketch Omega stored as BF16 in its first half
  // (every value is BF16-representable, so the 16-bit panel is the FP32
  // sketch exactly and Y = A Omega is the FP32 product bit for bit at half the
  // gathered bytes), then B.  PB: the BF16_POWER form's rounded Q.
  float *P1 = nullptr, *P2 = nullptr;
  __hip_bfloat16* PB = nullptr;
  float* part = nullptr;  // SpMM partial sums
  LN_CHECK(csr_plan(ctx, A));
  {
    WsCarve wc;
    wc.add((void**)&P1, pf);
    wc.add((void**)&P2, pf);
    if (bf_power) wc.add((void**)&PB, (size_t)n * k * 2);
    // the partial sums behind the panels when they fit before a CSR that
    // lives at the workspace's back (the pipeline's), else a buffer of their
    // own: at n = 2^26 they are 15 GB that would otherwise sit beside the
    // workspace
    const size_t part_b = (size_t)A->plan->nparts * k * sizeof(float);
    const auto& ws = ctx->bufs["ws"];
    const char* lo = (const char*)std::min({(const void*)A->row_ptr, (const void*)A->col, (const void*)A->val});
    const bool csr_in_ws = ws.p && lo >= (const char*)ws.p && lo < (const char*)ws.p + ws.bytes;
    const size_t limit = csr_in_ws ? (size_t)(lo - (const char*)ws.p) : (ws.p ? ws.bytes : 0);
    if (part_b > 0 && wc.end() + ((part_b + 255) & ~(size_t)255) <= limit) wc.add((void**)&part, part_b);
    LN_CHECK(wc.commit(ctx));
  }
  hipEvent_t* e = ctx->ev;
  double t_spmm = 0, t_qr = 0, t_gemm = 0;
  auto lap = [&](hipEvent_t a, hipEvent_t b) { return ev_seconds(a, b); };

  LN_HIP(ctx, hipEventRecord(e[0], s));
  auto* omega = reinterpret_cast<__hip_bfloat16*>(P2);
  if (!p->sparse_project) LN_CHECK(gaussian_fill_bf16(ctx, p->seed, 0u, n * k, omega));
  LN_HIP(ctx, hipEventRecord(e[1], s));
  LN_CHECK(csr_plan(ctx, A));
  LN_HIP(ctx, hipEventRecord(e[2], s));
  if (p->sparse_project)
    LN_CHECK(sparse_project_sketch(ctx, A->row_ptr, A->col, A->val, n, k, p->seed, p->sparse_project_s, P1, false));
  else
    LN_CHECK(csr_mm(ctx, A, (const void*)omega, PANEL_BF16, k, P1, part));
  LN_HIP(ctx, hipEventRecord(e[3], s));
  LN_HIP(ctx, hipEventSynchronize(e[3]));
  st.t_sketch = lap(e[0], e[1]);
  t_spmm += lap(e[2], e[3]);
  st.t_spmm_sketch = lap(e[2], e[3]);
  st.spmm_calls = 1;
  if (p->random_project_only) {
    *B_out = P1;
    S.clear();
    V.clear();
    st.t_spmm = t_spmm;
    st.t_total = lap(e[0], e[3]);
    return LIGHTNE_OK;
  }
  int passes = 0;
  for (int it = 0; it < q; ++it) {
    LN_HIP(ctx, hipEventRecord(e[4], s));
    LN_CHECK(cholqr2(ctx, P1, n, k, nullptr, &passes));
    st.qr_calls++;
    LN_HIP(ctx, hipEventRecord(e[5], s));
    // power iterations gather BF16 in the BF16_POWER form; the last product
    // B = A Q gathers the FP32 Q (the refinement pass)
    const bool gather_bf = bf_power && (it + 1 < q);
    if (gather_bf) LN_CHECK(f32_to_bf16(ctx, P1, PB, n * k));
    LN_CHECK(csr_mm(ctx, A, gather_bf ? (const void*)PB : (const void*)P1, gather_bf ? PANEL_BF16 : PANEL_F32, k, P2,
                    part));
    st.spmm_calls++;
    LN_HIP(ctx, hipEventRecord(e[6], s));
    LN_HIP(ctx, hipEventSynchronize(e[6]));
    t_qr += lap(e[4], e[5]);
    t_spmm += lap(e[5], e[6]);
    if (it + 1 < q) std::swap(P1, P2);  // next pass orthonormalises A Q
  }
  // Thin SVD of B: B^T B = V diag(S^2) V^T (FP64 Gram, host FP64 symmetric
  // eigensolver), U = B V S^-1.  The reference orthonormalises Z = B P for a
  // Gaussian k x k P and takes the SVD of C = Z^T B, U = Z U_C
  // (mklredsvd.cc:289-400); range(Z) = range(B) for every invertible P, so U
  // and S are B's left singular vectors and values in both formulations.
  LN_HIP(ctx, hipEventRecord(e[7], s));
  std::vector<double> G(k * k), w(k);
  V.assign(k * k, 0.0);
  LN_CHECK(gram_f64(ctx, P2, P2, n, k, k, 1, G.data()));
  LN_HIP(ctx, hipEventRecord(e[8], s));
  auto t0 = std::chrono::steady_clock::now();
  if (!sym_eig(G.data(), (int)k, w.data(), V.data()))
    return ctx_fail(ctx, LIGHTNE_ERR_NUMERIC, "lightne_rsvd: eigensolver did not converge");
  st.t_small_svd = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  S.resize(k);
  for (int64_t j = 0; j < k; ++j) S[j] = std::sqrt(std::max(w[j], 0.0));
  t_gemm += lap(e[7], e[8]);
  *B_out = P2;
  st.t_spmm = t_spmm;
  st.t_qr = t_qr;
  st.t_gemm = t_gemm;
  st.t_total = lap(e[0], e[8]) + st.t_small_svd;
  return LIGHTNE_OK;
}

// W = V[:, :cols] diag(S^-p) for the panel product B W; columns whose singular
// value is below 1e-6 S_0 (numerically zero for an FP32 B) are set to zero.
static std::vector<double> right_factor(const std::vector<double>& V, const std::vector<double>& S, int64_t k,
                                        int64_t cols, double pw) {
  std::vector<double> W(k * cols, 0.0);
  for (int64_t j = 0; j < cols; ++j) {
    if (!(S[j] > 1e-6 * S[0])) continue;
    const double f = std::pow(S[j], -pw);
    for (int64_t i = 0; i < k; ++i) W[i * cols + j] = V[i * k + j] * f;
  }
  return W;
}

int lightne_rsvd(lightne_ctx* ctx, const lightne_csr* Ac, const lightne_rsvd_params* p, float* U_out, float* S_host,
                 lightne_rsvd_stats* stats) {
  if (!ctx || !Ac || !p || !U_out) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_rsvd: NULL argument");
  auto* A = const_cast<lightne_csr*>(Ac);
  const int64_t n = A->n, k = p->rank;
  if (k <= 0 || k > n)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_rsvd: rank must be in [1, n] (got %lld)", (long long)k);
  if (p->precision != LIGHTNE_FP32 && p->precision != LIGHTNE_BF16 && p->precision != LIGHTNE_FP64 &&
      p->precision != LIGHTNE_BF16_POWER && p->precision != LIGHTNE_FP64_LEAN)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_rsvd: unknown precision");
  const int q = std::max(1, (int)p->power_iters);
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  if (p->upper) {
    // symmetric-upper storage (mklredsvd.cc:95-98): run on the expanded matrix
    lightne_csr F;
    LN_CHECK(csr_expand_upper(ctx, A, &F));
    lightne_rsvd_params pf = *p;
    pf.upper = 0;
    const int rc = lightne_rsvd(ctx, &F, &pf, U_out, S_host, stats);
    if (F.plan) {
      spmm_plan_free(*F.plan);
      delete F.plan;
    }
    return rc;
  }
  if (rsvd_is_wide(k)) {  // k > 256 or not a multiple of 16 (wide.hip)
    lightne_rsvd_stats st{};
    std::vector<double> S;
    LN_CHECK(rsvd_wide(ctx, A, p, k, 1.0, 0, U_out, k, S, &st));
    if (S_host && !p->random_project_only)
      for (int64_t j = 0; j < k; ++j) S_host[j] = (float)S[j];
    if (stats) *stats = st;
    return LIGHTNE_OK;
  }
  if (p->precision == LIGHTNE_FP64) return rsvd64(ctx, A, p, q, U_out, S_host, stats);
  if (p->precision == LIGHTNE_FP64_LEAN) {
    lightne_rsvd_stats st{};
    std::vector<double> S;
    LN_CHECK(rsvd64_lean(ctx, A, p, q, k, 1.0, 0, U_out, S, &st));
    if (S_host && !p->random_project_only)
      for (int64_t j = 0; j < k; ++j) S_host[j] = (float)S[j];
    if (stats) *stats = st;
    return LIGHTNE_OK;
  }
  lightne_rsvd_stats st{};
  const float* B = nullptr;
  std::vector<double> S, V;
  LN_CHECK(rsvd_core(ctx, A, p, q, &B, S, V, &st));
  hipStream_t s = ctx->stream;
  if (p->random_project_only) {
    LN_HIP(ctx, hipMemcpyAsync(U_out, B, (size_t)n * k * sizeof(float), hipMemcpyDeviceToDevice, s));
  } else {
    LN_HIP(ctx, hipEventRecord(ctx->ev[9], s));
    const std::vector<double> W = right_factor(V, S, k, k, 1.0);  // U = B V S^-1
    LN_CHECK(panel_times_small(ctx, B, n, k, k, W.data(), k, U_out, 0, nullptr));
    LN_HIP(ctx, hipEventRecord(ctx->ev[10], s));
    LN_HIP(ctx, hipEventSynchronize(ctx->ev[10]));
    const double tg = ev_seconds(ctx->ev[9], ctx->ev[10]);
    st.t_gemm += tg;
    st.t_total += tg;
    if (S_host)
      for (int64_t j = 0; j < k; ++j) S_host[j] = (float)S[j];
  }
  LN_HIP(ctx, hipStreamSynchronize(s));
  if (stats) *stats = st;
  return LIGHTNE_OK;
}

int lightne_dense_gram(lightne_ctx* ctx, const float* X, const float* Y, int64_t n, int64_t k1, int64_t k2,
                       int32_t symmetric, double* G_host) {
  if (!ctx || !X || !Y || !G_host || n < 0) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_dense_gram: bad argument");
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  return gram_f64(ctx, X, Y, n, k1, k2, symmetric, G_host);
}

int lightne_dense_gram_noalloc(lightne_ctx* ctx, const float* X, const float* Y, int64_t n, int64_t k1, int64_t k2,
                               int32_t symmetric, double* G_host) {
  if (!ctx || !X || !Y || !G_host || n < 0)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_dense_gram_noalloc: bad argument");
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  return gram_f64_noalloc(ctx, X, Y, n, k1, k2, symmetric, G_host);
}

int lightne_dense_panel_gemm(lightne_ctx* ctx, const float* X, int64_t n, int64_t k1, const double* W_host,
                             int64_t k2, float* Out, int32_t normalize_rows) {
  if (!ctx || !X || !W_host || !Out || n < 0)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_dense_panel_gemm: bad argument");
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  LN_CHECK(panel_times_small(ctx, X, n, k1, k1, W_host, k2, Out, normalize_rows, nullptr));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LIGHTNE_OK;
}

int lightne_dense_cholqr(lightne_ctx* ctx, float* Y, int64_t n, int64_t k, double* R_host, int32_t* passes_out) {
  if (!ctx || !Y || n < k || k <= 0 || k % 16 || k > 256)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_dense_cholqr: bad argument");
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  int passes = 0;
  LN_CHECK(cholqr2(ctx, Y, n, k, R_host, &passes));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (passes_out) *passes_out = passes;
  return LIGHTNE_OK;
}

int lightne_u_sigma_root(lightne_ctx* ctx, const float* U, const float* S_host, int64_t n, int64_t rank, int64_t dim,
                         int32_t normalize, float* emb_out) {
  if (!ctx || !U || !emb_out || n <= 0 || dim <= 0 || dim > rank)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_u_sigma_root: bad argument");
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  LN_CHECK(u_sigma_root_dev(ctx, U, S_host, n, rank, dim, normalize, emb_out));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LIGHTNE_OK;
}

// Stage 2 and the embedding on a CSR: emb = U[:, :dim] sqrt(S), row-normalised
// when asked (PathEmbed.h:371-408); A is freed here.
static int svd_embed(lightne_ctx* ctx, lightne_csr* A, const lightne_rsvd_params* rp, int64_t dim, int32_t normalize,
                     float* emb_out, std::vector<float>& S, lightne_rsvd_stats* rst) {
  const int64_t n = A->n, k = rp->rank;
  S.assign(k, 0.f);
  int rc = LIGHTNE_OK;
  if (rsvd_is_wide(k) && !rp->upper) {  // k > 256 or not a multiple of 16 (wide.hip)
    if (k > n) {
      lightne_csr_free(A);
      return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "embedding: rank exceeds n");
    }
    std::vector<double> Sd;
    if (rp->random_project_only) {  // Y = A Omega, then U sqrt(S) with S = 1
      auto* Y = (float*)ctx_buf(ctx, "netsmf_y", (size_t)n * k * sizeof(float));
      rc = Y ? rsvd_wide(ctx, A, rp, k, 1.0, 0, Y, k, Sd, rst) : LIGHTNE_ERR_OOM;
      if (rc == LIGHTNE_OK) rc = u_sigma_root_dev(ctx, Y, nullptr, n, k, dim, normalize, emb_out);
    } else {
      rc = rsvd_wide(ctx, A, rp, dim, 0.5, normalize, emb_out, dim, Sd, rst);
    }
    lightne_csr_free(A);
    if (rc != LIGHTNE_OK) return rc;
    for (int64_t j = 0; j < (int64_t)Sd.size(); ++j) S[j] = (float)Sd[j];
    return LIGHTNE_OK;
  }
  if (rp->precision == LIGHTNE_FP64) {
    // accuracy reference: explicit U, then U sqrt(S)
    auto* U = (float*)ctx_buf(ctx, "netsmf_u", (size_t)n * k * sizeof(float));
    rc = U ? lightne_rsvd(ctx, A, rp, U, S.data(), rst) : LIGHTNE_ERR_OOM;
    lightne_csr_free(A);
    if (rc != LIGHTNE_OK) return rc;
    return u_sigma_root_dev(ctx, U, rp->random_project_only ? nullptr : S.data(), n, k, dim, normalize, emb_out);
  }
  if (rp->precision == LIGHTNE_FP64_LEAN) {
    if (k > n || k % 16 || k <= 0 || k > 256 || rp->upper || rp->random_project_only) {
      lightne_csr_free(A);
      return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG,
                      "embedding: the lean FP64 form needs a rank in [16, 256] (multiple of 16), full storage and "
                      "the whole randomized SVD");
    }
    std::vector<double> Sd;
    rc = rsvd64_lean(ctx, A, rp, std::max(1, (int)rp->power_iters), dim, 0.5, normalize, emb_out, Sd, rst);
    lightne_csr_free(A);
    if (rc != LIGHTNE_OK) return rc;
    for (int64_t j = 0; j < k; ++j) S[j] = (float)Sd[j];
    return LIGHTNE_OK;
  }
  // emb = B (V S^-1/2)[:, :dim], row-normalised in the same pass
  if (k > n || k % 16 || k <= 0 || k > 256 || rp->upper ||
      (rp->precision != LIGHTNE_FP32 && rp->precision != LIGHTNE_BF16 && rp->precision != LIGHTNE_BF16_POWER)) {
    lightne_csr_free(A);
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "embedding: bad rank, precision or upper storage");
  }
  const float* B = nullptr;
  std::vector<double> Sd, V;
  rc = rsvd_core(ctx, A, rp, std::max(1, (int)rp->power_iters), &B, Sd, V, rst);
  lightne_csr_free(A);
  if (rc != LIGHTNE_OK) return rc;
  hipStream_t s = ctx->stream;
  LN_HIP(ctx, hipEventRecord(ctx->ev[9], s));
  if (rp->random_project_only) {
    LN_CHECK(u_sigma_root_dev(ctx, B, nullptr, n, k, dim, normalize, emb_out));
  } else {
    const std::vector<double> W = right_factor(V, Sd, k, dim, 0.5);
    LN_CHECK(panel_times_small(ctx, B, n, k, k, W.data(), dim, emb_out, normalize, nullptr));
    for (int64_t j = 0; j < k; ++j) S[j] = (float)Sd[j];
  }
  LN_HIP(ctx, hipEventRecord(ctx->ev[10], s));
  LN_HIP(ctx, hipEventSynchronize(ctx->ev[10]));
  const double tg = ev_seconds(ctx->ev[9], ctx->ev[10]);
  rst->t_gemm += tg;
  rst->t_total += tg;
  return LIGHTNE_OK;
}

int lightne_netsmf(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* sp,
                   const lightne_rsvd_params* rp, int64_t dim, int32_t normalize, float* emb_out, float* S_host,
                   lightne_netsmf_stats* stats) {
  if (!ctx || !g || !sp || !rp || !emb_out) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_netsmf: NULL argument");
  if (dim <= 0 || dim > rp->rank) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_netsmf: dim must be in [1, rank]");
  CallScope scope(ctx);
  lightne_netsmf_stats st{};
  auto t0 = std::chrono::steady_clock::now();
  lightne_csr* A = nullptr;
  // -upper 1 (PathEmbed.h:284-309 with mklredsvd.cc:95-98) denotes the same
  // symmetric matrix: stage 1 writes both triangles at once and stage 2 runs
  // on it; the reported nnz is the upper storage's, as the reference's
  lightne_sparsify_params spf = *sp;
  lightne_rsvd_params rpf = *rp;
  const bool upper = sp->upper || rp->upper;
  spf.upper = 0;
  rpf.upper = 0;
  sp = &spf;
  rp = &rpf;
  // stage 2 carves its panels from the workspace front (rsvd_core)
  const size_t pnl = ((size_t)g->n * rp->rank * sizeof(float) + 255) & ~(size_t)255;
  const bool bf_power = rp->precision == LIGHTNE_BF16_POWER;
  const size_t blk64 = ((size_t)std::min<int64_t>(g->n, (int64_t)1 << 20) * rp->rank * sizeof(double) + 255) &
                       ~(size_t)255;
  const size_t front = rsvd_is_wide(rp->rank)              ? rsvd_wide_front(g->n, rp->rank, rp->precision)
                       : rp->precision == LIGHTNE_FP64        ? 4 * 2 * pnl
                       : rp->precision == LIGHTNE_FP64_LEAN ? 2 * 2 * pnl + blk64
                                                            : 2 * pnl + (bf_power ? pnl / 2 + 256 : 0);
  // emb_out is written only by stage 2: stage 1 may use it as scratch when it
  // is device memory of this context's device, apart from the graph
  float* xs = nullptr;
  size_t xs_bytes = 0;
  {
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, emb_out) == hipSuccess && at.type == hipMemoryTypeDevice &&
        at.device == ctx->device) {
      const size_t b = (size_t)g->n * (size_t)dim * sizeof(float);
      auto apart = [&](const void* q, size_t qb) {
        const char *a0 = (const char*)emb_out, *q0 = (const char*)q;
        return !q || qb == 0 || a0 + b <= q0 || q0 + qb <= a0;
      };
      if (apart(g->offsets, (size_t)(g->n + 1) * 8) && apart(g->edges, (size_t)g->m * 4)) {
        xs = emb_out;
        xs_bytes = b;
      }
    } else {
      (void)hipGetLastError();  // a host pointer: the query's error is not this call's
    }
  }
  LN_CHECK(sparsify_impl(ctx, g, sp, &A, &st.sparsify, true, front, xs, xs_bytes));
  if (upper) st.sparsify.nnz = st.sparsify.kept_upper;
  auto t1 = std::chrono::steady_clock::now();
  std::vector<float> S;
  LN_CHECK(svd_embed(ctx, A, rp, dim, normalize, emb_out, S, &st.rsvd));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  auto t2 = std::chrono::steady_clock::now();
  if (S_host) std::copy(S.begin(), S.end(), S_host);
  st.t_sparsify = std::chrono::duration<double>(t1 - t0).count();
  st.t_svd = std::chrono::duration<double>(t2 - t1).count();
  st.t_total = std::chrono::duration<double>(t2 - t0).count();
  if (stats) *stats = st;
  return LIGHTNE_OK;
}

int lightne_ne_zhang(lightne_ctx* ctx, const lightne_graph* g, const lightne_rsvd_params* rp, int64_t dim,
                     float* emb_out, float* S_host, lightne_rsvd_stats* stats) {
  if (!ctx || !g || !rp || !emb_out) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_ne_zhang: NULL argument");
  if (dim <= 0 || dim > rp->rank) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_ne_zhang: dim must be in [1, rank]");
  if (g->n <= 0 || !g->offsets || (g->m > 0 && !g->edges) || g->n >= ((int64_t)1 << 31))
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_ne_zhang: bad graph (n must be < 2^31)");
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  // the matrix keeps the graph's row pointers and neighbour order
  auto* val = (float*)ctx_buf(ctx, "nz_val", (size_t)std::max<int64_t>(g->m, 1) * sizeof(float));
  if (!val) return LIGHTNE_ERR_OOM;
  LN_CHECK(ne_zhang_values(ctx, g, val));
  auto* A = new lightne_csr();
  A->n = g->n;
  A->nnz = g->m;
  A->owns = false;
  A->row_ptr = const_cast<int64_t*>(g->offsets);
  A->col = reinterpret_cast<int32_t*>(const_cast<uint32_t*>(g->edges));
  A->val = val;
  lightne_rsvd_stats st{};
  std::vector<float> S;
  LN_CHECK(svd_embed(ctx, A, rp, dim, 1, emb_out, S, &st));  // PathEmbed.h:459 normalises
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (S_host) std::copy(S.begin(), S.end(), S_host);
  if (stats) *stats = st;
  return LIGHTNE_OK;
}

}  // extern "C"
