// Stage 3: spectral propagation (LightNE/SpectralPropagation.h:45-229).
//
// M = (1 - mu) I - D~^-1 A~ over the INPUT graph with self loops (:51-75) is
// never materialised: the SpMM epilogue forms (M x)_i from the neighbour sum.
// The Chebyshev recurrence of order K (:94-208) runs as 2 gathers per step with
// the AXPYs fused into the second gather's epilogue; y = a - conv is written by
// the last step, z = (A + I) y by one more gather (:210-221), and the dense
// tall SVD (:16-43) is the FP64 Gram z^T z, its host FP64 eigendecomposition
// (sigma^2, V) and one panel product z V S^-1/2 with a row-normalising
// epilogue (z's column space is its left singular space: no tall QR).
#include "lightne_internal.h"
#include "spmm_epi.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace {

// modified Bessel function of the first kind by its power series,
// I_k(x) = sum_j (x/2)^(2j+k) / (j! (j+k)!)  (replaces boost::math::cyl_bessel_i,
// SpectralPropagation.h:8,12)
double bessel_i(int k, double x) {
  const double h = 0.5 * x;
  double term = 1.0;
  for (int i = 1; i <= k; ++i) term *= h / i;  // (x/2)^k / k!
  double sum = term;
  for (int j = 1; j < 200; ++j) {
    term *= (h * h) / ((double)j * (double)(j + k));
    sum += term;
    if (term < sum * 1e-18) break;
  }
  return sum;
}

struct GraphPlan {
  SpmmPlan P;
  ~GraphPlan() { spmm_plan_free(P); }
};

__global__ void k_y_from_conv(const float* __restrict__ a, const float* __restrict__ conv, float* __restrict__ y,
                              int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a[i] + (-1.0f) * conv[i];
}

int check_prop(lightne_ctx* ctx, const lightne_graph* g, const float* a, int64_t dim, const lightne_prop_params* p) {
  if (!ctx || !g || !a || !p) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "propagate: NULL argument");
  if (g->n <= 0 || !g->offsets || (g->m > 0 && !g->edges)) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "propagate: bad graph");
  if (dim <= 0 || dim > 1024) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "propagate: dim must be in [1, 1024]");
  if (p->precision != LIGHTNE_FP32 && p->precision != LIGHTNE_BF16 && p->precision != LIGHTNE_BF16_POWER)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "propagate: FP32, BF16 or BF16_POWER iterates");
  return LIGHTNE_OK;
}

}  // namespace

// Chebyshev weights as the reference forms them: c_0 = I_0(s),
// c_i = (i odd ? -2 : 2) I_i(s), each rounded to float at the AXPY (:132-197).
void cheb_coeffs(int64_t order, double theta, std::vector<float>& c) {
  const int64_t K = std::max<int64_t>(order, 2);
  c.assign(K, 0.f);
  c[0] = (float)bessel_i(0, theta);
  c[1] = (float)(-2 * bessel_i(1, theta));
  for (int64_t i = 2; i < K; ++i) c[i] = (float)(((i & 1) ? -2 : 2) * bessel_i((int)i, theta));
}

double lightne_bessel_i_impl(int k, double x) { return bessel_i(k, x); }
extern "C" double lightne_bessel_i(int k, double x) { return bessel_i(k, x); }

// Workspace propagate_forward carves (from offset 0).  FP32: two iterates,
// the intermediate M T_i and the accumulator (the new iterate overwrites
// T_{i-1} element by element in the epilogue that reads it).  BF16: the same
// in BF16 plus an FP32 accumulator and two FP32 panels for the FP32 first two
// steps (T_1, T_2), whose intermediate spans the second iterate and the
// intermediate.
size_t prop_ws_bytes(int64_t n, int64_t dim, bool bf) {
  const size_t pe = (size_t)n * dim;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return bf ? 3 * al(pe * 2) + 3 * al(pe * 4) : 4 * al(pe * 4);
}

// forward recurrence; y == NULL writes y over the accumulator (internal use)
int propagate_forward(lightne_ctx* ctx, const lightne_graph* g, const float* a, int64_t dim,
                      const lightne_prop_params* p, float* y, float* z, lightne_prop_stats* st, SpmmPlan& P) {
  const int64_t n = g->n, K = std::max<int64_t>(p->order, 2);
  const bool bf = p->precision != LIGHTNE_FP32;
  const int xt = bf ? PANEL_BF16 : PANEL_F32;
  const size_t esz = bf ? 2 : 4;
  const size_t pe = (size_t)n * dim;
  void* B[2];
  void* tmp = nullptr;
  float* conv = nullptr;
  float *x32 = nullptr, *x32b = nullptr;
  WsCarve wc;
  wc.add(&B[0], pe * esz);
  wc.add(&B[1], pe * esz);
  wc.add(&tmp, pe * esz);
  wc.add((void**)&conv, pe * sizeof(float));
  if (bf) {
    wc.add((void**)&x32, pe * sizeof(float));
    wc.add((void**)&x32b, pe * sizeof(float));
  }
  LN_CHECK(wc.commit(ctx));
  // FP32 first step: M a into tmp32, T_1 into lx1_32
  float* tmp32 = bf ? (float*)B[1] : (float*)tmp;  // BF16: spans B[1] and tmp
  float* lx1_32 = bf ? x32 : (float*)B[1];
  if (!y) y = conv;
  std::vector<float> c;
  cheb_coeffs(p->order, (double)p->theta, c);
  hipStream_t s = ctx->stream;
  // Lx0 = a: the FP32 steps read the caller's a
  const void* lx0 = a;
  SpmmArgs base;
  base.n = n;
  base.k = dim;
  base.rp = g->offsets;
  base.gcol = g->edges;
  base.impl_mode = IMPL_PROP;
  base.mu = p->mu;
  base.a_in = a;
  base.conv = conv;
  int calls = 0;
  LN_HIP(ctx, hipEventRecord(ctx->ev[13], s));
  // The first two steps (T_1 and T_2, which carry the three largest weights
  // c_0 = I_0, c_1 = -2 I_1, c_2 = 2 I_2; c_3 is 2 I_3(0.5) = 5e-3) always
  // gather and store FP32: in the BF16 form they are the refinement pass, and
  // the BF16 roundings (2^-9) of the later iterates enter y weighted by
  // |c_i| <= 5e-3.  Later iterates are gathered in the panel type.
  // tmp = M a   (:101-111)
  {
    SpmmArgs t = base;
    t.X = a;
    t.out = tmp32;
    t.alpha = 1.f;
    LN_CHECK(spmm_run(ctx, P, t, PANEL_F32, EPI_STORE, true));
    ++calls;
  }
  // Lx1 = 0.5 M tmp - a ; conv = c0 a + c1 Lx1   (:114-148)
  void* lx1 = lx1_32;
  {
    SpmmArgs t = base;
    t.X = tmp32;
    t.alpha = 0.5f;
    t.c0 = c[0];
    t.c1 = c[1];
    t.lx_out = lx1_32;
    LN_CHECK(spmm_run(ctx, P, t, PANEL_F32, EPI_CHEB1, true));
    ++calls;
  }
  float* yout = y;
  if (K <= 2) {
    hipLaunchKernelGGL(k_y_from_conv, dim3(grid_blocks((int64_t)pe, 256)), dim3(256), 0, s, a, conv, yout, (int64_t)pe);
    LN_LAUNCH(ctx);
  }
  // Every other step leaves its c_i T_i to the next one, which reads T_i as
  // its Lx1 anyway (one read and one write of conv fewer per pair of steps):
  // with FP32 iterates from T_2 on, the same sums in the same order; with
  // BF16 iterates from T_4 on, the next step adding c_i times T_i rounded to
  // BF16, whose weight |c_i| <= 2 I_4(1/2) = 3.2e-4 leaves that rounding
  // below the iterates' own.
  int64_t pend = -1;
  for (int64_t i = 2; i < K; ++i) {
    // the BF16 form's second FP32 step: T_2 from the FP32 T_1 and a
    const bool f32_step = bf && i == 2;
    const int xti = f32_step ? PANEL_F32 : xt;
    void* tmpi = f32_step ? (void*)tmp32 : tmp;
    // T_{i} overwrites T_{i-2} (never the caller's a)
    void* lx2 = f32_step ? (void*)x32b : (lx0 == (const void*)a) ? B[0] : const_cast<void*>(lx0);
    {
      SpmmArgs t = base;  // tmp = M Lx1  (:153-163)
      t.X = lx1;
      t.out = tmpi;
      t.alpha = 1.f;
      LN_CHECK(spmm_run(ctx, P, t, xti, EPI_STORE, true));
      ++calls;
    }
    SpmmArgs t = base;  // Lx2 = M tmp - Lx0 - 2 Lx1 ; conv += c_i Lx2  (:166-197)
    t.X = tmpi;
    t.alpha = 1.f;
    t.c1 = c[i];
    t.lx1 = lx1;
    t.lx0 = lx0;
    if (!bf || i >= 4) {
      if (pend >= 0) {
        t.conv_mode = 2;
        t.c_prev = c[pend];
        pend = -1;
      } else if (i + 1 < K) {
        t.conv_mode = 1;
        pend = i;
      }
    }
    if (i + 1 < K) {
      t.lx_out = lx2;
      LN_CHECK(spmm_run(ctx, P, t, xti, EPI_CHEB, true));
    } else {
      t.y_out = yout;
      LN_CHECK(spmm_run(ctx, P, t, xti, EPI_CHEB_LAST, true));
    }
    ++calls;
    if (f32_step && i + 1 < K) {
      // later steps gather BF16: T_1 and T_2 rounded into the BF16 iterates
      // (tmp32, spanning B[1] and tmp, is dead)
      LN_CHECK(f32_to_bf16(ctx, (const float*)lx1, (__hip_bfloat16*)B[0], (int64_t)pe));
      LN_CHECK(f32_to_bf16(ctx, (const float*)lx2, (__hip_bfloat16*)B[1], (int64_t)pe));
      lx0 = B[0];
      lx1 = B[1];
    } else {
      lx0 = lx1;
      lx1 = lx2;
    }
  }
  LN_HIP(ctx, hipEventRecord(ctx->ev[14], s));
  if (z) {
    SpmmArgs t = base;  // z = (A + I) y  (:210-221)
    t.impl_mode = IMPL_AGG;
    t.X = yout;
    t.out = z;
    t.alpha = 1.f;
    LN_CHECK(spmm_run(ctx, P, t, PANEL_F32, EPI_STORE_F32, true));
    ++calls;
  }
  LN_HIP(ctx, hipEventRecord(ctx->ev[15], s));
  LN_HIP(ctx, hipEventSynchronize(ctx->ev[15]));
  if (st) {
    st->t_cheb = ev_seconds(ctx->ev[13], ctx->ev[14]);
    st->t_aggregate = ev_seconds(ctx->ev[14], ctx->ev[15]);
    st->spmm_calls = calls;
  }
  return LIGHTNE_OK;
}

// get_embedding_dense (SpectralPropagation.h:16-43): thin SVD of mat (n x rank)
// from its FP64 Gram, mat^T mat = V S^2 V^T (U = mat V S^-1), then
// emb = rownorm(U[:, :dim] sqrt(S)) = rownorm(mat V[:, :dim] S^-1/2) in one
// panel product with the row-normalisation epilogue.  mat is not modified.
int embed_dense_dev(lightne_ctx* ctx, const float* mat, int64_t n, int64_t rank, int64_t dim, float* emb_out,
                    float* S_host) {
  if (dim <= 0 || dim > rank || n < rank)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "embed_dense: need 1 <= dim <= rank <= n");
  if (rsvd_is_wide(rank) || dim % 16)  // wide.hip: rocBLAS / rocSOLVER on FP64 row blocks
    return embed_dense_wide(ctx, mat, n, rank, rank, dim, emb_out, S_host);
  std::vector<double> G(rank * rank), V(rank * rank), w(rank), S(rank);
  LN_CHECK(gram_f64(ctx, mat, mat, n, rank, rank, 1, G.data()));
  if (!sym_eig(G.data(), (int)rank, w.data(), V.data()))
    return ctx_fail(ctx, LIGHTNE_ERR_NUMERIC, "embed_dense: eigensolver did not converge");
  for (int64_t j = 0; j < rank; ++j) S[j] = std::sqrt(std::max(w[j], 0.0));
  // columns whose singular value is numerically zero for an FP32 panel give
  // zero columns (their U sqrt(S) is zero to that precision)
  std::vector<double> W(rank * dim, 0.0);
  for (int64_t j = 0; j < dim; ++j) {
    if (!(S[j] > 1e-6 * S[0])) continue;
    const double f = 1.0 / std::sqrt(S[j]);
    for (int64_t i = 0; i < rank; ++i) W[i * dim + j] = V[i * rank + j] * f;
  }
  LN_CHECK(panel_times_small(ctx, mat, n, rank, rank, W.data(), dim, emb_out, 1, nullptr));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (S_host)
    for (int64_t j = 0; j < rank; ++j) S_host[j] = (float)S[j];
  return LIGHTNE_OK;
}

// A dim that is not a multiple of 16 runs on a copy padded with zero columns:
// the operator acts on each column alone, so the padding stays zero and the
// first dim columns are the result.  Returns the padded copy (context buffer)
// and its width.
static int pad_columns(lightne_ctx* ctx, const float* a, int64_t n, int64_t dim, const char* name, const float** ap,
                       int64_t* dp) {
  *dp = (dim + 15) / 16 * 16;
  *ap = a;
  if (*dp == dim) return LIGHTNE_OK;
  auto* buf = (float*)ctx_buf(ctx, name, (size_t)n * *dp * sizeof(float));
  if (!buf) return LIGHTNE_ERR_OOM;
  LN_HIP(ctx, hipMemset2DAsync(buf + dim, *dp * 4, 0, (*dp - dim) * 4, n, ctx->stream));
  LN_HIP(ctx, hipMemcpy2DAsync(buf, *dp * 4, a, dim * 4, dim * 4, n, hipMemcpyDeviceToDevice, ctx->stream));
  *ap = buf;
  return LIGHTNE_OK;
}

extern "C" {

int lightne_propagate(lightne_ctx* ctx, const lightne_graph* g, const float* a, int64_t dim,
                      const lightne_prop_params* p, float* y_out, float* z_out, lightne_prop_stats* stats) {
  LN_CHECK(check_prop(ctx, g, a, dim, p));
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  auto t0 = std::chrono::steady_clock::now();
  GraphPlan gp;
  LN_CHECK(spmm_plan_build(ctx, g->offsets, g->n, gp.P));
  lightne_prop_stats st{};
  const float* ap = a;
  int64_t dp = dim;
  LN_CHECK(pad_columns(ctx, a, g->n, dim, "prop_pad_a", &ap, &dp));
  if (dp != dim) {
    const size_t pb = (size_t)g->n * dp * sizeof(float);
    float* yp = y_out ? (float*)ctx_buf(ctx, "prop_pad_y", pb) : nullptr;
    float* zp = z_out ? (float*)ctx_buf(ctx, "prop_pad_z", pb) : nullptr;
    if ((y_out && !yp) || (z_out && !zp)) return LIGHTNE_ERR_OOM;
    LN_CHECK(propagate_forward(ctx, g, ap, dp, p, yp, zp, &st, gp.P));
    hipStream_t s = ctx->stream;
    if (y_out) LN_HIP(ctx, hipMemcpy2DAsync(y_out, dim * 4, yp, dp * 4, dim * 4, g->n, hipMemcpyDeviceToDevice, s));
    if (z_out) LN_HIP(ctx, hipMemcpy2DAsync(z_out, dim * 4, zp, dp * 4, dim * 4, g->n, hipMemcpyDeviceToDevice, s));
    LN_HIP(ctx, hipStreamSynchronize(s));
  } else {
    LN_CHECK(propagate_forward(ctx, g, a, dim, p, y_out, z_out, &st, gp.P));
  }
  auto t1 = std::chrono::steady_clock::now();
  st.t_total = std::chrono::duration<double>(t1 - t0).count();
  if (stats) *stats = st;
  return LIGHTNE_OK;
}

int lightne_embed_dense(lightne_ctx* ctx, const float* mat, int64_t n, int64_t rank, int64_t dim, float* emb_out,
                        float* S_host) {
  if (!ctx || !mat || !emb_out) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_embed_dense: NULL argument");
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  return embed_dense_dev(ctx, mat, n, rank, dim, emb_out, S_host);
}

int lightne_chebyshev_expansion(lightne_ctx* ctx, const lightne_graph* g, const float* a, int64_t dim,
                                const lightne_prop_params* p, float* emb_out, lightne_prop_stats* stats) {
  LN_CHECK(check_prop(ctx, g, a, dim, p));
  if (!emb_out) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "chebyshev_expansion: NULL output");
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  auto t0 = std::chrono::steady_clock::now();
  GraphPlan gp;
  LN_CHECK(spmm_plan_build(ctx, g->offsets, g->n, gp.P));
  // a dim that is not a multiple of 16 runs on a zero-padded copy of width dp
  const float* ap = a;
  int64_t dp = dim;
  LN_CHECK(pad_columns(ctx, a, g->n, dim, "prop_pad_a", &ap, &dp));
  // z after the recurrence's own workspace (y is written over its
  // accumulator)
  float* z = nullptr;
  WsCarve wc(prop_ws_bytes(g->n, dp, p->precision != LIGHTNE_FP32));
  wc.add((void**)&z, (size_t)g->n * dp * sizeof(float));
  LN_CHECK(wc.commit(ctx));
  lightne_prop_stats st{};
  LN_CHECK(propagate_forward(ctx, g, ap, dp, p, nullptr, z, &st, gp.P));
  auto t1 = std::chrono::steady_clock::now();
  // the zero padding columns of z add zero singular values only: the top dim
  // singular triplets of the padded z are z's
  LN_CHECK(embed_dense_dev(ctx, z, g->n, dp, dim, emb_out, nullptr));
  auto t2 = std::chrono::steady_clock::now();
  st.t_svd = std::chrono::duration<double>(t2 - t1).count();
  st.t_total = std::chrono::duration<double>(t2 - t0).count();
  if (stats) *stats = st;
  return LIGHTNE_OK;
}

}  // extern "C"
