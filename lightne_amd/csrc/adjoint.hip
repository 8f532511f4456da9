// Adjoint of the Chebyshev propagation y = a - sum_{i<K} c_i T_i a
// (SpectralPropagation.h:94-208), with T_0 = a, T_1 = X a,
// T_i = 2 X T_{i-1} - T_{i-2}, X = M^2/2 - I, M = (1-mu) I - D~^-1 A~.
// The reference has no gradients; this is the differentiation form.
//
// Reverse recurrence for lambda_i = dL/dT_i, g = dL/dy:
//   lambda_{K-1} = -c_{K-1} g,
//   lambda_{i-1} = -c_{i-1} g + coef_{i-1} X^T lambda_i - lambda_{i+1}
//   (coef = 2, or 1 for lambda_0 because T_1 = X T_0),
//   dL/da = g + lambda_0.
// M^T = D~ M D~^-1, so with m_i = D~^-1 lambda_i the recurrence uses the
// forward operator M (same SpMM kernel and plan):
//   m_{i-1} = -c_{i-1} D~^-1 g + coef_{i-1} X m_i - m_{i+1},  dL/da = g + D~ m_0.
// Parameters: dX/dmu = -M, so
//   dL/dmu    = -sum_{i>=1} coef_i <lambda_i, M T_{i-1}> = -sum coef_i <D~ M m_i, T_{i-1}>,
//   dL/dtheta = -sum_i c_i'(theta) <g, T_i>, c_0' = I_1, c_i' = (-1)^i (I_{i-1} + I_{i+1}).
// T_{i-1} is recomputed in the reverse sweep from pairs (T_{sC}, T_{sC+1})
// checkpointed every C steps of the forward sweep.
#include "lightne_internal.h"
#include "spmm_epi.h"

#include <algorithm>
#include <cmath>
#include <vector>

double lightne_bessel_i_impl(int k, double x);
void cheb_coeffs(int64_t order, double theta, std::vector<float>& c);

namespace {

constexpr int DOT_BLOCKS = 1024;

// per-block partial of sum_r w_r <x_r, y_r> (w_r = d_r + 1 when rp != NULL)
__global__ __launch_bounds__(256) void k_dot(const float* __restrict__ x, const float* __restrict__ y, int64_t n,
                                             int64_t d, const int64_t* __restrict__ rp, double* __restrict__ part) {
  __shared__ double s_w[4];
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(n, r0 + per);
  double acc = 0.0;
  for (int64_t r = r0; r < r1; ++r) {
    double s = 0.0;
    for (int64_t c = threadIdx.x; c < d; c += 256) s += (double)x[r * d + c] * (double)y[r * d + c];
    if (rp) s *= (double)(rp[r + 1] - rp[r] + 1);
    acc += s;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// out = coef * D~^-1 g  (mode 0)   or   out = g + D~ m  (mode 1)
__global__ void k_deg_scale(const float* __restrict__ g, const float* __restrict__ m, const int64_t* __restrict__ rp,
                            int64_t n, int64_t d, float coef, int mode, float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * d; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / d;
    const double deg = (double)(rp[r + 1] - rp[r] + 1);
    if (mode == 0)
      out[i] = coef * (float)((double)g[i] / deg);
    else
      out[i] = g[i] + (float)(deg * (double)m[i]);
  }
}

struct Adj {
  lightne_ctx* ctx;
  const lightne_graph* g;
  SpmmPlan P;
  int64_t n, d;
  float mu;
  double* part_d = nullptr;
  ~Adj() { spmm_plan_free(P); }

  SpmmArgs base() const {
    SpmmArgs a;
    a.n = n;
    a.k = d;
    a.rp = g->offsets;
    a.gcol = g->edges;
    a.impl_mode = IMPL_PROP;
    a.mu = mu;
    return a;
  }
  int mx(const float* x, float* out, float alpha) {  // out = alpha M x
    SpmmArgs a = base();
    a.X = x;
    a.out = out;
    a.alpha = alpha;
    return spmm_run(ctx, P, a, PANEL_F32, EPI_STORE, true);
  }
  // t1 = X t0 = 0.5 M (M t0) - t0  (tmp holds M t0)
  int first(const float* t0, float* tmp, float* t1) {
    LN_CHECK(mx(t0, tmp, 1.f));
    SpmmArgs a = base();
    a.X = tmp;
    a.alpha = 0.5f;
    a.a_in = t0;
    a.lx_out = t1;
    a.conv = nullptr;
    return spmm_run(ctx, P, a, PANEL_F32, EPI_CHEB1, true);
  }
  // t2 = 2 X t1 - t0 = M (M t1) - 2 t1 - t0
  int step(const float* t0, const float* t1, float* tmp, float* t2) {
    LN_CHECK(mx(t1, tmp, 1.f));
    SpmmArgs a = base();
    a.X = tmp;
    a.alpha = 1.f;
    a.lx1 = t1;
    a.lx0 = t0;
    a.lx_out = t2;
    a.conv = nullptr;
    return spmm_run(ctx, P, a, PANEL_F32, EPI_CHEB, true);
  }
  int dot(const float* x, const float* y, bool weighted, double* out) {
    hipLaunchKernelGGL(k_dot, dim3(DOT_BLOCKS), dim3(256), 0, ctx->stream, x, y, n, d,
                       weighted ? g->offsets : (const int64_t*)nullptr, part_d);
    LN_LAUNCH(ctx);
    std::vector<double> h(DOT_BLOCKS);
    LN_HIP(ctx, hipMemcpyAsync(h.data(), part_d, DOT_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (double v : h) s += v;
    *out = s;
    return LIGHTNE_OK;
  }
};

}  // namespace

extern "C" int lightne_propagate_backward(lightne_ctx* ctx, const lightne_graph* g, const float* a, int64_t dim,
                                          const lightne_prop_params* p, int32_t with_aggregate,
                                          const float* grad_out, float* grad_a, double* grad_mu_host,
                                          double* grad_theta_host, int32_t checkpoint_every) {
  if (!ctx || !g || !a || !p || !grad_out || !grad_a)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "propagate_backward: NULL argument");
  if (g->n <= 0 || !g->offsets || (g->m > 0 && !g->edges) || dim <= 0 || dim % 16 || dim > 1024)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "propagate_backward: bad graph or dim");
  if (p->precision != LIGHTNE_FP32)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "propagate_backward: the adjoint runs in FP32 only");
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  Adj A;
  A.ctx = ctx;
  A.g = g;
  A.n = g->n;
  A.d = dim;
  A.mu = p->mu;
  LN_CHECK(spmm_plan_build(ctx, g->offsets, g->n, A.P));
  const int64_t K = std::max<int64_t>(p->order, 2);
  const int64_t C = (checkpoint_every <= 1 || checkpoint_every >= K) ? K : checkpoint_every;
  const int64_t S = (K + C - 1) / C;
  const size_t pe = (size_t)g->n * dim * sizeof(float);
  std::vector<float> c;
  cheb_coeffs(p->order, (double)p->theta, c);
  A.part_d = (double*)ctx_buf(ctx, "adj_part", DOT_BLOCKS * sizeof(double));
  auto* tmp = (float*)ctx_buf(ctx, "adj_tmp", pe);
  auto* gb = (float*)ctx_buf(ctx, "adj_g", pe);
  if (!A.part_d || !tmp || !gb) return LIGHTNE_ERR_OOM;
  // g = dL/dy (through z = (A + I) y when aggregated: A + I is symmetric)
  const float* gbar = grad_out;
  if (with_aggregate) {
    SpmmArgs t = A.base();
    t.impl_mode = IMPL_AGG;
    t.X = grad_out;
    t.out = gb;
    t.alpha = 1.f;
    LN_CHECK(spmm_run(ctx, A.P, t, PANEL_F32, EPI_STORE_F32, true));
    gbar = gb;
  }
  // checkpoint pairs (T_{sC}, T_{sC+1}); T_0 is the caller's a
  std::vector<float*> ck(2 * S, nullptr);
  for (int64_t s = 0; s < S; ++s)
    for (int j = 0; j < 2; ++j) {
      if (s == 0 && j == 0) continue;
      char name[32];
      snprintf(name, sizeof(name), "adj_ck%lld_%d", (long long)s, j);
      ck[2 * s + j] = (float*)ctx_buf(ctx, name, pe);
      if (!ck[2 * s + j]) return LIGHTNE_ERR_OOM;
    }
  ck[0] = const_cast<float*>(a);
  // segment buffers for T_{sC+2} .. T_{sC+C-1}
  std::vector<float*> seg(std::max<int64_t>(C - 2, 0));
  for (size_t j = 0; j < seg.size(); ++j) {
    char name[32];
    snprintf(name, sizeof(name), "adj_seg%zu", j);
    seg[j] = (float*)ctx_buf(ctx, name, pe);
    if (!seg[j]) return LIGHTNE_ERR_OOM;
  }
  float* rot[3];
  for (int j = 0; j < 3; ++j) {
    char name[32];
    snprintf(name, sizeof(name), "adj_r%d", j);
    rot[j] = (float*)ctx_buf(ctx, name, pe);
    if (!rot[j]) return LIGHTNE_ERR_OOM;
  }
  // ---- forward sweep: checkpoints and <g, T_i> -------------------------------
  std::vector<double> gT(K, 0.0);
  {
    const float* t0 = a;
    float* t1 = ck[1];
    LN_CHECK(A.first(t0, tmp, t1));
    LN_CHECK(A.dot(gbar, t0, false, &gT[0]));
    LN_CHECK(A.dot(gbar, t1, false, &gT[1]));
    for (int64_t i = 2; i < K; ++i) {
      // T_i goes to its checkpoint slot if it starts a segment pair, else to a
      // rotating scratch buffer
      float* dst = nullptr;
      if (i % C == 0) {
        dst = ck[2 * (i / C)];
      } else if (i % C == 1) {
        dst = ck[2 * (i / C) + 1];
      } else {
        for (int j = 0; j < 3 && !dst; ++j)
          if (rot[j] != t0 && rot[j] != t1) dst = rot[j];
      }
      LN_CHECK(A.step(t0, t1, tmp, dst));
      LN_CHECK(A.dot(gbar, dst, false, &gT[i]));
      t0 = t1;
      t1 = dst;
    }
  }
  // ---- reverse sweep ------------------------------------------------------------
  float* m_next = nullptr;  // m_{i+1}
  float* m_cur = rot[0];    // m_i
  float* m_prev = rot[1];   // m_{i-1}
  float* spare = rot[2];
  {
    const int64_t tot = g->n * dim;
    hipLaunchKernelGGL(k_deg_scale, dim3(grid_blocks(tot, 256)), dim3(256), 0, ctx->stream, gbar,
                       (const float*)nullptr, g->offsets, g->n, dim, -c[K - 1], 0, m_cur);
    LN_LAUNCH(ctx);
  }
  double s_mu = 0.0;
  std::vector<const float*> T(C, nullptr);
  for (int64_t s = S - 1; s >= 0; --s) {
    const int64_t base = s * C, len = std::min(C, K - base);
    T[0] = ck[2 * s];
    if (len > 1) T[1] = ck[2 * s + 1];
    for (int64_t j = 2; j < len; ++j) {
      LN_CHECK(A.step(T[j - 2], T[j - 1], tmp, seg[j - 2]));
      T[j] = seg[j - 2];
    }
    const int64_t i_hi = std::min(base + C, K - 1), i_lo = std::max<int64_t>(base + 1, 1);
    for (int64_t i = i_hi; i >= i_lo; --i) {
      // p = M m_i ; dL/dmu += -coef_i <D~ p, T_{i-1}>
      LN_CHECK(A.mx(m_cur, tmp, 1.f));
      double dm;
      LN_CHECK(A.dot(tmp, T[i - 1 - base], true, &dm));
      s_mu += (i >= 2 ? 2.0 : 1.0) * dm;
      // m_{i-1} = -c_{i-1} D~^-1 g + coef_{i-1} (0.5 M p - m_i) - m_{i+1}
      SpmmArgs t = A.base();
      t.X = tmp;
      t.alpha = 0.5f;
      t.a_in = gbar;
      t.lx1 = m_cur;
      t.lx0 = m_next;
      t.c1 = -c[i - 1];
      t.c0 = (i - 1 >= 1) ? 2.f : 1.f;
      t.lx_out = m_prev;
      LN_CHECK(spmm_run(ctx, A.P, t, PANEL_F32, EPI_ADJ, true));
      float* old_next = m_next ? m_next : spare;
      m_next = m_cur;
      m_cur = m_prev;
      m_prev = old_next;
    }
  }
  // dL/da = g + D~ m_0
  {
    const int64_t tot = g->n * dim;
    hipLaunchKernelGGL(k_deg_scale, dim3(grid_blocks(tot, 256)), dim3(256), 0, ctx->stream, gbar,
                       (const float*)m_cur, g->offsets, g->n, dim, 0.f, 1, grad_a);
    LN_LAUNCH(ctx);
  }
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (grad_mu_host) *grad_mu_host = -s_mu;
  if (grad_theta_host) {
    const double th = (double)p->theta;
    double gt = -lightne_bessel_i_impl(1, th) * gT[0];
    for (int64_t i = 1; i < K; ++i) {
      const double dc = ((i & 1) ? -1.0 : 1.0) *
                        (lightne_bessel_i_impl((int)i - 1, th) + lightne_bessel_i_impl((int)i + 1, th));
      gt -= dc * gT[i];
    }
    *grad_theta_host = gt;
  }
  return LIGHTNE_OK;
}
