// Stage 2 (LightNE/mklredsvd/mklredsvd.cc:93-403) for sketch widths outside
// the 16-wide tiling of rsvd.hip: k > 256 (the reference's own runs use
// -rank 512, LightNE/youtube_lightne.sh:31, and -rank 4096,
// LightNE/blog_lightne.sh:31) or k not a multiple of 16.
//
// Layout: the two n x k panels are padded to kp = 16 ceil(k / 16) columns
// whose padding stays zero (a zero sketch column gives zero columns of Y and
// B and is never orthonormalised).  The sparse products run the CSR SpMM of
// spmm.hip on column slabs of at most 256 columns with row stride kp: every
// gathered slab row is 1 KB contiguous, so the gathers keep their coalescing
// and each slab's partial sums are as for a 256-wide panel.  The dense k x k
// work is O(k^3) and the n x k x k products O(n k^2), so at k = 4096 they are
// library calls on the device, in FP64 over row blocks of the FP32 panels:
//   Gram        rocblas_dsyrk on FP64 copies of row blocks, accumulated
//   Cholesky    rocsolver_dpotrf of the diagonally equilibrated Gram, with
//               the shifted CholeskyQR retry of rsvd.hip on failure
//   Y R^-1      rocblas_dtrsm on each FP64 row block, rounded back to FP32
//   B^T B = V S^2 V^T   rocsolver_dsyevd
//   B V S^-p    rocblas_dgemm on each FP64 row block
// (the reference: util.cc:41-79 geqrf/orgqr, mklredsvd.cc:289-400 gemm and
// gesvd through MKL).  The schedule is rsvd.hip's: CholeskyQR2 of Y, q
// products, and the thin SVD of B = A Q from its Gram instead of the explicit
// Z = orth(B P), C = Z^T B (the same U and S, range(Z) = range(B)).
#include "lightne_internal.h"
#include "spmm_epi.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace {

constexpr int64_t SLAB = 256;                        // SpMM columns per product
constexpr size_t BLOCK_BYTES = (size_t)1 << 30;      // FP64 row block of the dense steps

__global__ void k_rows_to_f64(const float* __restrict__ Y, int64_t ld, int64_t rows, int64_t k,
                              double* __restrict__ out) {
  const int64_t cnt = rows * k;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = (double)Y[(e / k) * ld + e % k];
}

__global__ void k_rows_to_f32(const double* __restrict__ in, int64_t rows, int64_t k, float* __restrict__ Y,
                              int64_t ld) {
  const int64_t cnt = rows * k;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x)
    Y[(e / k) * ld + e % k] = (float)in[e];
}

// X (rows x k, row-major) <- X diag(d)
__global__ void k_scale_cols(double* __restrict__ X, int64_t rows, int64_t k, const double* __restrict__ d) {
  const int64_t cnt = rows * k;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x)
    X[e] *= d[e % k];
}

// G (k x k) <- diag(d) G diag(d) + shift I
__global__ void k_sym_scale(double* __restrict__ G, int64_t k, const double* __restrict__ d, double shift) {
  const int64_t cnt = k * k;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % k, j = e / k;
    G[e] = d[i] * G[e] * d[j] + (i == j ? shift : 0.0);
  }
}

// W (k x cols, row-major) = V[:, k-1-c] f[c]: V holds the eigenvectors of the
// Gram as columns in ascending eigenvalue order (column-major)
__global__ void k_build_w(const double* __restrict__ V, int64_t k, int64_t cols, const double* __restrict__ f,
                          double* __restrict__ W) {
  const int64_t cnt = k * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / cols, c = e % cols;
    W[e] = V[i + (k - 1 - c) * k] * f[c];
  }
}

#define LN_BLAS(ctx, expr)                                                                      \
  do {                                                                                          \
    rocblas_status s_ = (expr);                                                                 \
    if (s_ != rocblas_status_success)                                                           \
      return ctx_fail((ctx), LIGHTNE_ERR_HIP, "%s:%d %s: rocBLAS status %d", __FILE__, __LINE__, \
                      #expr, (int)s_);                                                          \
  } while (0)

int handle(lightne_ctx* ctx, rocblas_handle* h) {
  if (!ctx->blas) {
    rocblas_handle nh = nullptr;
    LN_BLAS(ctx, rocblas_create_handle(&nh));
    ctx->blas = nh;
  }
  *h = (rocblas_handle)ctx->blas;
  LN_BLAS(ctx, rocblas_set_stream(*h, ctx->stream));
  return LIGHTNE_OK;
}

struct Wide {
  lightne_ctx* ctx;
  rocblas_handle h = nullptr;
  int64_t n = 0, k = 0, kp = 0, rb = 0;  // rows, rank, padded width, rows per FP64 block
  double *X64 = nullptr, *O64 = nullptr;  // row block; its product
  double *G = nullptr, *G0 = nullptr, *dv = nullptr, *D = nullptr, *E = nullptr, *W = nullptr;
  rocblas_int* info = nullptr;

  int init(int64_t n_, int64_t k_, int64_t cols) {
    n = n_;
    k = k_;
    kp = (k + 15) / 16 * 16;
    rb = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)(BLOCK_BYTES / ((size_t)std::max(k, cols) * 8))));
    LN_CHECK(handle(ctx, &h));
    X64 = (double*)ctx_buf(ctx, "wide_x64", (size_t)rb * k * 8);
    O64 = (double*)ctx_buf(ctx, "wide_o64", (size_t)rb * cols * 8);
    G = (double*)ctx_buf(ctx, "wide_g", (size_t)k * k * 8);
    G0 = (double*)ctx_buf(ctx, "wide_g0", (size_t)k * k * 8);
    dv = (double*)ctx_buf(ctx, "wide_d", (size_t)std::max(k, cols) * 8);
    D = (double*)ctx_buf(ctx, "wide_eval", (size_t)k * 8);
    E = (double*)ctx_buf(ctx, "wide_e", (size_t)k * 8);
    W = (double*)ctx_buf(ctx, "wide_w", (size_t)k * cols * 8);
    info = (rocblas_int*)ctx_buf(ctx, "wide_info", 256);
    if (!X64 || !O64 || !G || !G0 || !dv || !D || !E || !W || !info) return LIGHTNE_ERR_OOM;
    return LIGHTNE_OK;
  }

  int rows_to_f64(const float* Y, int64_t r0, int64_t rows) {
    hipLaunchKernelGGL(k_rows_to_f64, dim3(grid_blocks(rows * k, 256)), dim3(256), 0, ctx->stream, Y + r0 * kp, kp,
                       rows, k, X64);
    LN_LAUNCH(ctx);
    return LIGHTNE_OK;
  }

  // G = Y^T Y (k x k, FP64; upper triangle, column-major) over FP64 row blocks
  int gram(const float* Y) {
    const double one = 1.0, zero = 0.0;
    for (int64_t r0 = 0; r0 < n; r0 += rb) {
      const int64_t rows = std::min(rb, n - r0);
      LN_CHECK(rows_to_f64(Y, r0, rows));
      // column-major, the row block is the k x rows matrix X^T: G += X^T X
      LN_BLAS(ctx, rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_none, (rocblas_int)k, (rocblas_int)rows,
                                 &one, X64, (rocblas_int)k, r0 ? &one : &zero, G, (rocblas_int)k));
    }
    return LIGHTNE_OK;
  }

  int host_upper(std::vector<double>& Gh) {
    Gh.resize(k * k);
    LN_HIP(ctx, hipMemcpyAsync(Gh.data(), G, (size_t)k * k * 8, hipMemcpyDeviceToHost, ctx->stream));
    LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LIGHTNE_OK;
  }

  // Y <- Q with Y = Q R: CholeskyQR2 with an FP64 Gram, diagonal
  // equilibration and FP64 row-block triangular solves (rsvd.hip's cholqr2_t
  // with the k x k steps on the device)
  int cholqr(float* Y, int* passes_out) {
    std::vector<double> Gh, d(k);
    int passes = 0;
    for (int pass = 0; pass < 4; ++pass) {
      LN_CHECK(gram(Y));
      LN_CHECK(host_upper(Gh));
      double dev = 0.0, fro = 0.0;
      for (int64_t j = 0; j < k; ++j)
        for (int64_t i = 0; i <= j; ++i) {
          const double g = Gh[i + j * k];
          dev = std::max(dev, std::fabs(g - (i == j ? 1.0 : 0.0)));
        }
      if (pass >= 2 && dev < 1e-5) break;
      for (int64_t i = 0; i < k; ++i) d[i] = Gh[i + i * k] > 0.0 ? 1.0 / std::sqrt(Gh[i + i * k]) : 1.0;
      for (int64_t j = 0; j < k; ++j)
        for (int64_t i = 0; i <= j; ++i) {
          const double g = d[i] * Gh[i + j * k] * d[j];
          fro += (i == j ? 1.0 : 2.0) * g * g;
        }
      fro = std::sqrt(fro);
      LN_HIP(ctx, hipMemcpyAsync(dv, d.data(), (size_t)k * 8, hipMemcpyHostToDevice, ctx->stream));
      LN_HIP(ctx, hipMemcpyAsync(G0, G, (size_t)k * k * 8, hipMemcpyDeviceToDevice, ctx->stream));
      // shifted CholeskyQR (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa
      // 2020) when the equilibrated Gram is not numerically positive definite
      const double u = std::ldexp(1.0, -53);
      double shift = 0.0;
      bool ok = false;
      for (int attempt = 0; attempt < 9 && !ok; ++attempt) {
        if (attempt > 0) {
          shift = attempt == 1 ? 11.0 * ((double)n * k + (double)k * (k + 1)) * u * fro : shift * 100.0;
          LN_HIP(ctx, hipMemcpyAsync(G, G0, (size_t)k * k * 8, hipMemcpyDeviceToDevice, ctx->stream));
        }
        hipLaunchKernelGGL(k_sym_scale, dim3(grid_blocks(k * k, 256)), dim3(256), 0, ctx->stream, G, k,
                           (const double*)dv, attempt ? shift : 0.0);
        LN_LAUNCH(ctx);
        LN_BLAS(ctx, rocsolver_dpotrf(h, rocblas_fill_upper, (rocblas_int)k, G, (rocblas_int)k, info));
        rocblas_int ih = 0;
        LN_HIP(ctx, hipMemcpyAsync(&ih, info, sizeof(ih), hipMemcpyDeviceToHost, ctx->stream));
        LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ok = ih == 0;
      }
      if (!ok) return ctx_fail(ctx, LIGHTNE_ERR_NUMERIC, "CholeskyQR: Gram matrix is not positive definite");
      // Q = (Y D) Rs^-1 per row block: column-major the block is X^T, and
      // Rs^T Q^T = (Y D)^T
      const double one = 1.0;
      for (int64_t r0 = 0; r0 < n; r0 += rb) {
        const int64_t rows = std::min(rb, n - r0);
        LN_CHECK(rows_to_f64(Y, r0, rows));
        hipLaunchKernelGGL(k_scale_cols, dim3(grid_blocks(rows * k, 256)), dim3(256), 0, ctx->stream, X64, rows, k,
                           (const double*)dv);
        LN_LAUNCH(ctx);
        LN_BLAS(ctx, rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                                   rocblas_diagonal_non_unit, (rocblas_int)k, (rocblas_int)rows, &one, G,
                                   (rocblas_int)k, X64, (rocblas_int)k));
        hipLaunchKernelGGL(k_rows_to_f32, dim3(grid_blocks(rows * k, 256)), dim3(256), 0, ctx->stream,
                           (const double*)X64, rows, k, Y + r0 * kp, kp);
        LN_LAUNCH(ctx);
      }
      ++passes;
      if (pass >= 1 && dev < 1e-1) break;
    }
    if (passes_out) *passes_out = passes;
    return LIGHTNE_OK;
  }

  // B^T B = V diag(S^2) V^T: S descending on the host, V left in G
  int eig(const float* B, std::vector<double>& S) {
    LN_CHECK(gram(B));
    LN_BLAS(ctx, rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)k, G, (rocblas_int)k,
                                  D, E, info));
    std::vector<double> w(k);
    rocblas_int ih = 0;
    LN_HIP(ctx, hipMemcpyAsync(w.data(), D, (size_t)k * 8, hipMemcpyDeviceToHost, ctx->stream));
    LN_HIP(ctx, hipMemcpyAsync(&ih, info, sizeof(ih), hipMemcpyDeviceToHost, ctx->stream));
    LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ih != 0) return ctx_fail(ctx, LIGHTNE_ERR_NUMERIC, "lightne_rsvd: eigensolver did not converge (%d)", (int)ih);
    S.resize(k);
    for (int64_t j = 0; j < k; ++j) S[j] = std::sqrt(std::max(w[k - 1 - j], 0.0));
    return LIGHTNE_OK;
  }

  // out (n x cols, row stride ldo) = B V[:, :cols] diag(S^-pw); columns whose
  // singular value is below 1e-6 S_0 (numerically zero for an FP32 B) are zero
  int apply(const float* B, const std::vector<double>& S, int64_t cols, double pw, float* out, int64_t ldo) {
    std::vector<double> f(cols, 0.0);
    for (int64_t c = 0; c < cols; ++c)
      if (S[c] > 1e-6 * S[0]) f[c] = std::pow(S[c], -pw);
    LN_HIP(ctx, hipMemcpyAsync(dv, f.data(), (size_t)cols * 8, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_build_w, dim3(grid_blocks(k * cols, 256)), dim3(256), 0, ctx->stream, (const double*)G, k,
                       cols, (const double*)dv, W);
    LN_LAUNCH(ctx);
    const double one = 1.0, zero = 0.0;
    for (int64_t r0 = 0; r0 < n; r0 += rb) {
      const int64_t rows = std::min(rb, n - r0);
      LN_CHECK(rows_to_f64(B, r0, rows));
      // column-major: O^T (cols x rows) = W^T (cols x k) X^T (k x rows)
      LN_BLAS(ctx, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)cols,
                                 (rocblas_int)rows, (rocblas_int)k, &one, W, (rocblas_int)cols, X64, (rocblas_int)k,
                                 &zero, O64, (rocblas_int)cols));
      hipLaunchKernelGGL(k_rows_to_f32, dim3(grid_blocks(rows * cols, 256)), dim3(256), 0, ctx->stream,
                         (const double*)O64, rows, cols, out + r0 * ldo, ldo);
      LN_LAUNCH(ctx);
    }
    LN_HIP(ctx, hipStreamSynchronize(ctx->stream));  // f is a local
    return LIGHTNE_OK;
  }
};

int plan_of(lightne_ctx* ctx, lightne_csr* A) {
  if (A->plan) return LIGHTNE_OK;
  auto* P = new SpmmPlan();
  const int rc = spmm_plan_build(ctx, A->row_ptr, A->n, *P);
  if (rc != LIGHTNE_OK) {
    spmm_plan_free(*P);
    delete P;
    return rc;
  }
  A->plan = P;
  return LIGHTNE_OK;
}

// Y (n x kp) = A X (n x kp) in column slabs of at most SLAB columns
int mm_slabs(lightne_ctx* ctx, lightne_csr* A, const void* X, int xt, int64_t kp, float* Y) {
  LN_CHECK(plan_of(ctx, A));
  const size_t esz = xt == PANEL_BF16 ? 2 : 4;
  for (int64_t c0 = 0; c0 < kp; c0 += SLAB) {
    SpmmArgs a;
    a.n = A->n;
    a.k = std::min(SLAB, kp - c0);
    a.ldx = kp;
    a.ldo = kp;
    a.rp = A->row_ptr;
    a.col = A->col;
    a.val = A->val;
    a.impl_mode = IMPL_NONE;
    a.X = (const char*)X + c0 * esz;
    a.out = Y + c0;
    a.alpha = 1.f;
    LN_CHECK(spmm_run(ctx, *A->plan, a, xt, EPI_STORE_F32, false));
  }
  return LIGHTNE_OK;
}

}  // namespace

void wide_handle_free(lightne_ctx* ctx) {
  if (ctx && ctx->blas) {
    (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
  }
}

size_t rsvd_wide_front(int64_t n, int64_t k, int precision) {
  const int64_t kp = (k + 15) / 16 * 16;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const bool bf = precision == LIGHTNE_BF16 || precision == LIGHTNE_BF16_POWER;
  return 2 * al((size_t)n * kp * 4) + (bf ? al((size_t)n * kp * 2) : 0);
}

int rsvd_wide(lightne_ctx* ctx, lightne_csr* A, const lightne_rsvd_params* p, int64_t cols, double pw, int normalize,
              float* out, int64_t ldo, std::vector<double>& S, lightne_rsvd_stats* stp) {
  const int64_t n = A->n, k = p->rank;
  if (p->precision != LIGHTNE_FP32 && p->precision != LIGHTNE_BF16 && p->precision != LIGHTNE_BF16_POWER)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG,
                    "lightne_rsvd: rank %lld takes the wide path, which has FP32 and BF16 panels only (the FP64 "
                    "forms need a rank that is a multiple of 16 up to 256)", (long long)k);
  if (k <= 0 || k > n) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_rsvd: rank must be in [1, n]");
  if (k >= ((int64_t)1 << 15)) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_rsvd: rank above 32767");
  if (p->sparse_project && k > 256)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_rsvd: -sparse_project supports ranks up to 256");
  lightne_rsvd_stats& st = *stp;
  st = lightne_rsvd_stats{};
  hipStream_t s = ctx->stream;
  const bool bf = p->precision == LIGHTNE_BF16 || p->precision == LIGHTNE_BF16_POWER;
  const bool bf_power = p->precision == LIGHTNE_BF16_POWER;
  const int q = std::max(1, (int)p->power_iters);
  const int64_t kp = (k + 15) / 16 * 16;
  float *P1 = nullptr, *P2 = nullptr;  // Y / Q; Omega, then B
  __hip_bfloat16* PB = nullptr;
  {
    WsCarve wc;
    wc.add((void**)&P1, (size_t)n * kp * 4);
    wc.add((void**)&P2, (size_t)n * kp * 4);
    if (bf) wc.add((void**)&PB, (size_t)n * kp * 2);
    LN_CHECK(wc.commit(ctx));
  }
  Wide w{ctx};
  LN_CHECK(w.init(n, k, std::max(cols, k)));
  hipEvent_t* e = ctx->ev;
  auto lap = [&](hipEvent_t a, hipEvent_t b) { return ev_seconds(a, b); };
  double t_spmm = 0, t_qr = 0;
  LN_HIP(ctx, hipEventRecord(e[0], s));
  if (kp != k) {  // the padding columns stay zero throughout
    LN_HIP(ctx, hipMemset2DAsync(P1 + k, kp * 4, 0, (kp - k) * 4, n, s));
    LN_HIP(ctx, hipMemset2DAsync(P2 + k, kp * 4, 0, (kp - k) * 4, n, s));
  }
  if (!p->sparse_project) {
    LN_CHECK(gaussian_fill_ld(ctx, p->seed, 0u, n, k, kp, P2));
    if (bf) LN_CHECK(f32_to_bf16(ctx, P2, PB, n * kp));
  }
  LN_HIP(ctx, hipEventRecord(e[1], s));
  LN_CHECK(plan_of(ctx, A));
  LN_HIP(ctx, hipEventRecord(e[2], s));
  if (p->sparse_project) {
    // the sketch product n x k contiguous in P2, then into P1's rows
    LN_CHECK(sparse_project_sketch(ctx, A->row_ptr, A->col, A->val, n, k, p->seed, p->sparse_project_s, P2, false));
    LN_HIP(ctx, hipMemcpy2DAsync(P1, kp * 4, P2, k * 4, k * 4, n, hipMemcpyDeviceToDevice, s));
  } else {
    LN_CHECK(mm_slabs(ctx, A, bf ? (const void*)PB : (const void*)P2, bf ? PANEL_BF16 : PANEL_F32, kp, P1));
  }
  LN_HIP(ctx, hipEventRecord(e[3], s));
  LN_HIP(ctx, hipEventSynchronize(e[3]));
  st.t_sketch = lap(e[0], e[1]);
  t_spmm += lap(e[2], e[3]);
  st.t_spmm_sketch = lap(e[2], e[3]);
  st.spmm_calls = 1;
  if (p->random_project_only) {
    LN_HIP(ctx, hipMemcpy2DAsync(out, ldo * 4, P1, kp * 4, k * 4, n, hipMemcpyDeviceToDevice, s));
    LN_HIP(ctx, hipStreamSynchronize(s));
    S.clear();
    st.t_spmm = t_spmm;
    st.t_total = lap(e[0], e[3]);
    return LIGHTNE_OK;
  }
  int passes = 0;
  for (int it = 0; it < q; ++it) {
    LN_HIP(ctx, hipEventRecord(e[4], s));
    LN_CHECK(w.cholqr(P1, &passes));
    st.qr_calls++;
    LN_HIP(ctx, hipEventRecord(e[5], s));
    const bool gather_bf = bf_power && (it + 1 < q);
    if (gather_bf) LN_CHECK(f32_to_bf16(ctx, P1, PB, n * kp));
    LN_CHECK(mm_slabs(ctx, A, gather_bf ? (const void*)PB : (const void*)P1, gather_bf ? PANEL_BF16 : PANEL_F32, kp,
                      P2));
    st.spmm_calls++;
    LN_HIP(ctx, hipEventRecord(e[6], s));
    LN_HIP(ctx, hipEventSynchronize(e[6]));
    t_qr += lap(e[4], e[5]);
    t_spmm += lap(e[5], e[6]);
    if (it + 1 < q) std::swap(P1, P2);
  }
  LN_HIP(ctx, hipEventRecord(e[7], s));
  auto t0 = std::chrono::steady_clock::now();
  LN_CHECK(w.eig(P2, S));
  st.t_small_svd = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  LN_CHECK(w.apply(P2, S, cols, pw, out, ldo));
  if (normalize) {  // in place: each row is read whole before it is written
    if (ldo != cols) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "rsvd_wide: row normalisation needs ldo == cols");
    LN_CHECK(u_sigma_root_dev(ctx, out, nullptr, n, cols, cols, 1, out));
  }
  LN_HIP(ctx, hipEventRecord(e[8], s));
  LN_HIP(ctx, hipEventSynchronize(e[8]));
  st.t_spmm = t_spmm;
  st.t_qr = t_qr;
  st.t_gemm = std::max(0.0, lap(e[7], e[8]) - st.t_small_svd);
  st.t_total = lap(e[0], e[8]);
  return LIGHTNE_OK;
}

// get_embedding_dense (SpectralPropagation.h:16-43) for ranks and dims outside
// the 16-wide tiling: the thin SVD of mat (n x rank, row stride ld) from its
// FP64 Gram on the device, emb = rownorm(mat V[:, :dim] S^-1/2)
int embed_dense_wide(lightne_ctx* ctx, const float* mat, int64_t n, int64_t rank, int64_t ld, int64_t dim,
                     float* emb_out, float* S_host) {
  if (rank <= 0 || dim <= 0 || dim > rank || n < rank || ld < rank || rank >= ((int64_t)1 << 15))
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "embed_dense: need 1 <= dim <= rank <= n and rank < 32768");
  Wide w{ctx};
  LN_CHECK(w.init(n, rank, dim));
  w.kp = ld;
  std::vector<double> S;
  LN_CHECK(w.eig(mat, S));
  LN_CHECK(w.apply(mat, S, dim, 0.5, emb_out, dim));
  LN_CHECK(u_sigma_root_dev(ctx, emb_out, nullptr, n, dim, dim, 1, emb_out));  // rows read whole before written
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (S_host)
    for (int64_t j = 0; j < rank; ++j) S_host[j] = (float)S[j];
  return LIGHTNE_OK;
}
