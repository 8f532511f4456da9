// Host FP64 linear algebra on k x k matrices (k <= 256): the pieces of the
// randomized SVD that are too small for the device.
//   * chol_upper / tri_upper_inverse: CholeskyQR2, replacing the Householder
//     QR of LAPACKE_?geqrf + ?orgqr (LightNE/mklredsvd/util.cc:41-79).
//   * sym_eig / svd_left_via_eig: symmetric eigensolver used for the k x k
//     SVDs that replace LAPACKE_?gesvd on C (mklredsvd.cc:352-377) and on the
//     R factor of the tall panel of get_embedding_dense
//     (SpectralPropagation.h:16-43).
// Written from the textbook algorithms (Golub & Van Loan, 4th ed., 4.2.5 and
// 8.3); singular vectors are determined up to sign.  Host-only C++ (built by
// the host compiler); sym_eig carries an AVX2/FMA clone chosen at load time.
#include "linalg_host.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

// Right-looking: row j of R from the updated trailing block, then that block
// (upper triangle, row by row) loses the outer product of row j; every inner
// loop runs along a contiguous row.
bool chol_upper(const double* G, int k, double* R_out) {
  std::vector<double> Wbuf(G, G + (size_t)k * k);
  double* __restrict W = Wbuf.data();
  double* __restrict R = R_out;
  std::fill(R, R + (size_t)k * k, 0.0);
  for (int j = 0; j < k; ++j) {
    const double d = W[(size_t)j * k + j];
    if (!(d > 0.0) || !std::isfinite(d)) return false;
    const double rjj = std::sqrt(d);
    double* __restrict rj = R + (size_t)j * k;
    const double* __restrict wj = W + (size_t)j * k;
    rj[j] = rjj;
    for (int i = j + 1; i < k; ++i) rj[i] = wj[i] / rjj;
    for (int i = j + 1; i < k; ++i) {
      const double f = rj[i];
      double* __restrict wi = W + (size_t)i * k;
      for (int l = i; l < k; ++l) wi[l] -= f * rj[l];
    }
  }
  return true;
}

// Rows from the bottom up: row i of R^-1 is (e_i - sum_{p>i} R[i][p] row_p) /
// R[i][i], an update along contiguous rows.
void tri_upper_inverse(const double* R_in, int k, double* Rinv_out) {
  const double* __restrict R = R_in;
  double* __restrict X = Rinv_out;
  std::fill(X, X + (size_t)k * k, 0.0);
  for (int i = k - 1; i >= 0; --i) {
    double* __restrict xi = X + (size_t)i * k;
    xi[i] = 1.0;
    for (int p = i + 1; p < k; ++p) {
      const double f = R[(size_t)i * k + p];
      const double* __restrict xp = X + (size_t)p * k;
      for (int l = p; l < k; ++l) xi[l] -= f * xp[l];
    }
    const double inv = 1.0 / R[(size_t)i * k + i];
    for (int l = i; l < k; ++l) xi[l] *= inv;
  }
}

void matmul_host(const double* A, const double* B, int m, int k, int n, double* C) {
  std::vector<double> tmp((size_t)m * n, 0.0);
  for (int i = 0; i < m; ++i)
    for (int p = 0; p < k; ++p) {
      const double a = A[(size_t)i * k + p];
      if (a == 0.0) continue;
      for (int j = 0; j < n; ++j) tmp[(size_t)i * n + j] += a * B[(size_t)p * n + j];
    }
  std::copy(tmp.begin(), tmp.end(), C);
}

// x . y with four partial sums (independent adds the compiler can vectorize;
// the summation order is fixed, so results are deterministic)
static inline __attribute__((always_inline)) double dot4(const double* __restrict x, const double* __restrict y,
                                                         int m) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = 0;
  for (; i + 4 <= m; i += 4) {
    s0 += x[i] * y[i];
    s1 += x[i + 1] * y[i + 1];
    s2 += x[i + 2] * y[i + 2];
    s3 += x[i + 3] * y[i + 3];
  }
  for (; i < m; ++i) s0 += x[i] * y[i];
  return (s0 + s1) + (s2 + s3);
}

// sqrt(x^2 + z^2) without std::hypot's cost; hypot only where the squares
// leave the normal range
static inline __attribute__((always_inline)) double norm2(double x, double z) {
  const double q = x * x + z * z;
  if (q > 1e-290 && q < 1e290) return std::sqrt(q);
  return std::hypot(x, z);
}

// ---------------------------------------------------------------------------
// Symmetric eigendecomposition A = Z diag(w) Z^T (FP64), eigenvalues
// descending.  Householder tridiagonalisation, the reflectors accumulated
// afterwards from the last to the first (each touches only the trailing
// block: 4/3 n^3 flops instead of 2 n^3), then implicit symmetric QR steps
// with Wilkinson shifts on the tridiagonal, the rotations applied to
// contiguous columns of Q (Golub & Van Loan, 4th ed., Algorithms 5.1.5, 8.3.1 and
// 8.3.3).  Z is returned row-major (column j is the eigenvector of w[j]).
// ---------------------------------------------------------------------------
__attribute__((target_clones("arch=haswell", "default"))) bool sym_eig(const double* A_in, int n, double* w,
                                                                        double* Z_out) {
  // column-major working arrays: A (symmetric), the reflectors (column k
  // holds v_k in rows k+1..n-1) and Q
  std::vector<double> Abuf(A_in, A_in + (size_t)n * n), Vbuf((size_t)n * n, 0.0), Qbuf((size_t)n * n, 0.0);
  std::vector<double> tau_k(n, 0.0), pbuf(n), wbuf(n);
  double* __restrict A = Abuf.data();
  double* __restrict Vr = Vbuf.data();
  double* __restrict Q = Qbuf.data();
  double* __restrict p = pbuf.data();
  double* __restrict wv = wbuf.data();
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;  // length of x = A[k+1:, k]
    const double* __restrict x = A + (size_t)k * n + (k + 1);
    const double alpha = std::sqrt(dot4(x, x, m));
    if (alpha == 0.0) continue;
    const double beta = x[0] >= 0 ? -alpha : alpha;  // H x = beta e1
    double* __restrict v = Vr + (size_t)k * n + (k + 1);
    for (int i = 0; i < m; ++i) v[i] = x[i];
    v[0] -= beta;
    const double vv = dot4(v, v, m);
    if (vv == 0.0) continue;
    const double tau = 2.0 / vv;
    tau_k[k] = tau;
    // p = tau * A22 v ; K = tau/2 v^T p ; w = p - K v ; A22 -= v w^T + w v^T
    for (int i = 0; i < m; ++i) p[i] = 0.0;
    for (int j = 0; j < m; ++j) {
      const double vj = v[j];
      const double* __restrict col = A + (size_t)(k + 1 + j) * n + (k + 1);
      for (int i = 0; i < m; ++i) p[i] += col[i] * vj;
    }
    for (int i = 0; i < m; ++i) p[i] *= tau;
    const double K = 0.5 * tau * dot4(v, p, m);
    for (int i = 0; i < m; ++i) wv[i] = p[i] - K * v[i];
    for (int j = 0; j < m; ++j) {
      double* __restrict col = A + (size_t)(k + 1 + j) * n + (k + 1);
      const double vj = v[j], wj = wv[j];
      for (int i = 0; i < m; ++i) col[i] -= v[i] * wj + wv[i] * vj;
    }
    A[(size_t)k * n + (k + 1)] = beta;  // the subdiagonal entry; the rest of column k is not read again
  }
  // Q = H_0 H_1 ... H_{n-3}, built from the identity by H_k Q for k = n-3 .. 0:
  // H_{k+1} ... H_{n-3} is the identity outside rows/columns k+2.., so H_k
  // acts on the block [k+1:, k+1:] only
  for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
  for (int k = n - 3; k >= 0; --k) {
    const double tau = tau_k[k];
    if (tau == 0.0) continue;
    const int m = n - k - 1;
    const double* __restrict v = Vr + (size_t)k * n + (k + 1);
    for (int j = k + 1; j < n; ++j) {
      double* __restrict q = Q + (size_t)j * n + (k + 1);
      const double d = tau * dot4(v, q, m);
      for (int i = 0; i < m; ++i) q[i] -= d * v[i];
    }
  }
  std::vector<double> a(n), b(n, 0.0);
  for (int i = 0; i < n; ++i) a[i] = A[(size_t)i * n + i];
  for (int i = 0; i + 1 < n; ++i) b[i] = A[(size_t)i * n + (i + 1)];
  // implicit QR with Wilkinson shift; Q columns are contiguous (column-major)
  const double eps = std::numeric_limits<double>::epsilon();
  int h = n - 1, iters = 0;
  while (h > 0) {
    if (std::fabs(b[h - 1]) <= eps * (std::fabs(a[h - 1]) + std::fabs(a[h]))) {
      b[h - 1] = 0.0;
      --h;
      continue;
    }
    int l = h - 1;
    while (l > 0 && std::fabs(b[l - 1]) > eps * (std::fabs(a[l - 1]) + std::fabs(a[l]))) --l;
    if (++iters > 60 * n) return false;
    const double d = 0.5 * (a[h - 1] - a[h]);
    const double bb = b[h - 1] * b[h - 1];
    const double mu = a[h] - bb / (d + (d >= 0 ? 1.0 : -1.0) * norm2(d, b[h - 1]));
    double x = a[l] - mu, z = b[l];
    for (int k = l; k < h; ++k) {
      const double r = norm2(x, z);
      const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : z / r;
      if (k > l) b[k - 1] = r;  // the bulge is annihilated into b[k-1]
      const double ak = a[k], ak1 = a[k + 1], bk = b[k];
      a[k] = c * c * ak + 2.0 * c * s * bk + s * s * ak1;
      a[k + 1] = s * s * ak - 2.0 * c * s * bk + c * c * ak1;
      b[k] = c * s * (ak1 - ak) + (c * c - s * s) * bk;
      if (k + 1 < h) {
        x = b[k];
        z = s * b[k + 1];  // new bulge at (k, k+2)
        b[k + 1] *= c;
      }
      double* __restrict qk = Q + (size_t)k * n;
      double* __restrict qk1 = Q + (size_t)(k + 1) * n;
      for (int i = 0; i < n; ++i) {
        const double u = qk[i], t = qk1[i];
        qk[i] = c * u + s * t;
        qk1[i] = -s * u + c * t;
      }
    }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int i, int j) { return a[i] > a[j]; });
  for (int jj = 0; jj < n; ++jj) {
    const int j = ord[jj];
    w[jj] = a[j];
    for (int i = 0; i < n; ++i) Z_out[(size_t)i * n + jj] = Q[(size_t)j * n + i];
  }
  return true;
}

// Left singular vectors and singular values of a square A (n x n) through the
// eigendecomposition of A A^T: A A^T = U diag(S^2) U^T.  Relative accuracy of
// S_i is about eps (S_0/S_i)^2, ample for the FP32 panels it serves.
bool svd_left_via_eig(const double* A, int n, double* U, double* S) {
  std::vector<double> G((size_t)n * n, 0.0), w(n);
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      double s = 0.0;
      const double* ai = A + (size_t)i * n;
      const double* aj = A + (size_t)j * n;
      for (int k = 0; k < n; ++k) s += ai[k] * aj[k];
      G[(size_t)i * n + j] = s;
      G[(size_t)j * n + i] = s;
    }
  if (!sym_eig(G.data(), n, w.data(), U)) return false;
  for (int j = 0; j < n; ++j) S[j] = std::sqrt(std::max(w[j], 0.0));
  return true;
}
