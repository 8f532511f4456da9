// Host FP64 linear algebra on k x k matrices (k <= 256): the pieces of the
// randomized SVD that are too small for the device.
//   * chol_upper / tri_upper_inverse: CholeskyQR2, replacing the Householder
//     QR of LAPACKE_?geqrf + ?orgqr (LightNE/mklredsvd/util.cc:41-79).
//   * jacobi_svd: one-sided (Hestenes) Jacobi SVD, replacing LAPACKE_?gesvd
//     on C (mklredsvd.cc:352-377) and on the R factor of the tall panel of
//     get_embedding_dense (SpectralPropagation.h:16-43).
// Written from the textbook algorithms (Golub & Van Loan, 4th ed., 4.2.5 and
// 8.6.3); singular vectors are determined up to sign.
#include "lightne_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

bool chol_upper(const double* G, int k, double* R) {
  std::fill(R, R + (size_t)k * k, 0.0);
  for (int j = 0; j < k; ++j) {
    double d = G[(size_t)j * k + j];
    for (int p = 0; p < j; ++p) d -= R[(size_t)p * k + j] * R[(size_t)p * k + j];
    if (!(d > 0.0) || !std::isfinite(d)) return false;
    const double rjj = std::sqrt(d);
    R[(size_t)j * k + j] = rjj;
    for (int i = j + 1; i < k; ++i) {
      double s = G[(size_t)j * k + i];
      for (int p = 0; p < j; ++p) s -= R[(size_t)p * k + j] * R[(size_t)p * k + i];
      R[(size_t)j * k + i] = s / rjj;
    }
  }
  return true;
}

void tri_upper_inverse(const double* R, int k, double* Rinv) {
  std::fill(Rinv, Rinv + (size_t)k * k, 0.0);
  for (int j = k - 1; j >= 0; --j) {
    Rinv[(size_t)j * k + j] = 1.0 / R[(size_t)j * k + j];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int p = i + 1; p <= j; ++p) s += R[(size_t)i * k + p] * Rinv[(size_t)p * k + j];
      Rinv[(size_t)i * k + j] = -s / R[(size_t)i * k + i];
    }
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

void jacobi_svd(const double* A, int m, int n, double* U, double* S, double* V) {
  // work on columns: W = A (column-major copy for cache-friendly rotations)
  std::vector<double> W((size_t)m * n), Vw((size_t)n * n, 0.0);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) W[(size_t)j * m + i] = A[(size_t)i * n + j];
  for (int j = 0; j < n; ++j) Vw[(size_t)j * n + j] = 1.0;
  const double eps = 1e-15;
  for (int sweep = 0; sweep < 80; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double* wp = &W[(size_t)p * m];
        double* wq = &W[(size_t)q * m];
        double alpha = 0.0, beta = 0.0, gamma = 0.0;
        for (int i = 0; i < m; ++i) {
          alpha += wp[i] * wp[i];
          beta += wq[i] * wq[i];
          gamma += wp[i] * wq[i];
        }
        if (gamma == 0.0) continue;
        const double rel = std::fabs(gamma) / std::sqrt(alpha * beta);
        if (!(rel > eps)) continue;
        off = std::max(off, rel);
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < m; ++i) {
          const double x = wp[i], y = wq[i];
          wp[i] = c * x - s * y;
          wq[i] = s * x + c * y;
        }
        double* vp = &Vw[(size_t)p * n];
        double* vq = &Vw[(size_t)q * n];
        for (int i = 0; i < n; ++i) {
          const double x = vp[i], y = vq[i];
          vp[i] = c * x - s * y;
          vq[i] = s * x + c * y;
        }
      }
    }
    if (off <= eps) break;
  }
  std::vector<double> sv(n);
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += W[(size_t)j * m + i] * W[(size_t)j * m + i];
    sv[j] = std::sqrt(s);
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return sv[a] > sv[b]; });
  for (int jj = 0; jj < n; ++jj) {
    const int j = ord[jj];
    if (S) S[jj] = sv[j];
    if (U) {
      const double inv = sv[j] > 0.0 ? 1.0 / sv[j] : 0.0;
      for (int i = 0; i < m; ++i) U[(size_t)i * n + jj] = W[(size_t)j * m + i] * inv;
    }
    if (V)
      for (int i = 0; i < n; ++i) V[(size_t)i * n + jj] = Vw[(size_t)j * n + i];
  }
}
