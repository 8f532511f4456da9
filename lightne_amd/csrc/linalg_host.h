// Host FP64 small linear algebra (linalg_host.cpp): declarations only, so the
// file builds with the host C++ compiler.
#pragma once

// upper Cholesky R^T R = G (k x k row-major, in place into R).  Returns false
// if G is not numerically positive definite.
bool chol_upper(const double* G, int k, double* R);
// inverse of an upper-triangular matrix (row-major)
void tri_upper_inverse(const double* R, int k, double* Rinv);
// symmetric eigendecomposition, eigenvalues descending, Z row-major n x n
bool sym_eig(const double* A, int n, double* w, double* Z);
// left singular vectors / values of a square matrix via eig(A A^T)
bool svd_left_via_eig(const double* A, int n, double* U, double* S);
// C = A (m x k) * B (k x n), row-major
void matmul_host(const double* A, const double* B, int m, int k, int n, double* C);
