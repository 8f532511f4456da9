"""The MKL form of the CPU baseline (oracle/mkl_ref.py) computes what the
reference's MKL calls compute: each call against scipy/numpy, the randomized
SVD against the FP64 oracle, the recurrence against the FP64 propagation."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import mkl_ref as K
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.skipif(K.find() is None, reason="Intel MKL runtime not installed")


def test_mkl_calls_match_numpy():
    rng = np.random.default_rng(0)
    A = sp.random(500, 500, density=0.02, random_state=2, format="csr", dtype=np.float32)
    M = K.SparseCSR(A.indptr, A.indices, A.data, hint_k=24, hint_calls=2)
    X = rng.standard_normal((500, 24)).astype(np.float32)
    Y0 = rng.standard_normal((500, 24)).astype(np.float32)
    ref = A.astype(np.float64) @ X
    np.testing.assert_allclose(M.mm(X), ref, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(M.mm(X, 0.5, -1.0, Y0.copy()), 0.5 * ref - Y0, rtol=1e-4, atol=1e-5)
    Q = X.copy()
    K.gram_schmidt(Q)
    assert np.abs(Q.T.astype(np.float64) @ Q - np.eye(24)).max() < 1e-5
    assert H.principal_angles(Q, X) < 1e-5
    B = rng.standard_normal((24, 7)).astype(np.float32)
    np.testing.assert_allclose(K.sgemm(X, B), X.astype(np.float64) @ B, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(K.sgemm(X, Y0, trans_a=True), X.T.astype(np.float64) @ Y0, rtol=1e-4, atol=1e-3)
    U, S = K.gesvd_left(X.copy())
    np.testing.assert_allclose(S, np.linalg.svd(X.astype(np.float64), compute_uv=False), rtol=1e-5)
    y = Y0.copy()
    K.saxpy(-2.0, X, y)
    np.testing.assert_allclose(y, Y0 - 2 * X, rtol=1e-6)


def test_mkl_redsvd_matches_fp64_oracle():
    off, edg = H.golden_graph("cl2000")
    keys, cnt, wgh, _ = O.sample(off, edg, 3, 10, H.c1_sample_ratio(len(off) - 1, int(off[-1]), 10, 20.0), 0)
    rp, col, val, _ = O.trunc_log_csr(off, keys, wgh)
    n, k, d = len(rp) - 1, 96, 64
    A = K.SparseCSR(rp, col, val, hint_k=k, hint_calls=2)
    Om = O.gaussian(777, 0, n * k).reshape(n, k)
    Pm = O.gaussian(777, 1, k * k).reshape(k, k)
    U, S = K.redsvd(A, Om, Pm)
    U64, S64 = O.rsvd(rp, col, val, k, seed=777)
    np.testing.assert_allclose(S[:d], S64[:d], rtol=1e-4)
    assert H.principal_angles(U[:, :d], U64[:, :d]) < 1e-3


def test_mkl_propagation_matches_fp64_oracle():
    off, edg = H.golden_graph("cl2000")
    n = len(off) - 1
    a = np.random.default_rng(1).standard_normal((n, 16)).astype(np.float32)
    M = K.prop_operator(off, edg, 0.2)
    Mx = O.prop_matrix(off, edg, float(np.float32(0.2)))
    np.testing.assert_allclose(M.mm(a), Mx @ a.astype(np.float64), rtol=1e-5, atol=1e-5)
    y = K.chebyshev(M, a, O.cheb_coeffs(10, float(np.float32(0.5))))
    y64, z64 = O.propagate(off, edg, a.astype(np.float64))
    assert np.abs(y - y64).max() / np.abs(y64).max() < 1e-5
    z = K.aggregate_operator(off, edg).mm(y)
    assert np.abs(z - z64).max() / np.abs(z64).max() < 1e-5
