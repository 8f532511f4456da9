"""Dense panel kernels (FP64-MFMA Gram, FP32-MFMA panel GEMM, CholeskyQR2) vs
numpy in FP64."""
import numpy as np
import pytest

import lightne_amd as L

pytestmark = pytest.mark.gpu


def _panel(n, k, seed, scale=None):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, k)).astype(np.float32)
    if scale is not None:
        X *= scale[None, :].astype(np.float32)
    return X


@pytest.mark.parametrize("n,k1,k2", [(16, 16, 16), (1000, 64, 64), (12345, 144, 144), (5000, 256, 256),
                                     (777, 144, 128), (70000, 128, 128)])
def test_gram_matches_numpy(gpu, n, k1, k2):
    import torch

    X = _panel(n, k1, 1)
    Y = _panel(n, k2, 2)
    Xt, Yt = torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda()
    G = L.dense_gram(Xt, Yt, ctx=gpu)
    ref = X.astype(np.float64).T @ Y.astype(np.float64)
    np.testing.assert_allclose(G, ref, rtol=1e-12, atol=1e-9)
    if k1 == k2:
        Gs = L.dense_gram(Xt, ctx=gpu)
        np.testing.assert_allclose(Gs, X.astype(np.float64).T @ X.astype(np.float64), rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("n,k1,k2", [(16, 16, 16), (1000, 64, 64), (12345, 144, 144), (5000, 256, 256),
                                     (777, 144, 128), (3001, 128, 32)])
def test_panel_gemm_matches_numpy(gpu, n, k1, k2):
    import torch

    X = _panel(n, k1, 3)
    W = np.random.default_rng(4).standard_normal((k1, k2))
    out = L.dense_panel_gemm(torch.from_numpy(X).cuda(), W, ctx=gpu).cpu().numpy()
    ref = X.astype(np.float64) @ W.astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(out, ref, rtol=1e-4, atol=1e-4)
    if k1 * (k2 + 4) * 4 <= 150 * 1024:  # row normalisation needs W in one LDS block
        outn = L.dense_panel_gemm(torch.from_numpy(X).cuda(), W, normalize_rows=True, ctx=gpu).cpu().numpy()
        refn = ref / np.linalg.norm(ref, axis=1, keepdims=True)
        np.testing.assert_allclose(outn, refn, rtol=1e-4, atol=1e-5)


def test_panel_gemm_in_place(gpu):
    import torch

    X = _panel(4099, 144, 5)
    W = np.random.default_rng(6).standard_normal((144, 144))
    Xt = torch.from_numpy(X).cuda()
    L.dense_panel_gemm(Xt, W, out=Xt, ctx=gpu)
    np.testing.assert_allclose(Xt.cpu().numpy(), X.astype(np.float64) @ W.astype(np.float32), rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("n,k,cond", [(5000, 64, 1e2), (20000, 144, 1e5), (4096, 256, 1e3), (1000, 128, 1e7)])
def test_cholqr_orthonormal_and_reconstructs(gpu, n, k, cond):
    import torch

    rng = np.random.default_rng(int(cond))
    U = np.linalg.qr(rng.standard_normal((n, k)))[0]
    V = np.linalg.qr(rng.standard_normal((k, k)))[0]
    s = np.geomspace(1.0, 1.0 / cond, k)
    Y = ((U * s) @ V.T * 7.0).astype(np.float32)
    Yt = torch.from_numpy(Y.copy()).cuda()
    Q, R, passes = L.dense_cholqr(Yt, ctx=gpu)
    Qh = Q.cpu().numpy().astype(np.float64)
    orth = np.abs(Qh.T @ Qh - np.eye(k)).max()
    rec = np.abs(Qh @ R - Y).max() / np.abs(Y).max()
    print(f"n={n} k={k} cond={cond:g} passes={passes} orth={orth:.2e} rec={rec:.2e}")
    assert orth < 1e-5
    assert rec < 1e-5 * max(1.0, cond / 1e3)
    assert np.allclose(np.tril(R, -1), 0)
