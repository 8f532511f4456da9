"""Stage 3 (Chebyshev propagation + dense tall SVD) on the device vs the FP64 oracle."""
import numpy as np
import pytest

import lightne_amd as L
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _graph():
    return L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)


@pytest.mark.parametrize("order", [1, 2, 3, 10])
def test_propagate_matches_fp64_oracle(gpu, order):
    import torch

    hg = _graph()
    d = 128
    rng = np.random.default_rng(order)
    a = rng.standard_normal((hg.n, d)).astype(np.float32)
    at = torch.from_numpy(a).cuda()
    y, z, st = L.propagate(at, hg.to_device(), order=order, mu=0.2, theta=0.5, ctx=gpu)
    np.testing.assert_array_equal(at.cpu().numpy(), a)  # input not modified
    y64, z64 = O.propagate(hg.offsets, hg.edges, a, order=order)
    ey = np.abs(y.cpu().numpy() - y64).max() / np.abs(y64).max()
    ez = np.abs(z.cpu().numpy() - z64).max() / np.abs(z64).max()
    print(f"order={order} rel err y={ey:.2e} z={ez:.2e} spmm calls={st['spmm_calls']}")
    assert ey < 1e-5 and ez < 1e-5


def test_propagate_bf16_error(gpu):
    import torch

    hg = _graph()
    a = np.random.default_rng(0).standard_normal((hg.n, 128)).astype(np.float32)
    y, z, _ = L.propagate(torch.from_numpy(a).cuda(), hg.to_device(), precision=L.BF16, ctx=gpu)
    y64, z64 = O.propagate(hg.offsets, hg.edges, a)
    ey = np.abs(y.cpu().numpy() - y64).max() / np.abs(y64).max()
    ez = np.abs(z.cpu().numpy() - z64).max() / np.abs(z64).max()
    print(f"BF16 rel err y={ey:.2e} z={ez:.2e}")
    assert ey < 1e-3 and ez < 1e-3


@pytest.mark.parametrize("n,r,d", [(5000, 128, 128), (3000, 144, 128), (4096, 64, 32)])
def test_embed_dense_matches_numpy_svd(gpu, n, r, d):
    import torch

    rng = np.random.default_rng(n)
    # spread singular values so that singular vectors are well defined
    U = np.linalg.qr(rng.standard_normal((n, r)))[0]
    V = np.linalg.qr(rng.standard_normal((r, r)))[0]
    s = np.geomspace(100.0, 1.0, r)
    M = ((U * s) @ V.T).astype(np.float32)
    emb, S = L.get_embedding_dense(torch.from_numpy(M).cuda(), n, r, d, ctx=gpu)
    ref, S64 = O.embed_dense(M, d)
    np.testing.assert_allclose(S, S64, rtol=1e-4)
    e = H.sign_align(emb.cpu().numpy(), ref)
    assert np.abs(e - ref).max() < 2e-3


def test_chebyshev_expansion_end_to_end(gpu):
    import torch

    hg = _graph()
    d = 128
    a = np.random.default_rng(9).standard_normal((hg.n, d)).astype(np.float32)
    g = hg.to_device()
    emb, st = L.chebyshev_expansion(torch.from_numpy(a).cuda(), g, d, d, ctx=gpu)
    _, z, _ = L.propagate(torch.from_numpy(a).cuda(), g, ctx=gpu)
    emb2, _ = L.get_embedding_dense(z, hg.n, d, d, ctx=gpu)
    np.testing.assert_array_equal(emb.cpu().numpy(), emb2.cpu().numpy())
    # in place: the output may be the input itself
    b = torch.from_numpy(a).cuda()
    emb3, _ = L.chebyshev_expansion(b, g, d, d, out=b, ctx=gpu)
    assert emb3.data_ptr() == b.data_ptr()
    np.testing.assert_array_equal(emb3.cpu().numpy(), emb.cpu().numpy())
    _, z64 = O.propagate(hg.offsets, hg.edges, a)
    ref, S64 = O.embed_dense(z64, d)
    e = H.sign_align(emb.cpu().numpy(), ref)
    # rows are unit vectors; columns well separated for a random input
    gaps = S64[:-1] / S64[1:]
    good = np.where((np.r_[np.inf, gaps] > 1.001) & (np.r_[gaps, np.inf] > 1.001))[0]
    err = np.abs(e[:, good] - ref[:, good]).max()
    print(f"embedding max abs err over {len(good)} well-separated columns: {err:.2e}")
    assert err < 5e-3
    assert np.allclose(np.linalg.norm(emb.cpu().numpy(), axis=1), 1.0, atol=1e-5)
