"""Edge cases and error paths of the C ABI on the device: isolated vertices,
empty graphs, bad shapes, several contexts, errors leave the context usable."""
import numpy as np
import pytest

import lightne_amd as L
from lightne_amd import _lib
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _with_isolated(hg, extra):
    """hg plus `extra` isolated vertices appended."""
    off = np.concatenate([hg.offsets, np.full(extra, hg.offsets[-1], np.int64)])
    return L.HostGraph(off, hg.edges)


def test_isolated_vertices_give_zero_rows(gpu):
    hg = _with_isolated(L.chung_lu(3000, 2.1, 12.0, 300.0, 5), 37)
    g = hg.to_device()
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 20.0)
    emb, S, _ = L.NetSMF(g, walk_len=10, rank=64, dim=32, sample_ratio=ratio, seed=2, ctx=gpu)
    pro, _ = L.chebyshev_expansion(emb, g, 32, 32, ctx=gpu)
    e, p = emb.cpu().numpy(), pro.cpu().numpy()
    assert np.isfinite(e).all() and np.isfinite(p).all()
    assert np.all(e[-37:] == 0) and np.all(p[-37:] == 0)
    nrm = np.linalg.norm(p[:-37], axis=1)
    assert np.allclose(nrm[nrm > 1e-4], 1.0, atol=1e-5)


def test_graph_without_edges(gpu):
    """No edges: no samples, an empty sparsifier; the randomized SVD of the zero
    matrix is reported as an error (its Gram is not positive definite), and the
    context stays usable."""
    g = L.HostGraph(np.zeros(301, np.int64), np.zeros(0, np.uint32)).to_device()
    csr, st = L.generate_trunc_log_matrix(g, 1, walk_len=10, sample_ratio=1.0, ctx=gpu)
    assert st["samples"] == 0 and st["nnz"] == 0 and csr.nnz == 0
    with pytest.raises(L.LightNEError):
        L.MKLRedSVD(csr, 16, seed=1, ctx=gpu).run()
    hg = L.chung_lu(2000, 2.1, 12.0, 300.0, 5)
    emb, S, _ = L.NetSMF(hg.to_device(), walk_len=5, rank=32, dim=16, sample_ratio=0.2, seed=1, ctx=gpu)
    assert np.isfinite(emb.cpu().numpy()).all() and S[0] > 0


@pytest.mark.parametrize("rank,dim,prec", [(16, 32, "fp32"), (0, 0, "fp32"), (20, 16, "fp64"),
                                           (272, 128, "fp64_lean"), (3000, 64, "fp32")])
def test_bad_shapes_are_errors(gpu, rank, dim, prec):
    """dim above rank, rank 0, rank above n, and ranks outside the 16-wide
    tiling in the FP64 forms (any rank runs in FP32 and BF16)."""
    hg = L.chung_lu(2000, 2.1, 12.0, 300.0, 5)
    p = {"fp32": L.FP32, "fp64": L.FP64, "fp64_lean": L.FP64_LEAN}[prec]
    with pytest.raises(L.LightNEError) as e:
        L.NetSMF(hg.to_device(), walk_len=5, rank=rank, dim=dim, sample_ratio=0.2, seed=1, precision=p, ctx=gpu)
    assert e.value.code == _lib.ERR_BAD_ARG


def test_rank_above_n_and_bad_propagation_width(gpu):
    hg = L.chung_lu(100, 2.1, 6.0, 30.0, 5)
    g = hg.to_device()
    csr, _ = L.generate_trunc_log_matrix(g, 1, walk_len=5, sample_ratio=1.0, ctx=gpu)
    with pytest.raises(L.LightNEError) as e:
        L.MKLRedSVD(csr, 112, seed=1, ctx=gpu).run()
    assert e.value.code == _lib.ERR_BAD_ARG
    import torch

    a = torch.zeros(hg.n, 1040, device="cuda")  # widths up to 1024
    with pytest.raises(L.LightNEError) as e:
        L.propagate(a, g, ctx=gpu)
    assert e.value.code == _lib.ERR_BAD_ARG


def test_two_contexts_give_identical_results(gpu):
    hg = L.chung_lu(3000, 2.1, 12.0, 300.0, 5)
    g = hg.to_device()
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 20.0)
    c2 = L.Context(0)
    e1, S1, _ = L.NetSMF(g, walk_len=10, rank=64, dim=32, sample_ratio=ratio, seed=3, ctx=gpu)
    e2, S2, _ = L.NetSMF(g, walk_len=10, rank=64, dim=32, sample_ratio=ratio, seed=3, ctx=c2)
    np.testing.assert_array_equal(S1, S2)
    np.testing.assert_array_equal(e1.cpu().numpy(), e2.cpu().numpy())
    c2.close()
