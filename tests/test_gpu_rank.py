"""The reference's rank and dim domain (MKLRedSVD takes any rank,
mklredsvd.h:11-12; its own runs use -rank 4096, LightNE/blog_lightne.sh:31, and
-rank 512, LightNE/youtube_lightne.sh:31): ranks above 256 and ranks or dims
that are not multiples of 16, against the FP64 oracle of the same algorithm
with the same sketch."""
import os
import subprocess

import numpy as np
import pytest

import lightne_amd as L
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu
CLI = os.path.join(H.ROOT, "lightne_amd", "lightne")


@pytest.fixture(scope="module")
def c1(gpu):
    hg = L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 50.0)
    g = hg.to_device()
    csr, _ = L.generate_trunc_log_matrix(g, 1, walk_len=10, sample_ratio=ratio, ctx=gpu)
    return hg, g, ratio, csr, csr.to_numpy()


@pytest.mark.parametrize("k,d", [(100, 64), (200, 128), (512, 128), (4096, 128)])
def test_rsvd_any_rank_matches_fp64_oracle(gpu, c1, k, d):
    """Singular values within 1e-4 relative, top-d subspace within 1e-3 rad of
    the oracle's FP64 run (the reference's explicit schedule) with the same
    sketch; U's columns orthonormal."""
    hg, g, ratio, csr, (rp, col, val) = c1
    svd = L.MKLRedSVD(csr, k, seed=777, ctx=gpu).run()
    U = svd.matU.cpu().numpy()
    assert U.shape == (hg.n, k)
    U64, S64 = O.rsvd(rp, col, val, k, seed=777)
    rel = np.abs(svd.S[:d].astype(np.float64) - S64[:d]) / S64[:d]
    ang = H.principal_angles(U[:, :d], U64[:, :d])
    Ud = U[:, :d].astype(np.float64)
    orth = np.abs(Ud.T @ Ud - np.eye(d)).max()
    print(f"k={k}: max rel sv err {rel.max():.2e}, top-{d} angle {ang:.2e} rad, |U^T U - I| {orth:.1e}")
    assert rel.max() <= 1e-4
    assert ang <= 1e-3
    assert orth <= 1e-4
    assert np.all(np.diff(svd.S.astype(np.float64)) <= 0)


@pytest.mark.parametrize("k", [100, 300])
def test_sketch_any_rank(gpu, c1, k):
    """random_project_only: Y = A Omega with the padded sketch panel and the
    slabbed SpMM, against the oracle's product (the same Omega)."""
    hg, g, ratio, csr, (rp, col, val) = c1
    Y = L.MKLRedSVD(csr, k, random_project_only=True, seed=5, ctx=gpu).run().matU.cpu().numpy()
    Om = O.gaussian(5, 0, hg.n * k).reshape(hg.n, k)
    ref = O.spmm_f64(rp, col, val, Om.astype(np.float64))
    scale = np.abs(ref).max()
    assert np.abs(Y - ref).max() <= 2e-6 * scale


def test_bf16_form_wide_rank_is_the_fp32_result(gpu, c1):
    hg, g, ratio, csr, _ = c1
    a = L.MKLRedSVD(csr, 320, seed=3, ctx=gpu).run()
    b = L.MKLRedSVD(csr, 320, precision=L.BF16, seed=3, ctx=gpu).run()
    np.testing.assert_array_equal(a.S, b.S)
    np.testing.assert_array_equal(a.matU.cpu().numpy(), b.matU.cpu().numpy())


@pytest.mark.parametrize("k,d", [(100, 36), (512, 128)])
def test_netsmf_any_rank_is_u_sqrt_s(gpu, c1, k, d):
    """NetSMF's embedding at any rank is U[:, :d] sqrt(S) of the same rSVD
    (PathEmbed.h:371-408), plain and row-normalised."""
    hg, g, ratio, csr, _ = c1
    svd = L.MKLRedSVD(csr, k, seed=777, ctx=gpu).run()
    U = svd.matU.cpu().numpy().astype(np.float64)
    for norm in (False, True):
        emb, S, _ = L.NetSMF(g, walk_len=10, rank=k, dim=d, sample_ratio=ratio, seed=1, normalize=norm, ctx=gpu)
        np.testing.assert_array_equal(S, svd.S)
        ref = O.u_sigma_root(U, svd.S.astype(np.float64), d, norm)
        e = H.sign_align(emb.cpu().numpy().astype(np.float64), ref)
        assert np.abs(e - ref).max() <= 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("dim", [40, 100])
def test_propagation_any_dim(gpu, dim):
    """Propagation and the dense tall SVD with a dim that is not a multiple of
    16 (zero-padded columns) against the FP64 oracle."""
    hg = L.chung_lu(20000, 2.1, 10.0, 800.0, 4)
    g = hg.to_device()
    import torch

    a = torch.randn(hg.n, dim, generator=torch.Generator(device="cuda").manual_seed(3), device="cuda")
    y, z, _ = L.propagate(a, g, ctx=gpu)
    y64, z64 = O.propagate(hg.offsets, hg.edges, a.cpu().numpy().astype(np.float64))
    assert np.abs(z.cpu().numpy() - z64).max() <= 1e-5 * np.abs(z64).max()
    assert np.abs(y.cpu().numpy() - y64).max() <= 1e-5 * np.abs(y64).max()
    out, _ = L.chebyshev_expansion(a, g, dim, dim, ctx=gpu)
    e64, _ = O.embed_dense(z64, dim)
    assert H.principal_angles(out.cpu().numpy(), e64) <= 1e-3


def test_embed_dense_wide(gpu):
    """get_embedding_dense with rank above 256 and a dim not a multiple of 16."""
    import torch

    n, rank, dim = 30000, 300, 70
    rng = np.random.default_rng(2)
    M = (rng.standard_normal((n, rank)) * np.linspace(3, 0.1, rank)).astype(np.float32)
    emb, S = L.get_embedding_dense(torch.from_numpy(M).cuda(), n, rank, dim, ctx=gpu)
    ref, S64 = O.embed_dense(M.astype(np.float64), dim)
    np.testing.assert_allclose(S, S64, rtol=1e-5)
    e = H.sign_align(emb.cpu().numpy().astype(np.float64), ref)
    assert np.abs(e - ref).max() <= 1e-4


def test_blogcatalog_script_flags_through_the_cli(gpu, tmp_path):
    """LightNE/blog_lightne.sh's exact flag set (-rank 4096, -sample 1,
    -sample_ratio 2000, -tablesz 6679660000, ...) runs through the drop-in CLI
    on BlogCatalog and writes both embeddings."""
    off, edg = H.golden_graph("blogcatalog")
    adj = str(tmp_path / "blogcatalog.adj")
    L.write_adjacency_graph(adj, L.HostGraph(off, edg))
    ne, pro = str(tmp_path / "blog.ne"), str(tmp_path / "blog.pro")
    cmd = [CLI, "-walksperedge", "10000", "-walklen", "10", "-rounds", "1", "-s", "-m", "-ne_out", ne,
           "-pro_out", pro, "-ne_method", "netsmf", "-rank", "4096", "-dim", "128", "-order", "10", "-analyze", "1",
           "-sample", "1", "-sample_ratio", "2000", "-upper", "0", "-tablesz", "6679660000", adj]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=280)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    assert "### Running Time" in out.stdout
    n = len(off) - 1
    e = np.fromfile(ne, np.float32).reshape(n, 128)
    p = np.fromfile(pro, np.float32).reshape(n, 128)
    assert np.isfinite(e).all() and np.isfinite(p).all()
    nrm = np.linalg.norm(p, axis=1)
    assert np.abs(nrm[nrm > 1e-4] - 1).max() < 1e-5
    print([ln for ln in out.stdout.splitlines() if "Running Time" in ln or "nnz" in ln])
