"""Stage 2 (randomized SVD) on the device vs the FP64 oracle of the same
algorithm with the same Gaussian sketch."""
import numpy as np
import pytest
import scipy.sparse as sp

import lightne_amd as L
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _csr_np(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(np.float32)


def _power_law_sym(n, seed, hub_nnz=3000):
    """Symmetric sparse matrix with short rows, empty rows and hub rows longer
    than several SpMM chunks."""
    rng = np.random.default_rng(seed)
    deg = np.minimum((rng.pareto(1.3, n) * 3).astype(int), n - 1)
    deg[:3] = hub_nnz
    deg[10:20] = 0
    r = np.repeat(np.arange(n), deg)
    c = rng.integers(0, n, r.shape[0])
    c[deg[r] == 0] = 0
    v = rng.standard_normal(r.shape[0])
    A = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    A = A + A.T
    A.sum_duplicates()
    return A.astype(np.float32)


@pytest.mark.parametrize("k", [16, 128, 144, 256])
def test_spmm_sketch_matches_oracle(gpu, k):
    """random_project_only returns Y = A Omega: the SpMM (short, long and
    multi-chunk rows) and the Gaussian sketch in one check."""
    n = 6000
    rp, col, val = _csr_np(_power_law_sym(n, 1))
    assert np.diff(rp).max() > 2048
    csr = L.CSR.from_arrays(rp, col, val, ctx=gpu)
    svd = L.MKLRedSVD(csr, k, random_project_only=True, seed=777, ctx=gpu).run()
    Y = svd.matU.cpu().numpy()
    assert svd.S is None
    Om = O.gaussian(777, 0, n * k).reshape(n, k)
    ref = O.spmm_f64(rp, col, val, Om.astype(np.float64))
    scale = np.abs(val).max() * np.abs(Om).max() * np.sqrt(np.diff(rp).max())
    err = np.abs(Y - ref).max() / scale
    assert err < 2e-6, err


def test_gaussian_sketch_bitwise_close(gpu):
    n, k = 4096, 144
    eye = sp.identity(n, dtype=np.float32, format="csr")
    csr = L.CSR.from_arrays(*_csr_np(eye), ctx=gpu)
    Y = L.MKLRedSVD(csr, k, random_project_only=True, seed=2024, ctx=gpu).run().matU.cpu().numpy()
    ref = O.gaussian(2024, 0, n * k).reshape(n, k)
    d = np.abs(Y.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert d.max() <= 2, d.max()
    assert np.mean(d == 0) > 0.99


def test_spmm_bf16_panel(gpu):
    n, k = 5000, 144
    rp, col, val = _csr_np(_power_law_sym(n, 2))
    csr = L.CSR.from_arrays(rp, col, val, ctx=gpu)
    Y = L.MKLRedSVD(csr, k, random_project_only=True, precision=L.BF16, seed=5, ctx=gpu).run().matU.cpu().numpy()
    Om = O.gaussian(5, 0, n * k).reshape(n, k)
    import torch

    Omb = torch.from_numpy(Om).to(torch.bfloat16).to(torch.float64).numpy()
    ref = O.spmm_f64(rp, col, val, Omb)
    scale = np.abs(val).max() * 4 * np.sqrt(np.diff(rp).max())
    assert np.abs(Y - ref).max() / scale < 2e-6


def test_exact_rank_recovery(gpu):
    rng = np.random.default_rng(7)
    n, r, k = 2000, 24, 64
    Q = np.linalg.qr(rng.standard_normal((n, r)))[0]
    lam = np.linspace(9, 1, r) * rng.choice([-1, 1], r)
    A = ((Q * lam) @ Q.T).astype(np.float32)
    csr = L.CSR.from_arrays(*_csr_np(A), ctx=gpu)
    svd = L.MKLRedSVD(csr, k, seed=1, ctx=gpu).run()
    np.testing.assert_allclose(svd.S[:r], np.sort(np.abs(lam))[::-1], rtol=2e-5)
    assert svd.S[r:].max() < 1e-4 * svd.S[0]
    U = svd.matU.cpu().numpy()
    assert H.principal_angles(U[:, :r], Q) < 1e-3


def _c1_sparsifier(gpu):
    hg = L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 50.0)
    csr, st = L.generate_trunc_log_matrix(hg.to_device(), 1, walk_len=10, sample_ratio=ratio, ctx=gpu)
    return hg, csr


@pytest.mark.parametrize("q", [1, 2])
def test_rsvd_c1_matches_fp64_oracle(gpu, q):
    """Config 1: k = 144, d = 128.  Singular values within 1e-4 relative, top-d
    subspace within 1e-3 rad of the FP64 run with the same sketch."""
    hg, csr = _c1_sparsifier(gpu)
    k, d = 144, 128
    svd = L.MKLRedSVD(csr, k, power_iters=q, seed=777, ctx=gpu).run()
    rp, col, val = csr.to_numpy()
    U64, S64 = O.rsvd(rp, col, val, k, seed=777, power_iters=q)
    rel = np.abs(svd.S[:d] - S64[:d]) / S64[:d]
    ang = H.principal_angles(svd.matU.cpu().numpy()[:, :d], U64[:, :d])
    gap = S64[d - 1] / S64[d]
    print(f"q={q} max rel sv err={rel.max():.2e} angle={ang:.2e} rad gap={gap:.4f} S0={S64[0]:.4g}")
    assert rel.max() <= 1e-4
    assert ang <= 1e-3
    emb = L.compute_u_sigma_root(svd.matU, svd.S, hg.n, k, d, False, ctx=gpu).cpu().numpy()
    ref = O.u_sigma_root(svd.matU.cpu().numpy().astype(np.float64), svd.S.astype(np.float64), d, False)
    np.testing.assert_allclose(emb, ref, rtol=1e-6, atol=1e-7)
    embn = L.compute_u_sigma_root(svd.matU, svd.S, hg.n, k, d, True, ctx=gpu).cpu().numpy()
    refn = O.u_sigma_root(svd.matU.cpu().numpy().astype(np.float64), svd.S.astype(np.float64), d, True)
    np.testing.assert_allclose(embn, refn, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("q", [1, 2])
def test_rsvd_bf16_form_is_the_fp32_result(gpu, q):
    """BF16 form: only the BF16-representable sketch is gathered as BF16, with
    FP32 accumulation in the same order as the FP32 form, so the result is the
    FP32 form's bit for bit (and within the FP64 bars)."""
    hg, csr = _c1_sparsifier(gpu)
    k = 144
    svd = L.MKLRedSVD(csr, k, power_iters=q, precision=L.BF16, seed=777, ctx=gpu).run()
    f32 = L.MKLRedSVD(csr, k, power_iters=q, seed=777, ctx=gpu).run()
    np.testing.assert_array_equal(svd.S, f32.S)
    np.testing.assert_array_equal(svd.matU.cpu().numpy(), f32.matU.cpu().numpy())


@pytest.mark.parametrize("q", [1, 2])
def test_rsvd_bf16_power_error_vs_fp64(gpu, q):
    """BF16_POWER form: the power-iteration panels are gathered as BF16 too,
    the last product B = A Q in FP32.  Same sketch as the FP64 run."""
    hg, csr = _c1_sparsifier(gpu)
    k, d = 144, 128
    svd = L.MKLRedSVD(csr, k, power_iters=q, precision=L.BF16_POWER, seed=777, ctx=gpu).run()
    rp, col, val = csr.to_numpy()
    U64, S64 = O.rsvd(rp, col, val, k, seed=777, power_iters=q)
    rel = np.abs(svd.S[:d] - S64[:d]) / S64[:d]
    ang = H.principal_angles(svd.matU.cpu().numpy()[:, :d], U64[:, :d])
    gap = S64[d - 1] / S64[d]
    print(f"BF16_POWER q={q}: max rel sv err={rel.max():.2e} angle={ang:.2e} gap={gap:.4f}")
    assert rel.max() <= 1e-4
    # q = 1 has no power iteration (the bar is the FP32 one); a BF16 power
    # iteration perturbs Q by ~2^-9, amplified by 1/(gap - 1) ~ 800 here
    assert ang <= (1e-3 if q == 1 else 1e-2)


def test_netsmf_end_to_end(gpu):
    hg = L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 50.0)
    g = hg.to_device()
    emb, S, st = L.NetSMF(g, walk_len=10, rank=144, dim=128, sample_ratio=ratio, seed=1, ctx=gpu)
    csr, _ = L.generate_trunc_log_matrix(g, 1, walk_len=10, sample_ratio=ratio, ctx=gpu)
    svd = L.MKLRedSVD(csr, 144, seed=777, ctx=gpu).run()
    ref = L.compute_u_sigma_root(svd.matU, svd.S, hg.n, 144, 128, False, ctx=gpu).cpu().numpy()
    # NetSMF forms U sqrt(S) as one product B (V S^-1/2); MKLRedSVD + compute_u_sigma_root
    # as (B V S^-1) sqrt(S): the same values up to FP32 rounding
    e = emb.cpu().numpy()
    assert np.abs(e - ref).max() <= 1e-5 * np.abs(ref).max()
    np.testing.assert_array_equal(S, svd.S)
    assert st["t_total"] > 0


def test_spmm_fp64_sketch_matches_oracle(gpu):
    """FP64 form: Y = A Omega with FP64 panels and accumulation (short, long
    and multi-chunk rows) against the oracle's FP64 product."""
    n, k = 6000, 144
    rp, col, val = _csr_np(_power_law_sym(n, 3))
    csr = L.CSR.from_arrays(rp, col, val, ctx=gpu)
    Y = L.MKLRedSVD(csr, k, random_project_only=True, precision=L.FP64, seed=9, ctx=gpu).run().matU.cpu().numpy()
    Om = O.gaussian(9, 0, n * k).reshape(n, k).astype(np.float64)
    ref = O.spmm_f64(rp, col, val, Om)
    # the FP64 result is returned rounded to FP32: within one FP32 ulp of the
    # oracle's product rounded the same way (a lost FP64 digit would not be)
    r32 = ref.astype(np.float32)
    big = np.abs(ref) > 1e-6 * np.abs(ref).max()
    ulp = np.abs(Y.view(np.int32).astype(np.int64) - r32.view(np.int32).astype(np.int64))[big]
    assert ulp.max() <= 1, ulp.max()
    assert np.abs(Y - ref)[~big].max(initial=0.0) <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("q", [1, 2])
def test_rsvd_fp64_form_matches_oracle(gpu, q):
    """The FP64 form is the accuracy reference at sizes the host oracle cannot
    reach: on config 1 it agrees with the oracle's FP64 run (same sketch) to
    FP64 rounding amplified by the spectral gap, and to FP32 rounding of U."""
    hg, csr = _c1_sparsifier(gpu)
    k, d = 144, 128
    svd = L.MKLRedSVD(csr, k, power_iters=q, precision=L.FP64, seed=777, ctx=gpu).run()
    rp, col, val = csr.to_numpy()
    U64, S64 = O.rsvd(rp, col, val, k, seed=777, power_iters=q)
    rel = np.abs(svd.S[:d].astype(np.float64) - S64[:d]) / S64[:d]
    ang = H.principal_angles(svd.matU.cpu().numpy()[:, :d].astype(np.float64), U64[:, :d])
    print(f"FP64 form q={q}: max rel sv err={rel.max():.2e} angle={ang:.2e}")
    assert rel.max() <= 1e-7  # S is returned as FP32
    assert ang <= 1e-5


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_sparse_projection_sketch_matches_oracle(gpu, prec):
    """-sparse_project: Y = A Omega with the very sparse Omega formed without a
    dense panel (short, long and hub rows) against the oracle's FP64 product."""
    n, k = 6000, 144
    rp, col, val = _csr_np(_power_law_sym(n, 4))
    csr = L.CSR.from_arrays(rp, col, val, ctx=gpu)
    p = L.FP32 if prec == "fp32" else L.FP64
    Y = L.MKLRedSVD(csr, k, random_project_only=True, precision=p, seed=31, sparse_project=True,
                    sparse_project_s=16.0, ctx=gpu).run().matU.cpu().numpy()
    ref = O.spmm_f64(rp, col, val, O.sparse_projection(n, k, 31, 16.0))
    assert np.count_nonzero(ref) > 0.5 * ref.size
    # FP64 accumulation, one rounding to FP32 at the end
    scale = np.abs(ref).max()
    assert np.abs(Y - ref).max() <= 2e-7 * scale


def test_rsvd_sparse_projection_c1(gpu):
    """Config-1 sparsifier with -sparse_project (s = sqrt(n)): singular values and
    top-d subspace against the oracle's FP64 run with the same Omega."""
    hg, csr = _c1_sparsifier(gpu)
    k, d = 144, 128
    svd = L.MKLRedSVD(csr, k, power_iters=1, seed=5, sparse_project=True, ctx=gpu).run()
    rp, col, val = csr.to_numpy()
    U64, S64 = O.rsvd(rp, col, val, k, seed=5, power_iters=1, sparse_project=True)
    rel = np.abs(svd.S[:d] - S64[:d]) / S64[:d]
    ang = H.principal_angles(svd.matU.cpu().numpy()[:, :d], U64[:, :d])
    print(f"sparse_project: max rel sv err={rel.max():.2e} angle={ang:.2e}")
    assert rel.max() <= 1e-4
    assert ang <= 1e-3


def test_upper_storage_is_the_same_matrix(gpu):
    """-upper 1 (PathEmbed.h:284-309, mklredsvd.cc:95-98): upper-triangle storage
    denotes U + U^T - diag(U).  The randomized SVD of the upper CSR equals the
    one of the full CSR up to the order of the FP32 sums over a row (the
    expansion orders each row by column, the full build keeps the canonical
    table order), and NetSMF gives the same embedding with the upper storage's
    nnz reported."""
    hg = L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 50.0)
    g = hg.to_device()
    up, su = L.generate_trunc_log_matrix(g, 1, walk_len=10, sample_ratio=ratio, upper=True, ctx=gpu)
    full, sf = L.generate_trunc_log_matrix(g, 1, walk_len=10, sample_ratio=ratio, ctx=gpu)
    assert su["nnz"] == sf["kept_upper"]
    assert sf["nnz"] == 2 * su["nnz"] - _diag_count(up)
    a = L.MKLRedSVD(up, 144, upper=True, seed=3, ctx=gpu).run()
    b = L.MKLRedSVD(full, 144, seed=3, ctx=gpu).run()
    np.testing.assert_allclose(a.S, b.S, rtol=1e-5)
    assert H.principal_angles(a.matU.cpu().numpy()[:, :128], b.matU.cpu().numpy()[:, :128]) < 1e-4
    e1, S1, st1 = L.NetSMF(g, walk_len=10, rank=144, dim=128, sample_ratio=ratio, seed=1, upper=True, ctx=gpu)
    e0, S0, st0 = L.NetSMF(g, walk_len=10, rank=144, dim=128, sample_ratio=ratio, seed=1, ctx=gpu)
    np.testing.assert_allclose(S1, S0, rtol=1e-5)
    assert H.principal_angles(e1.cpu().numpy(), e0.cpu().numpy()) < 1e-4
    assert st1["sparsify"]["nnz"] == su["nnz"]


def _diag_count(csr):
    rp, col, _ = csr.to_numpy()
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return int(np.count_nonzero(rows == col))


def test_netsmf_is_deterministic(gpu):
    """Two runs with the same seeds give the same embedding bit for bit (the
    table's slot order varies with the order of concurrent inserts; the pass
    that puts every cluster of the table in canonical order, and fixed-order
    reductions, keep every sum in one order)."""
    hg = L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 50.0)
    g = hg.to_device()
    e1, S1, _ = L.NetSMF(g, walk_len=10, rank=144, dim=128, sample_ratio=ratio, seed=4, ctx=gpu)
    e2, S2, _ = L.NetSMF(g, walk_len=10, rank=144, dim=128, sample_ratio=ratio, seed=4, ctx=gpu)
    np.testing.assert_array_equal(S1, S2)
    np.testing.assert_array_equal(e1.cpu().numpy(), e2.cpu().numpy())


def test_sketch_product_is_the_fp32_product_bit_for_bit(gpu):
    """Y = A Omega gathers the sketch from its BF16 storage: the values are
    BF16-representable, so the result is the FP32 product itself.  Rows of at
    most 40 nonzeros are summed one nonzero at a time in index order, which a
    host restatement reproduces exactly (with and without fused multiply-add);
    the device must equal one of the two in every element."""
    n, k, seed = 4000, 144, 99
    rng = np.random.default_rng(3)
    deg = rng.integers(0, 41, n)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(deg)
    col = np.concatenate([np.sort(rng.choice(n, d, replace=False)) for d in deg]).astype(np.int32)
    val = rng.standard_normal(rp[-1]).astype(np.float32)
    csr = L.CSR.from_arrays(rp, col, val, ctx=gpu)
    Y = L.MKLRedSVD(csr, k, random_project_only=True, seed=seed, ctx=gpu).run().matU.cpu().numpy()
    Om = O.gaussian(seed, 0, n * k).reshape(n, k)
    fma = np.zeros((n, k), np.float32)
    plain = np.zeros((n, k), np.float32)
    for j in range(int(deg.max())):
        r = np.nonzero(deg > j)[0]
        w = val[rp[r] + j][:, None]
        x = Om[col[rp[r] + j]]
        # w * x is exact in FP64; the sum is rounded once to FP64, then to FP32
        fma[r] = (w.astype(np.float64) * x.astype(np.float64) + fma[r].astype(np.float64)).astype(np.float32)
        plain[r] = plain[r] + w * x
    same = (Y.view(np.int32) == fma.view(np.int32)) | (Y.view(np.int32) == plain.view(np.int32))
    print(f"equal to the fused order: {np.mean(Y.view(np.int32) == fma.view(np.int32)):.6f}, "
          f"to the unfused: {np.mean(Y.view(np.int32) == plain.view(np.int32)):.6f}")
    assert same.all()
