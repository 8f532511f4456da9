"""The lean FP64 form of the randomized SVD (two FP64 panels, the FP32 form's
schedule): the accuracy reference at sizes where the explicit FP64 form's four
panels do not fit beside the sparsifier."""
import numpy as np
import pytest
import torch

import lightne_amd as L
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _c1(gpu):
    hg = L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 50.0)
    g = hg.to_device()
    csr, _ = L.generate_trunc_log_matrix(g, 1, walk_len=10, sample_ratio=ratio, ctx=gpu)
    return hg, g, ratio, csr


@pytest.mark.parametrize("q", [1, 2])
def test_lean_fp64_matches_oracle_and_explicit_form(gpu, q):
    """Same sketch: singular values within FP32 rounding of the oracle's FP64
    run, top-d subspace within 1e-5 rad; the explicit-schedule FP64 form gives
    the same S and subspace (range(Z) = range(B)); host and device outputs are
    the same bytes."""
    hg, g, ratio, csr = _c1(gpu)
    k, d = 144, 128
    lean = L.MKLRedSVD(csr, k, power_iters=q, precision=L.FP64_LEAN, seed=777, ctx=gpu).run()
    host = np.empty((hg.n, k), np.float32)
    lean_h = L.MKLRedSVD(csr, k, power_iters=q, precision=L.FP64_LEAN, seed=777, ctx=gpu).run(out=host)
    np.testing.assert_array_equal(lean.matU.cpu().numpy(), host)
    np.testing.assert_array_equal(lean.S, lean_h.S)
    rp, col, val = csr.to_numpy()
    U64, S64 = O.rsvd(rp, col, val, k, seed=777, power_iters=q)
    rel = np.abs(lean.S[:d].astype(np.float64) - S64[:d]) / S64[:d]
    ang = H.principal_angles(host[:, :d].astype(np.float64), U64[:, :d])
    print(f"lean FP64 q={q}: max rel sv err={rel.max():.2e} angle={ang:.2e}")
    assert rel.max() <= 1e-7
    assert ang <= 1e-5
    expl = L.MKLRedSVD(csr, k, power_iters=q, precision=L.FP64, seed=777, ctx=gpu).run()
    np.testing.assert_allclose(lean.S, expl.S, rtol=1e-7)
    assert H.principal_angles(host[:, :d], expl.matU.cpu().numpy()[:, :d]) <= 1e-5
    # U has orthonormal columns to FP32 rounding
    Ud = host[:, :d].astype(np.float64)
    assert np.abs(Ud.T @ Ud - np.eye(d)).max() < 1e-5


def test_lean_fp64_netsmf_host_output(gpu):
    """NetSMF in the lean FP64 form writes emb = U[:, :d] sqrt(S) to a host
    array; it equals the explicit FP64 form's embedding up to column signs
    (singular vectors) and FP32 rounding, with or without row normalisation."""
    hg, g, ratio, _ = _c1(gpu)
    for norm in (False, True):
        out = np.empty((hg.n, 128), np.float32)
        e_l, S_l, st = L.NetSMF(g, walk_len=10, rank=144, dim=128, sample_ratio=ratio, seed=1, normalize=norm,
                                precision=L.FP64_LEAN, out=out, ctx=gpu)
        assert e_l is out
        e_x, S_x, _ = L.NetSMF(g, walk_len=10, rank=144, dim=128, sample_ratio=ratio, seed=1, normalize=norm,
                               precision=L.FP64, ctx=gpu)
        np.testing.assert_allclose(S_l, S_x, rtol=1e-7)
        ex = e_x.cpu().numpy()
        if not norm:
            al = H.sign_align(out.astype(np.float64), ex)
            assert np.abs(al - ex).max() <= 1e-5 * np.abs(ex).max()
        assert H.principal_angles(out, ex) <= 1e-5
    with pytest.raises(L.LightNEError):  # host output only for the lean FP64 form
        L.NetSMF(g, walk_len=10, rank=144, dim=128, sample_ratio=ratio, seed=1, out=np.empty((hg.n, 128), np.float32),
                 ctx=gpu)


def test_wrappers_reject_bad_tensors(gpu):
    """Wrong dtype, non-contiguous views, wrong shapes and int32 row pointers are
    rejected before anything is launched."""
    X = torch.randn(64, 32, device="cuda")
    with pytest.raises(L.LightNEError, match="dtype"):
        L.dense_gram(X.double(), ctx=gpu)
    with pytest.raises(L.LightNEError, match="contiguous"):
        L.dense_gram(torch.randn(32, 64, device="cuda").T, ctx=gpu)
    with pytest.raises(L.LightNEError, match="shape"):
        L.dense_panel_gemm(X, np.eye(32), out=torch.empty(64, 16, device="cuda"), ctx=gpu)
    with pytest.raises(L.LightNEError, match="contiguous"):
        L.dense_cholqr(torch.randn(64, 64, device="cuda")[:, :32], ctx=gpu)
    rp = torch.tensor([0, 1, 2], dtype=torch.int32, device="cuda")
    with pytest.raises(L.LightNEError, match="row_ptr"):
        L.CSR.from_arrays(rp, torch.zeros(2, dtype=torch.int32, device="cuda"),
                          torch.ones(2, device="cuda"), ctx=gpu)
    hg = L.chung_lu(200, 2.1, 6.0, 50.0, 1)
    g = hg.to_device()
    with pytest.raises(L.LightNEError, match="dtype"):
        L.propagate(torch.randn(hg.n, 16, device="cuda", dtype=torch.float64), g, ctx=gpu)
    with pytest.raises(L.LightNEError, match="shape"):
        L.propagate(torch.randn(hg.n + 1, 16, device="cuda"), g, ctx=gpu)


def test_calls_on_another_torch_stream_are_ordered(gpu):
    """An input produced on a side stream and consumed by the library (bound to
    the default stream) is read after it is written, and the result is visible
    on the side stream without an explicit synchronisation."""
    hg = L.chung_lu(20000, 2.1, 8.0, 400.0, 3)
    g = hg.to_device()
    side = torch.cuda.Stream()
    ref_in = torch.randn(hg.n, 32, device="cuda")
    y_ref, _, _ = L.propagate(ref_in, g, ctx=gpu)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        a = torch.empty_like(ref_in)
        for _ in range(20):  # keep the side stream busy before the input is complete
            a.copy_(ref_in * 2.0)
        a.copy_(ref_in)
        y, _, _ = L.propagate(a, g, ctx=gpu)
        same = torch.equal(y, y_ref)
    side.synchronize()
    assert same
