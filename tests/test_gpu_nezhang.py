"""The ne_zhang_et_al method (PathEmbed.h:410-463) on the device vs the oracle."""
import numpy as np
import pytest

import lightne_amd as L
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", ["cl2000", "c1"])
def test_ne_zhang_matrix_matches_oracle(gpu, graph):
    """Pattern identical (graph row pointers, adjacency order), values within
    one FP32 ulp (the FP64 sums may round in another order)."""
    if graph == "c1":
        hg = L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)
    else:
        hg = L.chung_lu(2000, 2.1, 12.0, 300.0, 5)
    csr = L.ne_zhang_matrix(hg.to_device(), ctx=gpu)
    rp, col, val = csr.to_numpy()
    erp, ecol, evl = O.ne_zhang_matrix(hg.offsets, hg.edges)
    assert np.array_equal(rp, erp) and np.array_equal(col, ecol)
    ulp = np.abs(val.view(np.int32).astype(np.int64) - evl.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1, ulp.max()


def test_ne_zhang_embedding_matches_fp64_oracle(gpu):
    """emb = rownorm(U[:, :d] sqrt(S)) (PathEmbed.h:445-460) on config 1's graph:
    singular values within 1e-4, top-d subspace within 1e-3 rad of the FP64
    randomized SVD of the oracle's matrix (same sketch)."""
    hg = L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)
    k, d = 144, 128
    emb, S, st = L.NE_Zhang_et_al(hg.to_device(), rank=k, dim=d, ctx=gpu)
    rp, col, val = O.ne_zhang_matrix(hg.offsets, hg.edges)
    U64, S64 = O.rsvd(rp, col, val, k, seed=777, power_iters=1)
    rel = np.abs(S[:d] - S64[:d]) / S64[:d]
    e = emb.cpu().numpy()
    ref = O.u_sigma_root(U64, S64, d, True)
    ang = H.principal_angles(e, ref)
    print(f"ne_zhang: max rel sv err={rel.max():.2e} angle={ang:.2e} gap={S64[d - 1] / S64[d]:.4f}")
    assert rel.max() <= 1e-4
    assert ang <= 1e-3
    nrm = np.linalg.norm(e, axis=1)
    assert np.all(np.abs(nrm[nrm > 1e-3] - 1.0) < 1e-5)
