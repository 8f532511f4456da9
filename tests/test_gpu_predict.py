"""Downstream quality: BlogCatalog node classification (the LightNE/predict.py
protocol) with the device embedding and with the oracle's FP64 embedding of the
same sparsifier; micro / macro F1 must agree."""
import numpy as np
import pytest

import lightne_amd as L
from lightne_amd import predict as P
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_blogcatalog_f1_device_vs_fp64_oracle(gpu):
    off, edg = H.golden_graph("blogcatalog")
    hg = L.HostGraph(off, edg)
    g = hg.to_device()
    k, d, T = 256, 128, 10
    ratio = H.c1_sample_ratio(hg.n, hg.m, T, 20.0)
    # device: NetSMF (stages 1 + 2) and the spectral propagation
    emb, S, _ = L.NetSMF(g, walk_len=T, rank=k, dim=d, sample_ratio=ratio, seed=1, ctx=gpu)
    pro, _ = L.chebyshev_expansion(emb, g, d, d, order=10, mu=0.2, s=0.5, ctx=gpu)
    # oracle: FP64 randomized SVD and propagation of the same sparsifier (bit-exact
    # with the oracle's own stage 1, which the stage-1 tests check)
    csr, _ = L.generate_trunc_log_matrix(g, 1, walk_len=T, sample_ratio=ratio, ctx=gpu)
    rp, col, val = csr.to_numpy()
    U64, S64 = O.rsvd(rp, col, val, k, seed=777)
    e64 = O.u_sigma_root(U64, S64, d, False)
    y64, z64 = O.propagate(off, edg, e64, order=10, mu=0.2, theta=0.5)
    p64, _ = O.embed_dense(z64, d)
    y = P.load_label(H.golden_path("blogcatalog_labels.npz"))
    keep = np.asarray(y.getnnz(axis=1)) > 0
    res = {}
    for name, X in [("device_ne", emb.cpu().numpy()), ("oracle_ne", e64), ("device_pro", pro.cpu().numpy()),
                    ("oracle_pro", p64)]:
        res[name] = P.predict_cv(np.asarray(X, np.float64)[keep], y[keep], train_ratio=0.5, n_splits=3, C=10.0)
    print({k_: (round(v[0], 2), round(v[1], 2)) for k_, v in res.items()})
    for stage in ["ne", "pro"]:
        dv, ov = res[f"device_{stage}"], res[f"oracle_{stage}"]
        assert abs(dv[0] - ov[0]) <= 1.0 and abs(dv[1] - ov[1]) <= 1.5, (stage, dv, ov)
    assert res["device_pro"][0] > 25.0  # well above the label prior
