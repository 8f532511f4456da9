"""The CLI (drop-in for the reference runner) against the Python API on the same seeds."""
import os
import subprocess

import numpy as np
import pytest

import lightne_amd as L
from tests import helpers as H

pytestmark = pytest.mark.gpu
CLI = os.path.join(H.ROOT, "lightne_amd", "lightne")


def test_cli_matches_api(gpu, tmp_path):
    import torch

    hg = L.chung_lu(3000, 2.1, 12.0, 300.0, 5)
    adj = str(tmp_path / "g.adj")
    L.write_adjacency_graph(adj, hg)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 20.0)
    ne, pro = str(tmp_path / "ne.emb"), str(tmp_path / "pro.emb")
    out = subprocess.run([CLI, "-walklen", "10", "-rank", "64", "-dim", "32", "-order", "10",
                          "-sample_ratio", repr(ratio), "-seed", "9", "-rounds", "1", "-ne_out", ne, "-pro_out", pro,
                          "-s", adj], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    for line in ["### Application: LightNE_mkl", "# total samples", "# nnz in sparsifier", "### Running Time"]:
        assert line in out.stdout
    e1 = np.fromfile(ne, np.float32).reshape(-1, 32)  # predict.py --binary layout
    p1 = np.fromfile(pro, np.float32).reshape(-1, 32)
    g = hg.to_device()
    emb, S, _ = L.NetSMF(g, walk_len=10, rank=64, dim=32, sample_ratio=ratio, seed=9)
    np.testing.assert_array_equal(e1, emb.cpu().numpy())
    p2, _ = L.chebyshev_expansion(emb, g, 32, 32)
    np.testing.assert_array_equal(p1, p2.cpu().numpy())
    nrm = np.linalg.norm(p1, axis=1)
    assert np.allclose(nrm[nrm > 1e-4], 1.0, atol=1e-5)  # rows of isolated vertices stay zero
    # -ne_in runs stage 3 alone from the saved stage-2 embedding
    pro2 = str(tmp_path / "pro2.emb")
    out = subprocess.run([CLI, "-dim", "32", "-rank", "64", "-rounds", "1", "-ne_in", ne, "-pro_out", pro2, "-s", adj],
                         capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    np.testing.assert_array_equal(np.fromfile(pro2, np.float32), p1.ravel())


def test_cli_ne_zhang_matches_api(gpu, tmp_path):
    """-ne_method ne_zhang_et_al (LightNE.cc:106) writes the NE_Zhang_et_al embedding."""
    hg = L.chung_lu(3000, 2.1, 12.0, 300.0, 5)
    adj = str(tmp_path / "g.adj")
    L.write_adjacency_graph(adj, hg)
    ne = str(tmp_path / "ne.emb")
    out = subprocess.run([CLI, "-ne_method", "ne_zhang_et_al", "-rank", "64", "-dim", "32", "-rounds", "1",
                          "-ne_out", ne, "-s", adj], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    assert "### Running Time" in out.stdout
    e1 = np.fromfile(ne, np.float32).reshape(-1, 32)
    emb, S, _ = L.NE_Zhang_et_al(hg.to_device(), rank=64, dim=32)
    np.testing.assert_array_equal(e1, emb.cpu().numpy())
