"""Stage 1 on the device vs the oracle: per-key sample counts bit-exact, CSR
exact (values within one float ulp)."""
import numpy as np
import pytest

import lightne_amd as L
from lightne_amd import _lib
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def dgraph(off, edg):
    return L.HostGraph(off, edg).to_device()


def case_graph(rec, ctx):
    """The case's graph on the device; compressed-format cases go through the
    device decoder of the encoded file, as `lightne -c` does."""
    off, edg = H.golden_graph(rec["graph"])
    if not rec.get("compressed"):
        return off, edg, dgraph(off, edg), 0
    import torch
    from lightne_amd import bytepd

    img = torch.from_numpy(bytepd.encode(off, edg)).cuda()
    return off, edg, L.decode_compressed_graph(img, ctx=ctx), 1


@pytest.mark.parametrize("rec", H.manifest(), ids=lambda r: r["case"])
def test_counts_match_golden_and_oracle(gpu, rec):
    off, edg, g, ith = case_graph(rec, gpu)
    t = L.sparsify_counts(g, int(rec["seed"]), rec["walk_len"], bool(rec["sample"]),
                          sample_ratio=rec["sample_ratio"], ctx=gpu)
    keys, cnt, wgh, st = O.sample(off, edg, int(rec["seed"]), rec["walk_len"], rec["sample_ratio"], rec["sample"],
                                  threads=1, ith=ith)
    np.testing.assert_array_equal(t.keys, keys)
    np.testing.assert_array_equal(t.counts.astype(np.uint64), cnt)
    if rec["sample"] == 0:
        np.testing.assert_array_equal(t.weights, wgh)
    else:  # FP64 sums of 1/p in a different order
        np.testing.assert_allclose(t.weights, wgh, rtol=1e-12)
    assert t.stats["samples"] == st["samples"]
    assert t.stats["walk_steps"] == st["walk_steps"]
    assert t.stats["attempts"] == st["attempts"]
    gd = H.golden_case(rec["case"])
    if rec["full"]:
        np.testing.assert_array_equal(t.keys, gd["keys"])


def _compare_csr(csr, ref, ulps=1):
    rp, col, val = csr.to_numpy()
    erp, ecol, evl, _ = ref
    np.testing.assert_array_equal(rp, erp)
    np.testing.assert_array_equal(col, ecol)
    d = np.abs(val.view(np.int32).astype(np.int64) - evl.view(np.int32).astype(np.int64))
    assert d.max(initial=0) <= ulps, d.max()
    return int(np.sum(d > 0))


@pytest.mark.parametrize("rec", H.manifest(), ids=lambda r: r["case"])
def test_csr_matches_oracle(gpu, rec):
    off, edg, g, ith = case_graph(rec, gpu)
    csr, st = L.generate_trunc_log_matrix(g, int(rec["seed"]), walk_len=rec["walk_len"], sample=bool(rec["sample"]),
                                          sample_ratio=rec["sample_ratio"], ctx=gpu)
    keys, cnt, wgh, ost = O.sample(off, edg, int(rec["seed"]), rec["walk_len"], rec["sample_ratio"], rec["sample"],
                                   threads=1, ith=ith)
    if rec["sample"] == 0:
        assert st["M_exact"] == ost["samples"]
        ref = O.trunc_log_csr(off, keys, wgh)
        assert ref[3]["factor"] == st["factor"]
    else:
        ref = O.trunc_log_csr(off, keys, wgh, M_exact=st["M_exact"])
    assert st["distinct_keys"] == len(keys)
    assert st["kept_upper"] == ref[3]["kept_upper"]
    assert st["nnz"] == csr.nnz == ref[3]["nnz"]
    _compare_csr(csr, ref)
    csr.free()


def _c1_graph():
    return L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)


def test_c1_counts_and_csr_bit_exact(gpu):
    """Config 1 shape: Chung-Lu n=10,312, T=10, 50 samples per directed edge."""
    hg = _c1_graph()
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 50.0)
    g = hg.to_device()
    t = L.sparsify_counts(g, 1, 10, False, sample_ratio=ratio, ctx=gpu)
    keys, cnt, wgh, st = O.sample(hg.offsets, hg.edges, 1, 10, ratio, 0)
    assert st["samples"] > 30_000_000
    np.testing.assert_array_equal(t.keys, keys)
    np.testing.assert_array_equal(t.counts.astype(np.uint64), cnt)
    assert t.stats["walk_steps"] == st["walk_steps"]
    csr, s2 = L.generate_trunc_log_matrix(g, 1, walk_len=10, sample_ratio=ratio, ctx=gpu)
    ref = O.trunc_log_csr(hg.offsets, keys, wgh)
    nd = _compare_csr(csr, ref)
    print(f"C1: samples={st['samples']} keys={len(keys)} nnz={csr.nnz} value mismatches(<=1ulp)={nd}")
    # deterministic across runs
    csr2, _ = L.generate_trunc_log_matrix(g, 1, walk_len=10, sample_ratio=ratio, ctx=gpu)
    a, b = csr.to_numpy(), csr2.to_numpy()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_c2_counts_bit_exact(gpu):
    """Config 2: Chung-Lu n=1.14M, m=2.99M, T=10, M = 20 m (10 per directed edge)."""
    hg = L.chung_lu(1_140_000, 2.0, 5.25, 2500.0, 2)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 10.0)
    g = hg.to_device()
    t = L.sparsify_counts(g, 20240601, 10, False, sample_ratio=ratio, ctx=gpu)
    keys, cnt, wgh, st = O.sample(hg.offsets, hg.edges, 20240601, 10, ratio, 0)
    np.testing.assert_array_equal(t.keys, keys)
    np.testing.assert_array_equal(t.counts.astype(np.uint64), cnt)
    assert t.stats["samples"] == st["samples"] and t.stats["walk_steps"] == st["walk_steps"]
    print(f"C2: samples={st['samples']} keys={len(keys)}")


def test_table_full_is_an_error_and_context_survives(gpu):
    off, edg = H.golden_graph("cl2000")
    g = dgraph(off, edg)
    with pytest.raises(L.LightNEError) as e:
        L.sparsify_counts(g, 3, 10, False, sample_ratio=0.9, max_table_bytes=64 * 1024, ctx=gpu)
    assert e.value.code == _lib.ERR_TABLE_FULL
    t = L.sparsify_counts(g, 3, 10, False, sample_ratio=0.9, ctx=gpu)
    keys, cnt, _, _ = O.sample(off, edg, 3, 10, 0.9, 0, threads=1)
    np.testing.assert_array_equal(t.keys, keys)


def test_asymmetric_graph_is_rejected(gpu):
    # 0 -> 1 but 1 has no neighbours: a walk from 1 would draw `rand % 0`
    off = np.array([0, 1, 1], np.int64)
    edg = np.array([1], np.uint32)
    g = dgraph(off, edg)
    with pytest.raises(L.LightNEError) as e:
        L.sparsify_counts(g, 1, 10, False, sample_ratio=50.0, ctx=gpu)
    assert e.value.code == _lib.ERR_BAD_GRAPH
