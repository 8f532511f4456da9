"""Stage 1 on the device vs the oracle: per-key sample counts bit-exact, CSR
rows equal as sets (values within one float ulp), and each row's entries in
the order the canonical table layout fixes."""
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


@pytest.mark.parametrize("case", ["toy_t3_s1", "cl2000_t10_s0", "cl2000_t5_s1", "blog_t10_s0", "ccl2000_t10_s0",
                                  "ccl2000_t5_s1"])
def test_walk_layouts_draw_the_same_samples(gpu, case):
    """The walks over neighbour rows (each directed edge's target with its row,
    the default), over 32-bit row pointers and over the graph's 64-bit offsets
    draw the same samples: equal counts, tables and CSR, bit for bit with
    -sample 0 (with -sample 1 the FP64 weights are sums in atomic order)."""
    rec = next(r for r in H.manifest() if r["case"] == case)
    off, edg, g, ith = case_graph(rec, gpu)
    kw = dict(walk_len=rec["walk_len"], sample=bool(rec["sample"]), sample_ratio=rec["sample_ratio"], ctx=gpu)
    runs = {}
    for w in (0, 2, 1):
        t = L.sparsify_counts(g, int(rec["seed"]), rec["walk_len"], bool(rec["sample"]),
                              sample_ratio=rec["sample_ratio"], wide_offsets=w, ctx=gpu)
        csr, st = L.generate_trunc_log_matrix(g, int(rec["seed"]), wide_offsets=w, **kw)
        runs[w] = (t, csr.to_numpy(), st)
        csr.free()
        assert st["walk_rows"] == (1 if w == 0 else 0) and t.stats["walk_rows"] == st["walk_rows"]
    t0, c0, s0 = runs[0]
    for w in (2, 1):
        t, c, st = runs[w]
        np.testing.assert_array_equal(t.keys, t0.keys)
        np.testing.assert_array_equal(t.counts, t0.counts)
        for k in ("samples", "walk_steps", "attempts", "distinct_keys"):
            assert st[k] == s0[k], k
        if not rec["sample"]:
            np.testing.assert_array_equal(t.weights, t0.weights)
            for x, y in zip(c, c0):
                np.testing.assert_array_equal(x, y)
        else:
            np.testing.assert_allclose(t.weights, t0.weights, rtol=1e-12)
    keys, cnt, wgh, ost = O.sample(off, edg, int(rec["seed"]), rec["walk_len"], rec["sample_ratio"], rec["sample"],
                                   threads=1, ith=ith)
    np.testing.assert_array_equal(t0.keys, keys)
    np.testing.assert_array_equal(t0.counts.astype(np.uint64), cnt)


def _compare_csr(csr, ref, ulps=1):
    """The device CSR against a reference, row by row as sets of (column,
    value): the same columns, values within `ulps`."""
    rp, col, val = H.sort_rows(*csr.to_numpy())
    erp, ecol, evl = H.sort_rows(*ref[:3])
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


def _canonical_row_order(keys, kept, cap):
    """Host model of the CSR's within-row order: linear probing of `keys` in a
    table of `cap` slots, each cluster (maximal run of occupied slots) sorted
    by key, then the trunc-log emission's order, which is slot order (every
    wave quarter of the partition writes its own run, the quarters in slot
    order), the kept keys' entries (u, v) then (v, u), and a stable sort by
    row.  Returns the entries' (row, col) in CSR order."""
    h = np.array([(O.hash64_2(int(k)) * cap) >> 64 for k in keys], np.int64)
    slot = np.full(cap, -1, np.int64)
    for i, hi in enumerate(h):
        p = int(hi)
        while slot[p] >= 0:
            p = p + 1 if p + 1 < cap else 0
        slot[p] = i
    # clusters: maximal runs of occupied slots (the table has an empty slot)
    occ = slot >= 0
    start = int(np.flatnonzero(~occ)[0]) + 1
    order = [(start + j) % cap for j in range(cap)]
    run = []
    for p in order + [None]:
        if p is not None and occ[p]:
            run.append(p)
            continue
        if len(run) > 1:
            ids = slot[run]
            slot[run] = ids[np.argsort(keys[ids])]
        run = []
    pos = np.flatnonzero(occ)
    rows, cols = [], []
    for p in pos:
        i = slot[p]
        if not kept[i]:
            continue
        u, v = int(keys[i] >> 32), int(keys[i] & 0xFFFFFFFF)
        rows.append(u)
        cols.append(v)
        if u != v:
            rows.append(v)
            cols.append(u)
    rows, cols = np.array(rows, np.int64), np.array(cols, np.int32)
    o = np.argsort(rows, kind="stable")
    return rows[o], cols[o]


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("case", ["cl2000_t10_s0", "cl200_t5_s1"])
def test_csr_row_order_is_the_canonical_emission_order(gpu, case, layout):
    """The entries of every CSR row come in the order a host model of the
    canonical table layout and of the emission predicts, for both slot
    layouts, so that order (and every FP32 sum over a row) does not depend on
    the order in which concurrent inserts claimed slots."""
    rec = _rec(case)
    off, edg = H.golden_graph(rec["graph"])
    csr, st = L.generate_trunc_log_matrix(dgraph(off, edg), int(rec["seed"]), walk_len=rec["walk_len"],
                                          sample=bool(rec["sample"]), sample_ratio=rec["sample_ratio"],
                                          table_layout=layout, ctx=gpu)
    assert st["table_layout"] == layout and st["row_order"] == 1
    rp, col, _ = csr.to_numpy()
    keys, _, wgh, _ = O.sample(off, edg, int(rec["seed"]), rec["walk_len"], rec["sample_ratio"], rec["sample"],
                               threads=1)
    ref = O.trunc_log_csr(off, keys, wgh, M_exact=st["M_exact"]) if rec["sample"] else \
        O.trunc_log_csr(off, keys, wgh)
    erp, ecol = ref[0], ref[1]
    erows = np.repeat(np.arange(len(erp) - 1, dtype=np.int64), np.diff(erp))
    kept_pairs = set(zip(erows[erows <= ecol].tolist(), ecol[erows <= ecol].tolist()))
    kept = np.array([(int(k >> 32), int(k & 0xFFFFFFFF)) in kept_pairs for k in keys])
    rows, cols = _canonical_row_order(keys.astype(np.uint64), kept, int(st["table_capacity"]))
    np.testing.assert_array_equal(rp, erp)
    np.testing.assert_array_equal(col, cols)
    csr.free()


def test_crowded_table_rows_fall_back_to_column_order(gpu):
    """A table capped near full (load 0.97) has clusters longer than the
    canonical pass stages: the rows are then sorted by column (row_order 2),
    equal the oracle's row by row, and a second run gives the same CSR."""
    rec = _rec("cl2000_t10_s0")
    off, edg = H.golden_graph(rec["graph"])
    keys, _, wgh, _ = O.sample(off, edg, int(rec["seed"]), rec["walk_len"], rec["sample_ratio"], rec["sample"],
                               threads=1)
    cap = (int(len(keys) / 0.97) + 255) // 256 * 256
    g = dgraph(off, edg)
    kw = dict(walk_len=rec["walk_len"], sample_ratio=rec["sample_ratio"], max_table_bytes=cap * 8, table_layout=2,
              ctx=gpu)
    a, st = L.generate_trunc_log_matrix(g, int(rec["seed"]), **kw)
    assert st["table_capacity"] == cap and st["distinct_keys"] == len(keys)
    assert st["row_order"] == 2
    _compare_csr(a, O.trunc_log_csr(off, keys, wgh))
    rp, col, val = a.to_numpy()
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.int64), np.diff(rp))
    assert np.all((np.diff(col) > 0) | (np.diff(rows) != 0))  # each row by column
    b, st2 = L.generate_trunc_log_matrix(g, int(rec["seed"]), **kw)
    assert st2["row_order"] == 2
    for x, y in zip((rp, col, val), b.to_numpy()):
        np.testing.assert_array_equal(x, y)
    a.free()
    b.free()


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


def test_sparse_graph_csr_sorted_beside_the_table(gpu):
    """A sparse graph whose sort buffers do not fit the dead table (so the sort
    gets buffers of its own) and whose 26-bit keys take an odd number of radix
    passes (the emission then writes the values into the sort's temporary):
    the CSR still equals the oracle's."""
    hg = L.chung_lu(8000, 2.1, 4.0, 200.0, 11)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 10.0)
    csr, st = L.generate_trunc_log_matrix(hg.to_device(), 5, walk_len=10, sample_ratio=ratio, ctx=gpu)
    assert st["table_layout"] == 2 and st["nnz"] * 12 > st["table_capacity"] * 8  # keys + values exceed the table
    keys, cnt, wgh, _ = O.sample(hg.offsets, hg.edges, 5, 10, ratio, 0)
    ref = O.trunc_log_csr(hg.offsets, keys, wgh)
    assert st["nnz"] == csr.nnz == ref[3]["nnz"]
    _compare_csr(csr, ref)
    csr.free()


def test_c2_counts_bit_exact(gpu):
    """Config 2: Chung-Lu n=1.14M, m=2.99M, T=10, M = 20 m (10 per directed edge):
    per-key counts bit-exact and the CSR equal to the oracle's row by row."""
    hg = L.chung_lu(1_140_000, 2.0, 5.25, 2500.0, 2)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 10.0)
    g = hg.to_device()
    t = L.sparsify_counts(g, 20240601, 10, False, sample_ratio=ratio, ctx=gpu)
    keys, cnt, wgh, st = O.sample(hg.offsets, hg.edges, 20240601, 10, ratio, 0)
    np.testing.assert_array_equal(t.keys, keys)
    np.testing.assert_array_equal(t.counts.astype(np.uint64), cnt)
    assert t.stats["samples"] == st["samples"] and t.stats["walk_steps"] == st["walk_steps"]
    print(f"C2: samples={st['samples']} keys={len(keys)}")
    # the C2 sparsifier against the oracle's trunc-log CSR, row by row as sets
    # (rows sorted by (row, col) on the device: 9.7e7 entries)
    import torch

    csr, s2 = L.generate_trunc_log_matrix(g, 20240601, walk_len=10, sample_ratio=ratio, ctx=gpu)
    erp, ecol, evl, info = O.trunc_log_csr(hg.offsets, keys, wgh)
    assert s2["factor"] == info["factor"] and s2["nnz"] == info["nnz"] == csr.nnz
    rp, col, val = csr.to_numpy()
    np.testing.assert_array_equal(rp, erp)
    rows = torch.repeat_interleave(torch.arange(hg.n, device="cuda"), torch.from_numpy(np.diff(rp)).cuda())
    order = torch.sort(rows * hg.n + torch.from_numpy(col).cuda().long(), stable=True).indices
    dcol = torch.from_numpy(col).cuda()[order]
    dval = torch.from_numpy(val.view(np.int32)).cuda()[order].long()
    assert torch.equal(dcol, torch.from_numpy(ecol).cuda())
    ulps = (dval - torch.from_numpy(evl.view(np.int32)).cuda().long()).abs().max().item()
    print(f"C2 CSR: nnz={csr.nnz} row order {s2['row_order']}, max value difference {ulps} ulp")
    assert ulps <= 1


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


def _rec(case):
    return next(r for r in H.manifest() if r["case"] == case)


@pytest.mark.parametrize("case", ["toy_t3_s1", "cl2000_t10_s0", "cl2000_t5_s1", "blog_t10_s0"])
def test_table_layouts_agree(gpu, case):
    """8-byte and 16-byte slots hold the same key -> (count, weight) multiset."""
    rec = _rec(case)
    off, edg = H.golden_graph(rec["graph"])
    g = dgraph(off, edg)
    kw = dict(sample_ratio=rec["sample_ratio"], ctx=gpu)
    tw = L.sparsify_counts(g, int(rec["seed"]), rec["walk_len"], bool(rec["sample"]), table_layout=1, **kw)
    tp = L.sparsify_counts(g, int(rec["seed"]), rec["walk_len"], bool(rec["sample"]), table_layout=2, **kw)
    assert tw.stats["table_layout"] == 1 and tp.stats["table_layout"] == 2
    np.testing.assert_array_equal(tp.keys, tw.keys)
    np.testing.assert_array_equal(tp.counts, tw.counts)
    if rec["sample"] == 0:
        np.testing.assert_array_equal(tp.weights, tw.weights)
    else:  # FP64 sums of 1/p in the order the atomics landed
        np.testing.assert_allclose(tp.weights, tw.weights, rtol=1e-12)


@pytest.mark.parametrize("walk_len,ratio", [(10, 1100.0), (1, 3000.0)])
def test_packed_counts_wrap_exactly(gpu, walk_len, ratio):
    """Counts far above the packed slot's 10 bits (several wraps into the side
    table) stay exact: packed == wide == oracle.  With T = 1 every sample is
    its edge, so each directed edge adds its whole count (thousands) to the
    key in one add: a new slot that starts wrapped, and an add that wraps
    several times."""
    off, edg = H.golden_graph("toy")
    g = dgraph(off, edg)
    tp = L.sparsify_counts(g, 11, walk_len, False, sample_ratio=ratio, table_layout=2, ctx=gpu)
    tw = L.sparsify_counts(g, 11, walk_len, False, sample_ratio=ratio, table_layout=1, ctx=gpu)
    keys, cnt, _, st = O.sample(off, edg, 11, walk_len, ratio, 0, threads=8)
    assert int(cnt.max()) > 3 * 1024
    np.testing.assert_array_equal(tp.keys, keys)
    np.testing.assert_array_equal(tp.counts.astype(np.uint64), cnt)
    np.testing.assert_array_equal(tw.counts, tp.counts)
    assert tp.stats["samples"] == st["samples"]


@pytest.mark.parametrize("case,parts,upper", [("cl2000_t10_s0", 5, False), ("cl2000_t5_s1", 3, False),
                                              ("blog_t10_s0", 9, False), ("blog_t10_s0", 4, True),
                                              ("toy_t10_s0", 3, False)])
def test_csr_by_row_ranges_matches_one_sort(gpu, case, parts, upper):
    """Building the CSR range by range (one sort per row range) gives the matrix
    of the single global sort."""
    rec = _rec(case)
    off, edg = H.golden_graph(rec["graph"])
    g = dgraph(off, edg)
    kw = dict(walk_len=rec["walk_len"], sample=bool(rec["sample"]), sample_ratio=rec["sample_ratio"], upper=upper,
              ctx=gpu)
    a, sa = L.generate_trunc_log_matrix(g, int(rec["seed"]), **kw)
    b, sb = L.generate_trunc_log_matrix(g, int(rec["seed"]), csr_parts=parts, **kw)
    assert sa["csr_parts"] == 1 and sb["csr_parts"] >= 2
    assert sa["nnz"] == sb["nnz"] == a.nnz == b.nnz
    rb = b.to_numpy()
    if rec["sample"] == 0:
        for x, y in zip(a.to_numpy(), rb):
            np.testing.assert_array_equal(x, y)
    else:  # M_exact and the weights are FP64 sums in atomic order
        _compare_csr(a, (*rb, None))
    a.free()
    b.free()


@pytest.mark.parametrize("sample", [False, True])
def test_netsmf_row_ranges_in_a_fresh_context(gpu, sample):
    """The pipeline's context-owned build by row ranges (the workspace grows and
    the sampler runs again) matches the single-sort pipeline, on the first and
    on a repeated call: bit for bit with -sample 0; with -sample 1 the FP64
    weights are sums in atomic order, so the layout is planned again after the
    growth and the result agrees to rounding."""
    hg = L.chung_lu(20000, 2.1, 10.0, 500.0, 7)
    g = hg.to_device()
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 20.0)
    kw = dict(walk_len=10, rank=64, dim=32, sample_ratio=ratio, seed=3, sample=sample)
    e1, S1, st1 = L.NetSMF(g, ctx=gpu, **kw)
    ctx2 = L.Context(0)
    try:
        for _ in range(2):
            e2, S2, st2 = L.NetSMF(g, csr_parts=6, ctx=ctx2, **kw)
            assert st2["sparsify"]["csr_parts"] >= 2 and st1["sparsify"]["csr_parts"] == 1
            assert st2["sparsify"]["distinct_keys"] == st1["sparsify"]["distinct_keys"]
            if not sample:
                np.testing.assert_array_equal(S2, S1)
                np.testing.assert_array_equal(e2.cpu().numpy(), e1.cpu().numpy())
            else:
                assert abs(st2["sparsify"]["nnz"] - st1["sparsify"]["nnz"]) <= 1e-4 * st1["sparsify"]["nnz"]
                np.testing.assert_allclose(S2, S1, rtol=1e-4)
                assert H.principal_angles(e2.cpu().numpy()[:, :16], e1.cpu().numpy()[:, :16]) < 1e-2
    finally:
        ctx2.close()


@pytest.mark.parametrize("parts,sample", [(0, False), (6, False), (0, True), (6, True)])
def test_value_cache_gives_the_same_embedding(gpu, parts, sample):
    """The count pass's per-slot value cache gives the pipeline the CSR (so the
    embedding) that reading the degrees again gives, bit for bit, in one sort
    and by row ranges, with plain and effective-resistance (-sample 1) weights;
    auto mode gives the same."""
    import torch

    hg = L.chung_lu(20000, 2.1, 10.0, 500.0, 7)
    g = hg.to_device()
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 20.0)
    kw = dict(walk_len=10, rank=64, dim=32, sample_ratio=ratio, seed=3, csr_parts=parts, sample=sample, ctx=gpu)
    e0, S0, st0 = L.NetSMF(g, value_cache=-1, **kw)
    e1, S1, st1 = L.NetSMF(g, value_cache=1, **kw)
    assert st0["sparsify"]["value_cache"] == 0 and st1["sparsify"]["value_cache"] == 1
    assert st1["sparsify"]["nnz"] == st0["sparsify"]["nnz"]
    np.testing.assert_array_equal(S1, S0)
    assert torch.equal(e1, e0)
    e2, S2, _ = L.NetSMF(g, **kw)
    np.testing.assert_array_equal(S2, S0)
    assert torch.equal(e2, e0)


@pytest.mark.parametrize("rank,dim", [(64, 32), (256, 256)])
def test_value_cache_in_the_output_buffer(gpu, rank, dim):
    """A fresh context's first pipeline call by row ranges has no workspace
    room for the value cache: the pipeline's output buffer (written only by
    stage 2) holds the values of the first occupied slots it has room for (all
    of them at dim 256; at dim 32 the rest read the degrees), value_cache = 2,
    and the embedding is the one built without a cache, bit for bit."""
    import torch

    hg = L.chung_lu(20000, 2.1, 10.0, 500.0, 7)
    g = hg.to_device()
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 20.0)
    kw = dict(walk_len=10, rank=rank, dim=dim, sample_ratio=ratio, seed=3, csr_parts=6)
    e0, S0, st0 = L.NetSMF(g, value_cache=-1, ctx=gpu, **kw)
    c2 = L.Context(0)
    try:
        e1, S1, st1 = L.NetSMF(g, ctx=c2, **kw)
        keys = st1["sparsify"]["distinct_keys"]
        print(f"value_cache {st1['sparsify']['value_cache']}, distinct keys {keys}, output floats {hg.n * dim}")
        assert st1["sparsify"]["value_cache"] == 2
        if dim == 32:
            assert keys > hg.n * dim  # the cache covers the first occupied slots only
        else:
            assert keys <= hg.n * dim  # the whole cache
        assert st1["sparsify"]["nnz"] == st0["sparsify"]["nnz"]
        np.testing.assert_array_equal(S1, S0)
        assert torch.equal(e1, e0)
    finally:
        c2.close()


def test_csr_prefix_copy(gpu):
    """The entries of the first rows holding a fraction of the nonzeros come to
    the host alone (row_ptr whole); out-of-range counts are errors."""
    off, edg = H.golden_graph("cl2000")
    csr, _ = L.generate_trunc_log_matrix(dgraph(off, edg), 42, sample_ratio=0.9, ctx=gpu)
    rp, col, val = csr.to_numpy()
    rp2, col2, val2 = csr.to_numpy(nnz_fraction=0.3)
    np.testing.assert_array_equal(rp2, rp)
    nz = len(col2)
    assert 0 < nz < len(col) and nz == rp[np.searchsorted(rp, 0.3 * len(col))]
    np.testing.assert_array_equal(col2, col[:nz])
    np.testing.assert_array_equal(val2, val[:nz])
    with pytest.raises(L.LightNEError):
        gpu.check(gpu.lib.lightne_csr_copy_prefix(gpu.ptr, csr.handle, csr.nnz + 1, None, None))
    csr.free()
