"""Device decoder of Ligra's compressed graph format (-c) and the runner's
compressed and binary inputs, against the source CSR, the oracle's decoder
and the API (GPU)."""
import os
import subprocess

import numpy as np
import pytest

import lightne_amd as L
from lightne_amd import bytepd
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu
CLI = os.path.join(H.ROOT, "lightne_amd", "lightne")


def _decode(img_np, ctx):
    import torch

    return L.decode_compressed_graph(torch.from_numpy(img_np).cuda(), ctx=ctx)


def _check_same(g, off, edg):
    np.testing.assert_array_equal(g.offsets.cpu().numpy(), off)
    np.testing.assert_array_equal(g.edges.cpu().numpy().view(np.uint32), edg)


@pytest.mark.parametrize("name", ["toy", "cl200", "cl2000", "blogcatalog"])
def test_decode_fixture_graphs(gpu, name):
    off, edg = H.golden_graph(name)
    img = bytepd.encode(off, edg)
    g = _decode(img, gpu)
    assert g.ith_rule == L.ITH_BYTEPD and (g.n, g.m) == (off.shape[0] - 1, edg.shape[0])
    _check_same(g, off, edg)
    if name in ("toy", "cl200"):
        doff, dedg = O.decode_bytepd(img)
        _check_same(g, doff, dedg)


def _hard_graph(seed=5, n=1 << 21):
    """Block boundaries (degrees 1, 63, 64, 65, 128, 129, 5000), unsorted lists
    (wrapped 5-byte differences), first edges up to 2^21 away (4 bytes),
    isolated vertices."""
    rng = np.random.default_rng(seed)
    hubs = {0: 5000, 1: 63, 2: 64, 3: 65, 4: 128, 5: 129, n - 1: 1, n // 2: 700}
    adj = {}
    for h, d in hubs.items():
        nb = rng.choice(np.arange(6, n - 1), size=d, replace=False)
        if h % 2 == 0:
            nb = np.sort(nb)  # sorted lists on even hubs, random order on odd ones
        adj[h] = [int(x) for x in nb]
    # symmetric: every neighbour of a hub lists the hub back
    back = {}
    for h, lst in adj.items():
        for v in lst:
            back.setdefault(v, []).append(h)
    for v, lst in back.items():
        if v in adj:
            adj[v] = adj[v] + [x for x in lst if x not in adj[v]]
        else:
            adj[v] = lst[::-1]
    deg = np.zeros(n, np.int64)
    for v, lst in adj.items():
        deg[v] = len(lst)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(deg, out=off[1:])
    edg = np.zeros(off[-1], np.uint32)
    for v, lst in adj.items():
        edg[off[v]:off[v + 1]] = lst
    return off, edg


def test_decode_block_boundaries_and_long_varints(gpu):
    off, edg = _hard_graph()
    img = bytepd.encode(off, edg)
    g = _decode(img, gpu)
    _check_same(g, off, edg)
    # the same file decoded twice gives the same arrays (no state kept between calls)
    _check_same(_decode(img, gpu), off, edg)


def _bad(img, ctx, match):
    with pytest.raises(L.LightNEError, match=match):
        _decode(img, ctx)


def test_decode_rejects_malformed_files(gpu):
    """Every malformed file is an error with a reason; the decoder's reads are
    bounds-checked, so none of these reads outside the image."""
    import torch

    off, edg = H.golden_graph("cl2000")
    img = bytepd.encode(off, edg)
    n, m, voff, deg, e = O._bytepd_sections(img)
    head = 24 + 8 * (n + 1) + 4 * n
    _bad(img[:-50].copy(), gpu, "BAD_GRAPH")  # the last records run past the end
    _bad(img[:head - 4].copy(), gpu, "shorter than its header")
    v = int(np.argmax(deg))
    x = img.copy()
    x[head + int(voff[v])] ^= 1  # the record's own degree no longer matches
    _bad(x, gpu, "degree")
    x = img.copy()
    d0 = 24 + 8 * (n + 1)
    x[d0:d0 + 4] = np.frombuffer(np.uint32(deg[0] + 1).tobytes(), np.uint8)  # degree array != records, sum != m
    _bad(x, gpu, "BAD_GRAPH")
    u = int(np.nonzero(deg > 0)[0][0])
    first = head + int(voff[u]) + 4 * (1 + (int(deg[u]) - 1) // 64) + 4  # u's first edge
    x = img.copy()
    x[first:first + 6] = 0xFF  # a varint that does not end within 5 bytes
    _bad(x, gpu, "BAD_GRAPH")
    x = img.copy()
    x[first:first + 3] = [0x80, 0x80, 0x01]  # first neighbour u + 2^20 >= n
    _bad(x, gpu, "neighbour id")
    t = torch.from_numpy(img).cuda()
    buf = torch.zeros(img.size + 8, dtype=torch.uint8, device="cuda")
    buf[1:1 + img.size] = t
    with pytest.raises(L.LightNEError, match="aligned"):
        gpu.check(gpu.lib.lightne_graph_decode_compressed(gpu.ptr, buf.data_ptr() + 1, img.size, n, m,
                                                         buf.data_ptr(), buf.data_ptr()))
    offd = torch.empty(n + 2, dtype=torch.int64, device="cuda")
    edgd = torch.empty(m, dtype=torch.int32, device="cuda")
    with pytest.raises(L.LightNEError, match="header"):
        gpu.check(gpu.lib.lightne_graph_decode_compressed(gpu.ptr, t.data_ptr(), img.size, n + 1, m,
                                                         offd.data_ptr(), edgd.data_ptr()))
    # the context still works after the errors
    _check_same(_decode(img, gpu), off, edg)


def test_netsmf_on_compressed_graph_uses_reference_lookup(gpu):
    """A decoded -c graph samples with the compressed lookup; the same CSR with the
    plain lookup gives other counts (the two inputs differ in the reference too)."""
    off, edg = H.golden_graph("cl2000")
    g = _decode(bytepd.encode(off, edg), gpu)
    t1 = L.sparsify_counts(g, 9, 10, False, sample_ratio=0.9, ctx=gpu)
    k1, c1, _, _ = O.sample(off, edg, 9, 10, 0.9, 0, threads=1, ith=1)
    np.testing.assert_array_equal(t1.keys, k1)
    np.testing.assert_array_equal(t1.counts.astype(np.uint64), c1)
    t0 = L.sparsify_counts(L.HostGraph(off, edg).to_device(), 9, 10, False, sample_ratio=0.9, ctx=gpu)
    assert not np.array_equal(t0.keys, t1.keys) or not np.array_equal(t0.counts, t1.counts)


def test_cli_compressed_and_binary_inputs(gpu, tmp_path):
    hg = L.chung_lu(3000, 2.1, 12.0, 300.0, 5)
    adj, cmp_, binf = str(tmp_path / "g.adj"), str(tmp_path / "g.cmp"), str(tmp_path / "g.bin")
    L.write_adjacency_graph(adj, hg)
    L.write_compressed_graph(cmp_, hg)
    L.write_graph_bin(binf, hg)
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 20.0)
    common = ["-walklen", "10", "-rank", "64", "-dim", "32", "-order", "10", "-sample_ratio", repr(ratio),
              "-seed", "9", "-rounds", "1"]
    outs = {}
    for tag, extra, path in [("adj", [], adj), ("bin", [], binf), ("cmp", ["-c"], cmp_),
                             ("cmp_plain", ["-c", "-ith_rule", "plain"], cmp_)]:
        ne = str(tmp_path / f"ne_{tag}.emb")
        r = subprocess.run([CLI] + common + ["-ne_out", ne, "-s"] + extra + [path], capture_output=True, text=True,
                           timeout=240)
        assert r.returncode == 0, r.stderr
        outs[tag] = np.fromfile(ne, np.float32).reshape(-1, 32)
    np.testing.assert_array_equal(outs["bin"], outs["adj"])
    np.testing.assert_array_equal(outs["cmp_plain"], outs["adj"])  # same CSR, same lookup
    g = L.read_compressed_graph(cmp_, ctx=gpu)
    emb, S, _ = L.NetSMF(g, walk_len=10, rank=64, dim=32, sample_ratio=ratio, seed=9)
    np.testing.assert_array_equal(outs["cmp"], emb.cpu().numpy())
    assert not np.array_equal(outs["cmp"], outs["adj"])
