"""Ligra's compressed graph format (-c) on the host: the encoder against bytes
derived by hand from the format and against files the reference's own reader
decoded (reference-header harness fixtures); the oracle's decoder and its
neighbour lookup against the same fixtures; the binary CSR cache (CPU only)."""
import hashlib
import os

import numpy as np
import pytest

import lightne_amd as L
from lightne_amd import bytepd
from oracle import oracle as O
from tests import helpers as H


def _csr(adj_lists):
    n = len(adj_lists)
    off = np.zeros(n + 1, np.int64)
    for u, a in enumerate(adj_lists):
        off[u + 1] = off[u] + len(a)
    edges = np.array([v for a in adj_lists for v in a], np.uint32)
    return off, edges


def _record(img, v):
    n, m, voff, deg, e = O._bytepd_sections(img)
    return e[int(voff[v]):int(voff[v + 1])].tolist()


def test_encoder_bytes_by_hand():
    # vertex 0: [2, 1] -> degree 2, block index 0, first edge +2, delta (1 - 2) mod 2^32
    # vertex 1: [0, 2] -> first edge -1 (0x40 | 1), delta 2
    off, edg = _csr([[2, 1], [0, 2], [1, 0]])
    img = bytepd.encode(off, edg)
    assert bytepd.header(img) == (3, 6)
    assert _record(img, 0) == [2, 0, 0, 0, 0, 0, 0, 0, 0x02, 0xFF, 0xFF, 0xFF, 0xFF, 0x0F]
    assert _record(img, 1) == [2, 0, 0, 0, 0, 0, 0, 0, 0x41, 0x02]
    # first edge with more than 6 bits: 150 = 2 << 6 | 22 -> 0x80 | 22, then 2
    adj = [[] for _ in range(151)]
    adj[0], adj[150] = [150], [0]
    img = bytepd.encode(*_csr(adj))
    assert _record(img, 0) == [1, 0, 0, 0, 0, 0, 0, 0, 0x96, 0x02]
    assert _record(img, 150) == [1, 0, 0, 0, 0, 0, 0, 0, 0xD6, 0x02]
    assert _record(img, 1) == []  # degree 0: no record
    # two blocks: a star centre of degree 65; block 1 starts at byte 4 + 4 + 4 + 1 + 63
    adj = [list(range(1, 66))] + [[0] for _ in range(65)]
    img = bytepd.encode(*_csr(adj))
    r = _record(img, 0)
    assert r[:8] == [65, 0, 0, 0, 76, 0, 0, 0]
    assert r[8:12] == [0, 0, 0, 0] and r[12:76] == [0x01] * 64
    assert r[76:80] == [64, 0, 0, 0] and r[80:] == [0x81, 0x01]


def test_encoder_rejects_what_the_format_cannot_hold():
    with pytest.raises(ValueError):
        bytepd.encode(*_csr([[1, 1], [0, 0]]))  # a repeated neighbour encodes to no bytes
    with pytest.raises(ValueError):
        bytepd.encode(np.array([0, 1], np.int64), np.array([5], np.uint32))  # id >= n


@pytest.mark.parametrize("name", ["toy", "cl2000", "blogcatalog"])
def test_encoder_output_read_by_the_reference(name):
    """The harness read our file with gbbs_io::read_compressed_symmetric_graph and
    decoded it with the reference's block decoder: same graph back, and the file
    is byte-identical to the one it read."""
    off, edg = H.golden_graph(name)
    bp = dict(np.load(H.golden_path(f"bytepd_{name}.npz")))
    img = bytepd.encode(off, edg)
    assert img.size == int(bp["image_bytes"][0])
    assert hashlib.sha256(img.tobytes()).hexdigest() == str(bp["sha256_image"])
    if "image" in bp:
        np.testing.assert_array_equal(img, bp["image"])
    np.testing.assert_array_equal(bp["offsets"], off)
    if "decoded" in bp:
        np.testing.assert_array_equal(bp["decoded"], edg)
    else:
        assert hashlib.sha256(edg.tobytes()).hexdigest() == str(bp["sha256_decoded"])


@pytest.mark.parametrize("name", ["toy", "cl2000"])
def test_oracle_decoder_matches_reference_decode(name):
    off, edg = H.golden_graph(name)
    bp = dict(np.load(H.golden_path(f"bytepd_{name}.npz")))
    img = bytepd.encode(off, edg)
    doff, dedg = O.decode_bytepd(img)
    np.testing.assert_array_equal(doff, bp["offsets"])
    np.testing.assert_array_equal(dedg, bp["decoded"])


@pytest.mark.parametrize("name", ["toy", "cl2000", "blogcatalog"])
def test_compressed_lookup_rule_matches_reference(name):
    """The reference's get_ith_out_neighbor on compressed graphs, dumped at every
    index, equals neighbour i - 1 unless 64 | i -- the rule the oracle's and the
    device sampler's walks use for ith = 1."""
    off, edg = H.golden_graph(name)
    bp = dict(np.load(H.golden_path(f"bytepd_{name}.npz")))
    m = edg.shape[0]
    row = np.repeat(np.arange(off.shape[0] - 1), np.diff(off))
    i = np.arange(m) - off[row]
    rule = edg[off[row] + i - ((i % 64) != 0)]
    if "ith" in bp:
        np.testing.assert_array_equal(rule, bp["ith"])
    else:
        assert hashlib.sha256(rule.tobytes()).hexdigest() == str(bp["sha256_ith"])
    assert not np.array_equal(rule, edg)  # the rule is not the plain order


def test_literal_lookup_restatement_agrees_with_rule():
    off, edg = H.golden_graph("cl2000")
    img = bytepd.encode(off, edg)
    bp = dict(np.load(H.golden_path("bytepd_cl2000.npz")))
    deg = np.diff(off)
    for v in [int(np.argmax(deg)), 0, 7, 1999]:
        for i in range(int(deg[v])):
            assert O.ith_neighbor_bytepd(img, v, i) == bp["ith"][off[v] + i]


def test_oracle_sampler_compressed_lookup_changes_samples():
    off, edg = H.golden_graph("toy")
    k0, c0, _, s0 = O.sample(off, edg, 3, 10, 3.0, 0, threads=1, ith=0)
    k1, c1, _, s1 = O.sample(off, edg, 3, 10, 3.0, 0, threads=1, ith=1)
    assert s0["samples"] == s1["samples"]  # same RNG chains, same number of draws
    assert not (np.array_equal(k0, k1) and np.array_equal(c0, c1))


def test_binary_csr_round_trip(tmp_path):
    hg = L.chung_lu(3000, 2.1, 8.0, 200.0, 4)
    p = str(tmp_path / "g.bin")
    L.write_graph_bin(p, hg)
    back = L.read_graph_bin(p)
    np.testing.assert_array_equal(back.offsets, hg.offsets)
    np.testing.assert_array_equal(back.edges, hg.edges)
    assert os.path.getsize(p) == 24 + 8 * (hg.n + 1) + 4 * hg.m
    raw = np.fromfile(p, np.uint8)
    raw[:-4].tofile(str(tmp_path / "short.bin"))  # truncated
    with pytest.raises(L.LightNEError):
        L.read_graph_bin(str(tmp_path / "short.bin"))
    bad = raw.copy()
    bad[-4:] = 0xFF  # a neighbour id >= n
    bad.tofile(str(tmp_path / "bad.bin"))
    with pytest.raises(L.LightNEError):
        L.read_graph_bin(str(tmp_path / "bad.bin"))
    with pytest.raises(L.LightNEError):
        L.read_graph_bin(H.golden_path("manifest.json"))  # no magic
