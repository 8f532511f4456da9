"""Writes the stage-1 golden vectors under tests/golden/ from the reference-header
harness (oracle/_ref/ref_harness, built by `make -C oracle ref` from the
reference's own pbbslib / ligra / par_table headers).

Run in the build container only (it needs /root/reference and the harness):
    python tests/golden/make_golden.py

Inputs are seeded graphs (written as AdjacencyGraph files for the harness) and
the BlogCatalog network shipped with the reference (data_bin/blogcatalog.mat,
read with scipy.io.loadmat, which executes nothing from the file).  Outputs:
    graph_<name>.npz          offsets/edges of every input graph
    golden_<case>.npz         scalars and either the full key->value table or a
                              SHA-256 digest plus a strided subset
    manifest.json             the case list
    blogcatalog_labels.npz    BlogCatalog's node labels
    bytepd_<name>.npz         for graphs written in Ligra's compressed format by
                              lightne_amd.bytepd: the reference reader's decode
                              and its get_ith_out_neighbor at every index (full,
                              or a digest), and the file image of the toy graph

Compressed cases (``"compressed": true`` in the manifest) run the harness on
the compressed file (``-c``), so their walks use the reference's neighbour
lookup for that format.  ``python tests/golden/make_golden.py compressed``
rewrites only those.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from lightne_amd import api  # noqa: E402  (generators and AdjacencyGraph writer)

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
BLOG = "/root/reference/data_bin/blogcatalog.mat"
FULL_LIMIT = 20000  # tables up to this many keys are stored in full


def toy_graph() -> api.HostGraph:
    # triangle 0-1-2, star centred at 3 (leaves 4..7), path 7-8-9, bridge 2-3,
    # neighbour lists deliberately NOT sorted (input order matters)
    adj = {0: [2, 1], 1: [0, 2], 2: [3, 1, 0], 3: [7, 4, 2, 5, 6], 4: [3], 5: [3], 6: [3],
           7: [8, 3], 8: [7, 9], 9: [8]}
    n = 10
    off = np.zeros(n + 1, np.int64)
    for u in range(n):
        off[u + 1] = off[u] + len(adj[u])
    edges = np.array([v for u in range(n) for v in adj[u]], np.uint32)
    return api.HostGraph(off, edges)


def blog_graph() -> api.HostGraph:
    import scipy.io

    A = scipy.io.loadmat(BLOG)["network"].tocsr()
    A.sort_indices()
    A.setdiag(0)
    A.eliminate_zeros()
    return api.HostGraph(A.indptr.astype(np.int64), A.indices.astype(np.uint32))


def run_harness(adj_path, seed, T, sample, ratio, out_path, compressed=False):
    env = dict(os.environ, NUM_THREADS="1")
    subprocess.run([HARNESS, adj_path, str(seed), str(T), str(sample), repr(ratio), "1.0", "0", "1", out_path]
                   + (["-c"] if compressed else []), check=True, env=env, stdout=subprocess.DEVNULL)
    with open(out_path, "rb") as f:
        raw = f.read()
    magic, nk = np.frombuffer(raw[:16], np.uint64)
    assert int(magic) == 0x31444C4F47454E4C
    nk = int(nk)
    o = 16
    keys = np.frombuffer(raw[o:o + 8 * nk], np.uint64).copy(); o += 8 * nk
    vals = np.frombuffer(raw[o:o + 8 * nk], np.float64).copy(); o += 8 * nk
    tl = np.frombuffer(raw[o:o + 8 * nk], np.float64).copy(); o += 8 * nk
    sc = np.frombuffer(raw[o:o + 48], np.float64).copy(); o += 48
    su = np.frombuffer(raw[o:o + 24], np.uint64).copy()
    return keys, vals, tl, sc, su


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def write_blog_labels():
    """BlogCatalog's node labels (variable ``group``, 10,312 x 39) as a scipy
    sparse .npz, for the node-classification check (LightNE/predict.py)."""
    import scipy.io
    import scipy.sparse as sp

    y = sp.csr_matrix(scipy.io.loadmat(BLOG)["group"]).astype(np.bool_)
    sp.save_npz(os.path.join(HERE, "blogcatalog_labels.npz"), y, compressed=True)


def read_csr_dump(path):
    raw = np.fromfile(path, np.uint8)
    n, m = (int(x) for x in np.frombuffer(raw[:16].tobytes(), np.uint64))
    o = 16
    off = np.frombuffer(raw[o:o + 8 * (n + 1)].tobytes(), np.uint64).astype(np.int64); o += 8 * (n + 1)
    dec = np.frombuffer(raw[o:o + 4 * m].tobytes(), np.uint32).copy(); o += 4 * m
    ith = np.frombuffer(raw[o:o + 4 * m].tobytes(), np.uint32).copy()
    return off, dec, ith


CASES = [
    # name, graph, seed, T, sample, sample_ratio, compressed
    ("toy_t10_s0", "toy", 1, 10, 0, 3.0, False),
    ("toy_t1_s0", "toy", 7, 1, 0, 5.5, False),
    ("toy_t3_s1", "toy", 99, 3, 1, 4.0, False),
    ("cl200_t10_s0", "cl200", 12345, 10, 0, 0.35, False),
    ("cl200_t5_s1", "cl200", 2**63 + 5, 5, 1, 2.0, False),
    ("cl2000_t10_s0", "cl2000", 42, 10, 0, 0.9, False),
    ("cl2000_t5_s1", "cl2000", 17, 5, 1, 1.3, False),
    ("blog_t10_s0", "blogcatalog", 1, 10, 0, 1.0, False),
    ("blog_t10_s1", "blogcatalog", 5, 10, 1, 2.0, False),
    # compressed-format input (-c): walks use that format's neighbour lookup
    ("ctoy_t10_s0", "toy", 3, 10, 0, 3.0, True),
    ("ccl2000_t10_s0", "cl2000", 42, 10, 0, 0.9, True),
    ("ccl2000_t5_s1", "cl2000", 23, 5, 1, 1.3, True),
    ("cblog_t10_s0", "blogcatalog", 1, 10, 0, 1.0, True),
]


def main(only_compressed=False):
    assert os.path.exists(HARNESS), "build the harness first: make -C oracle ref"
    from lightne_amd import bytepd

    if not only_compressed:
        write_blog_labels()
    graphs = {
        "toy": toy_graph(),
        "cl200": api.chung_lu(200, 2.1, 6.0, 40.0, 3),
        "cl2000": api.chung_lu(2000, 2.1, 10.0, 150.0, 11),
        "blogcatalog": blog_graph(),
    }
    cases = [c for c in CASES if c[6] or not only_compressed]
    manifest = {"cases": []}
    if only_compressed:
        with open(os.path.join(HERE, "manifest.json")) as f:
            manifest["cases"] = [r for r in json.load(f)["cases"] if not r.get("compressed")]
    with tempfile.TemporaryDirectory() as td:
        for name, g in graphs.items():
            if not only_compressed:
                np.savez_compressed(os.path.join(HERE, f"graph_{name}.npz"), offsets=g.offsets, edges=g.edges)
            api.write_adjacency_graph(os.path.join(td, f"{name}.adj"), g)
            bytepd.write(os.path.join(td, f"{name}.cmp"), g.offsets, g.edges)
        dumped = set()
        for case, gname, seed, T, sample, ratio, comp in cases:
            out = os.path.join(td, case + ".bin")
            src = os.path.join(td, f"{gname}.cmp" if comp else f"{gname}.adj")
            keys, vals, tl, sc, su = run_harness(src, seed, T, sample, ratio, out, comp)
            if comp and gname not in dumped:
                dumped.add(gname)
                off, dec, ith = read_csr_dump(out + ".csr")
                img = np.fromfile(src, np.uint8)
                bp = dict(offsets=off, sha256_image=np.array(digest(img)), image_bytes=np.array([img.size]))
                if dec.shape[0] <= 100000:
                    bp.update(decoded=dec, ith=ith)
                else:
                    bp.update(sha256_decoded=np.array(digest(dec)), sha256_ith=np.array(digest(ith)))
                if img.size <= 4096:
                    bp.update(image=img)
                np.savez_compressed(os.path.join(HERE, f"bytepd_{gname}.npz"), **bp)
            rec = dict(case=case, graph=gname, seed=str(seed), walk_len=T, sample=sample, sample_ratio=ratio,
                       nkeys=int(keys.shape[0]), kept=int(np.isfinite(tl).sum()))
            if comp:
                rec["compressed"] = True
            payload = dict(scalars=sc, scalars_u=su, nkeys=np.array([keys.shape[0]], np.int64))
            if keys.shape[0] <= FULL_LIMIT:
                payload.update(keys=keys, vals=vals, trunclog=tl)
                rec["full"] = True
            else:
                idx = np.arange(0, keys.shape[0], 997)
                payload.update(sub_idx=idx, sub_keys=keys[idx], sub_vals=vals[idx], sub_trunclog=tl[idx])
                rec["full"] = False
                rec["sha256_keys_vals"] = digest(keys, vals)
                rec["sha256_trunclog"] = digest(tl)
            np.savez_compressed(os.path.join(HERE, f"golden_{case}.npz"), **payload)
            manifest["cases"].append(rec)
            print(case, rec["nkeys"], rec["kept"])
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main(only_compressed=len(sys.argv) > 1 and sys.argv[1] == "compressed")
