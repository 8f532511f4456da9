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


def run_harness(adj_path, seed, T, sample, ratio, out_path):
    env = dict(os.environ, NUM_THREADS="1")
    subprocess.run([HARNESS, adj_path, str(seed), str(T), str(sample), repr(ratio), "1.0", "0", "1", out_path],
                   check=True, env=env, stdout=subprocess.DEVNULL)
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


def main():
    assert os.path.exists(HARNESS), "build the harness first: make -C oracle ref"
    write_blog_labels()
    graphs = {
        "toy": toy_graph(),
        "cl200": api.chung_lu(200, 2.1, 6.0, 40.0, 3),
        "cl2000": api.chung_lu(2000, 2.1, 10.0, 150.0, 11),
        "blogcatalog": blog_graph(),
    }
    cases = [
        # name, graph, seed, T, sample, sample_ratio
        ("toy_t10_s0", "toy", 1, 10, 0, 3.0),
        ("toy_t1_s0", "toy", 7, 1, 0, 5.5),
        ("toy_t3_s1", "toy", 99, 3, 1, 4.0),
        ("cl200_t10_s0", "cl200", 12345, 10, 0, 0.35),
        ("cl200_t5_s1", "cl200", 2**63 + 5, 5, 1, 2.0),
        ("cl2000_t10_s0", "cl2000", 42, 10, 0, 0.9),
        ("cl2000_t5_s1", "cl2000", 17, 5, 1, 1.3),
        ("blog_t10_s0", "blogcatalog", 1, 10, 0, 1.0),
        ("blog_t10_s1", "blogcatalog", 5, 10, 1, 2.0),
    ]
    manifest = {"cases": []}
    with tempfile.TemporaryDirectory() as td:
        for name, g in graphs.items():
            np.savez_compressed(os.path.join(HERE, f"graph_{name}.npz"), offsets=g.offsets, edges=g.edges)
            api.write_adjacency_graph(os.path.join(td, f"{name}.adj"), g)
        for case, gname, seed, T, sample, ratio in cases:
            out = os.path.join(td, case + ".bin")
            keys, vals, tl, sc, su = run_harness(os.path.join(td, f"{gname}.adj"), seed, T, sample, ratio, out)
            rec = dict(case=case, graph=gname, seed=str(seed), walk_len=T, sample=sample, sample_ratio=ratio,
                       nkeys=int(keys.shape[0]), kept=int(np.isfinite(tl).sum()))
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
    main()
