"""Shared test helpers: golden fixtures, graphs, subspace comparisons."""
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["cases"]


def golden_path(name):
    return os.path.join(GOLDEN, name)


def golden_graph(name):
    d = np.load(os.path.join(GOLDEN, f"graph_{name}.npz"))
    return d["offsets"].astype(np.int64), d["edges"].astype(np.uint32)


def golden_case(case):
    return dict(np.load(os.path.join(GOLDEN, f"golden_{case}.npz")))


def c1_sample_ratio(n, m_directed, T=10, per_edge=50.0):
    """sample_ratio giving `per_edge` samples per directed edge: M = T n ln(n) ratio."""
    return float(np.float32(per_edge * m_directed / (T * n * np.log(n))))


def sort_rows(rp, col, val):
    """The CSR with every row's entries in column order (the device keeps
    them in its table's canonical slot order; the oracle sorts by column)."""
    rows = np.repeat(np.arange(len(rp) - 1, dtype=np.int64), np.diff(rp))
    o = np.lexsort((col, rows))
    return rp, col[o], val[o]


def principal_angles(A, B):
    """Largest principal angle (rad) between the column spaces of A and B."""
    qa = np.linalg.qr(np.asarray(A, np.float64))[0]
    qb = np.linalg.qr(np.asarray(B, np.float64))[0]
    s = np.linalg.svd(qa.T @ qb, compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1.0, 1.0)))


def sign_align(X, ref):
    """Flip column signs of X to match ref (singular vectors are defined up to sign)."""
    s = np.sign(np.sum(np.asarray(X, np.float64) * np.asarray(ref, np.float64), axis=0))
    s[s == 0] = 1.0
    return X * s[None, :]


def same_bits(A, B, chunk=1 << 20) -> bool:
    """A and B (torch device tensors or host arrays, float32) hold the same
    bits, compared in row chunks on the device."""
    import torch

    def dev(X, r):
        x = X[r:r + chunk]
        return torch.from_numpy(np.ascontiguousarray(x)).cuda() if isinstance(x, np.ndarray) else x

    if tuple(A.shape) != tuple(B.shape):
        return False
    for r in range(0, A.shape[0], chunk):
        if not torch.equal(dev(A, r).view(torch.int32), dev(B, r).view(torch.int32)):
            return False
    return True
