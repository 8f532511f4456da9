"""Node-classification evaluation of an embedding: the protocol of
LightNE/predict.py (micro / macro F1 of one-vs-rest logistic regression over
repeated random train/test splits, each test node predicted with as many
labels as it truly has).

    python -m lightne_amd.predict --label labels.mat --embedding pro.emb \\
        --binary --dim 128 --seed 0 --C 10

Same flags and output lines as the reference script (predict.py:119-224);
embeddings are the raw FP32 n x dim files the CLI writes (``--binary``) or
``.npy``.  Host-side (scikit-learn); nothing here runs on the GPU.
"""
from __future__ import annotations

import argparse
import os

import numpy as np
import scipy.io
import scipy.sparse as sp


def top_k_indicator(scores: np.ndarray, y) -> sp.csr_matrix:
    """predict.py:27-43: row i predicts its num_labels(i) highest-scoring labels."""
    y = sp.csr_matrix(y)
    k = np.asarray(y.sum(axis=1)).ravel().astype(np.int64)
    order = np.argsort(-scores, axis=1, kind="stable")
    rows = np.repeat(np.arange(scores.shape[0]), k)
    cols = np.concatenate([order[i, :k[i]] for i in range(scores.shape[0])]) if k.sum() else np.zeros(0, np.int64)
    return sp.csr_matrix((np.ones(rows.shape[0], np.bool_), (rows, cols)), shape=y.shape)


def predict_cv(X: np.ndarray, y, train_ratio: float = 0.2, n_splits: int = 10, random_state: int = 0,
               C: float = 1.0, num_workers: int = 1):
    """predict.py:87-117: mean and std of micro / macro F1 (percent) over
    ``n_splits`` shuffled splits with ``train_ratio`` of the nodes for training."""
    from sklearn.linear_model import LogisticRegression
    from sklearn.metrics import f1_score
    from sklearn.model_selection import ShuffleSplit
    from sklearn.multiclass import OneVsRestClassifier

    y = sp.csr_matrix(y)
    micro, macro = [], []
    splits = ShuffleSplit(n_splits=n_splits, test_size=1 - train_ratio, random_state=random_state)
    for tr, te in splits.split(X):
        clf = OneVsRestClassifier(LogisticRegression(C=C, solver="liblinear"), n_jobs=num_workers)
        clf.fit(X[tr], y[tr])
        pred = top_k_indicator(clf.predict_proba(X[te]), y[te])
        micro.append(100.0 * f1_score(y[te], pred, average="micro"))
        macro.append(100.0 * f1_score(y[te], pred, average="macro"))
    return float(np.mean(micro)), float(np.mean(macro)), float(np.std(micro)), float(np.std(macro))


def load_label(path: str, variable_name: str = "group") -> sp.csr_matrix:
    """Labels from a .mat (variable ``group``), .npz (scipy sparse) or .npy file."""
    if path.endswith(".npz"):
        return sp.load_npz(path).tocsr().astype(np.bool_)
    if path.endswith(".npy"):
        return sp.csr_matrix(np.load(path)).astype(np.bool_)
    return sp.csr_matrix(scipy.io.loadmat(path)[variable_name]).astype(np.bool_)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--label", required=True)
    ap.add_argument("--embedding", required=True)
    ap.add_argument("--matfile-variable-name", default="group")
    ap.add_argument("--seed", type=int, required=True)
    ap.add_argument("--start-train-ratio", type=float, default=10)
    ap.add_argument("--stop-train-ratio", type=float, default=90)
    ap.add_argument("--num-train-ratio", type=int, default=9)
    ap.add_argument("--C", type=float, default=1.0)
    ap.add_argument("--num-split", type=int, default=10)
    ap.add_argument("--num-workers", type=int, default=1)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--norm", action="store_true")
    ap.add_argument("--binary", action="store_true")
    a = ap.parse_args(argv)
    y = load_label(a.label, a.matfile_variable_name)
    if a.binary:
        X = np.fromfile(a.embedding, dtype=np.float32).reshape(-1, a.dim)
    elif os.path.splitext(a.embedding)[1] == ".npy":
        X = np.load(a.embedding)
    else:
        raise SystemExit("embedding must be --binary raw FP32 or .npy")
    keep = np.asarray(y.getnnz(axis=1)) > 0  # nodes without labels are dropped (predict.py:185-191)
    X, y = X[keep], y[keep]
    if a.norm:
        X = X / np.maximum(np.linalg.norm(X, axis=1, keepdims=True), 1e-12)
    res = [predict_cv(X, y, train_ratio=tr / 100.0, n_splits=a.num_split, C=a.C, random_state=a.seed,
                      num_workers=a.num_workers)
           for tr in np.linspace(a.start_train_ratio, a.stop_train_ratio, a.num_train_ratio)]
    micro, macro = [r[0] for r in res], [r[1] for r in res]
    print(" ".join(str(x) for x in micro))
    print(" ".join(str(x) for x in macro))
    print(" & ".join("%.2f" % x for x in micro))
    print(" & ".join("%.2f" % x for x in macro))
    return micro, macro


if __name__ == "__main__":
    main()
