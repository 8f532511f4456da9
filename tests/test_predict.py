"""The node-classification harness (LightNE/predict.py protocol), CPU only."""
import numpy as np
import scipy.sparse as sp

from lightne_amd import predict as P


def test_top_k_indicator_predicts_as_many_labels_as_true():
    scores = np.array([[0.1, 0.9, 0.5, 0.3], [0.7, 0.2, 0.6, 0.1]])
    y = sp.csr_matrix(np.array([[1, 0, 1, 0], [0, 0, 0, 1]], np.bool_))
    pred = P.top_k_indicator(scores, y).toarray()
    assert pred.tolist() == [[False, True, True, False], [True, False, False, False]]


def test_predict_cv_separable_and_random():
    rng = np.random.default_rng(0)
    n, c = 600, 4
    lab = rng.integers(0, c, n)
    y = sp.csr_matrix((np.ones(n, np.bool_), (np.arange(n), lab)), shape=(n, c))
    X = np.eye(c)[lab] * 5 + 0.1 * rng.standard_normal((n, c))
    mi, ma, _, _ = P.predict_cv(X, y, train_ratio=0.5, n_splits=2, C=1.0)
    assert mi > 99 and ma > 99
    mi_r, _, _, _ = P.predict_cv(rng.standard_normal((n, 8)), y, train_ratio=0.5, n_splits=2)
    assert mi_r < 40  # chance level is 25


def test_cli_on_the_blogcatalog_labels(tmp_path):
    from tests import helpers as H

    y = P.load_label(H.golden_path("blogcatalog_labels.npz"))
    assert y.shape == (10312, 39)
    emb = np.random.default_rng(1).standard_normal((10312, 16)).astype(np.float32)
    f = tmp_path / "e.emb"
    emb.tofile(f)
    micro, macro = P.main(["--label", H.golden_path("blogcatalog_labels.npz"), "--embedding", str(f), "--binary",
                           "--dim", "16", "--seed", "0", "--num-train-ratio", "1", "--start-train-ratio", "50",
                           "--stop-train-ratio", "50", "--num-split", "1"])
    assert len(micro) == 1 and 0 < micro[0] < 30  # a random embedding predicts near the label prior
