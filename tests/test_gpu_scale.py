"""Size-independent properties at the full R-MAT scale 24 workload (n = 16.8M,
2m = 521M directed edges, T = 5, M = 10 m, k = 144, d = 128, q = 2), where the
oracle cannot follow: the CSR built by row ranges and the single global sort
give the same embedding bit for bit, the BF16 form's stage 2 is bit-identical
to the FP32 form's (its BF16 sketch is exact), a repeated run is identical,
the sparsifier's counts obey their invariants, and the BF16 propagation stays
within its bound of the FP32 one (GPU)."""
import numpy as np
import pytest

import lightne_amd as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat24():
    import torch

    hg = L.rmat(24, 16.0, 0.57, 0.19, 0.19, 3)
    g = hg.to_device()
    ratio = float(np.float32(5.0 * hg.m / (5 * hg.n * np.log(hg.n))))  # 5 samples per directed edge
    print(f"R-MAT 24: n={hg.n} 2m={hg.m} sample_ratio={ratio}", flush=True)
    yield hg, g, ratio
    del g
    torch.cuda.empty_cache()


def test_full_size_stage1_and_2_properties(gpu, rmat24):
    import torch

    hg, g, ratio = rmat24
    kw = dict(walk_len=5, rank=144, dim=128, sample_ratio=ratio, seed=1, svd_seed=777, power_iters=2, ctx=gpu)
    e1, S1, st1 = L.NetSMF(g, **kw)
    sp = st1["sparsify"]
    print(f"samples={sp['samples']} distinct={sp['distinct_keys']} nnz={sp['nnz']}", flush=True)
    # -sample 0: every attempt is a sample; the CSR holds both triangles of the kept keys
    assert sp["samples"] == sp["attempts"] and sp["M_exact"] == sp["samples"]
    assert 2 * sp["kept_upper"] - hg.n <= sp["nnz"] <= 2 * sp["kept_upper"]
    assert sp["kept_upper"] <= sp["distinct_keys"] <= sp["samples"]
    assert np.all(np.diff(S1.astype(np.float64)) <= 0) and S1[-1] > 0
    e2, S2, st2 = L.NetSMF(g, csr_parts=8, **kw)
    assert st2["sparsify"]["csr_parts"] >= 2 and st2["sparsify"]["nnz"] == sp["nnz"]
    np.testing.assert_array_equal(S2, S1)
    assert torch.equal(e2, e1)
    del e2
    e3, S3, _ = L.NetSMF(g, precision=L.BF16, **kw)
    np.testing.assert_array_equal(S3, S1)
    assert torch.equal(e3, e1)
    del e3
    e4, S4, _ = L.NetSMF(g, **kw)
    assert torch.equal(e4, e1)


def test_full_size_bf16_propagation_bound(gpu, rmat24):
    import torch

    hg, g, _ = rmat24
    gen = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randn(hg.n, 128, generator=gen, device="cuda")
    _, z32, _ = L.propagate(a, g, order=10, mu=0.2, theta=0.5, precision=L.FP32, ctx=gpu)
    _, z16, _ = L.propagate(a, g, order=10, mu=0.2, theta=0.5, precision=L.BF16, ctx=gpu)
    err = float((torch.linalg.vector_norm(z16 - z32) / torch.linalg.vector_norm(z32)).item())
    print(f"stage-3 z, BF16 iterates vs FP32: relative error {err:.2e}", flush=True)
    assert err < 1e-3
