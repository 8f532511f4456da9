"""Size-independent properties at the full R-MAT scale 26 workload (n = 67.1M,
2m = 1.84B directed edges, T = 5, M = 10 m = 9.2e9 samples, k = 144, d = 128,
q = 1), where no host oracle can follow.  This is the only size whose
sparsifier (1.5e10 nonzeros) must be built in row ranges by itself and whose
propagation writes over its input: the embedding does not depend on the
number of row ranges, the first call (which sizes the workspace and runs the
deterministic sampler a second time after growing it) equals later calls on
the warm workspace, the BF16 form's stage 2 is the FP32 form's bit for bit,
the counts obey their invariants, the in-place propagation equals the
out-of-place one, and the BF16 propagation stays within its bound (GPU)."""
import numpy as np
import pytest

import lightne_amd as L
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat26(gpu):
    import torch

    gpu.release()  # what earlier tests left cached on the shared context
    torch.cuda.empty_cache()
    hg = L.rmat(26, 14.0, 0.57, 0.19, 0.19, 5)
    g = hg.to_device()
    ratio = float(np.float32(5.0 * hg.m / (5 * hg.n * np.log(hg.n))))  # 5 samples per directed edge
    print(f"R-MAT 26: n={hg.n} 2m={hg.m} sample_ratio={ratio}", flush=True)
    yield hg, g, ratio
    del g
    gpu.release()
    torch.cuda.empty_cache()


def test_c5_stage1_and_2_properties(gpu, rmat26):
    hg, g, ratio = rmat26
    kw = dict(walk_len=5, rank=144, dim=128, sample_ratio=ratio, seed=1, svd_seed=777, power_iters=1, ctx=gpu)
    e1, S1, st1 = L.NetSMF(g, **kw)  # empty workspace: grows, samples twice
    sp = st1["sparsify"]
    print(f"samples={sp['samples']} distinct={sp['distinct_keys']} nnz={sp['nnz']} ranges={sp['csr_parts']}",
          flush=True)
    assert sp["csr_parts"] >= 4 and sp["table_layout"] == 2  # one sort would need 529 GB
    assert sp["samples"] == sp["attempts"] and sp["M_exact"] == sp["samples"]
    assert 2 * sp["kept_upper"] - hg.n <= sp["nnz"] <= 2 * sp["kept_upper"]
    assert sp["kept_upper"] <= sp["distinct_keys"] <= sp["samples"]
    assert sp["distinct_keys"] <= 0.8 * sp["table_capacity"]
    assert np.all(np.diff(S1.astype(np.float64)) <= 0) and S1[-1] > 0
    h1 = e1.cpu().numpy()
    del e1
    # more, narrower row ranges: the same matrix, so the same embedding
    e2, S2, st2 = L.NetSMF(g, csr_parts=sp["csr_parts"] + 8, **kw)
    assert st2["sparsify"]["csr_parts"] > sp["csr_parts"] and st2["sparsify"]["nnz"] == sp["nnz"]
    np.testing.assert_array_equal(S2, S1)
    assert H.same_bits(e2, h1)
    del e2
    e3, S3, _ = L.NetSMF(g, precision=L.BF16, **kw)
    np.testing.assert_array_equal(S3, S1)
    assert H.same_bits(e3, h1)
    del e3
    e4, S4, st4 = L.NetSMF(g, **kw)  # warm workspace, one sampler run
    np.testing.assert_array_equal(S4, S1)
    assert H.same_bits(e4, h1)
    assert st4["sparsify"]["distinct_keys"] == sp["distinct_keys"]


def test_c5_propagation_in_place_and_bf16_bound(gpu, rmat26):
    import torch

    hg, g, _ = rmat26
    gpu.release()  # the stage-1 workspace is not needed here
    gen = torch.Generator(device="cuda").manual_seed(7)
    a = torch.randn(hg.n, 128, generator=gen, device="cuda")
    out, _ = L.chebyshev_expansion(a, g, 128, 128, ctx=gpu)
    ref = out.cpu().numpy()
    del out
    L.chebyshev_expansion(a, g, 128, 128, out=a, ctx=gpu)  # written over its input
    assert H.same_bits(a, ref)
    del ref
    a = torch.randn(hg.n, 128, generator=gen, device="cuda")
    y32, z32, _ = L.propagate(a, g, order=10, mu=0.2, theta=0.5, precision=L.FP32, ctx=gpu)
    z32h = z32.cpu().numpy()
    del y32, z32
    y16, z16, _ = L.propagate(a, g, order=10, mu=0.2, theta=0.5, precision=L.BF16, ctx=gpu)
    del y16
    num = den = 0.0
    for r in range(0, hg.n, 1 << 22):
        x = torch.from_numpy(z32h[r:r + (1 << 22)]).cuda()
        num += float(torch.sum((z16[r:r + (1 << 22)] - x).double() ** 2))
        den += float(torch.sum(x.double() ** 2))
    err = (num / den) ** 0.5
    print(f"R-MAT 26 stage-3 z, BF16 iterates vs FP32: relative error {err:.2e}", flush=True)
    assert err < 1e-4
