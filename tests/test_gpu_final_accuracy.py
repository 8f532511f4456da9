"""End-to-end accuracy against FP64: the final embedding (propagation + dense
tall SVD, all 128 columns) of the FP32 and BF16 forms against the oracle's
FP64 propagation and library SVD of the same stage-2 embedding, and the
orthonormality of the FP32 left singular vectors U = B V S^-1."""
import numpy as np
import pytest

import lightne_amd as L
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

CASES = {  # n, beta, mean, maxdeg, seed, T, samples per directed edge
    "c1": (10_312, 2.1, 65.0, 3900.0, 1, 10, 50.0),
    "c2": (1_140_000, 2.0, 5.25, 2500.0, 2, 10, 10.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_final_embedding_vs_fp64(gpu, case):
    n, beta, mean, maxdeg, seed, T, per = CASES[case]
    hg = L.chung_lu(n, beta, mean, maxdeg, seed)
    g = hg.to_device()
    ratio = H.c1_sample_ratio(hg.n, hg.m, T, per)
    emb, S, _ = L.NetSMF(g, walk_len=T, rank=144, dim=128, sample_ratio=ratio, seed=1, ctx=gpu)
    a64 = emb.cpu().numpy().astype(np.float64)
    _, z64 = O.propagate(hg.offsets, hg.edges, a64, order=10, mu=0.2, theta=0.5)
    e64, S64 = O.embed_dense(z64, 128)
    for prec in (L.FP32, L.BF16):
        out, _ = L.chebyshev_expansion(emb, g, 128, 128, precision=prec, ctx=gpu)
        e = out.cpu().numpy()
        ang = H.principal_angles(e, e64)
        print(f"{case} precision={prec}: final-embedding angle vs FP64 {ang:.2e} rad "
              f"(z singular values {S64[0]:.3g} .. {S64[127]:.3g})")
        assert ang <= 1e-3
        # rows are unit vectors (get_embedding_dense normalises, SpectralPropagation.h:40)
        nrm = np.linalg.norm(e, axis=1)
        assert np.abs(nrm[nrm > 0] - 1).max() < 1e-5


def test_stage2_u_orthonormal_c2(gpu):
    """U = B V S^-1 is formed in FP32 from an FP64 Gram: its columns stay
    orthonormal to FP32 rounding amplified by sigma_0 / sigma_j."""
    n, beta, mean, maxdeg, seed, T, per = CASES["c2"]
    hg = L.chung_lu(n, beta, mean, maxdeg, seed)
    g = hg.to_device()
    csr, _ = L.generate_trunc_log_matrix(g, 1, walk_len=T, sample_ratio=H.c1_sample_ratio(hg.n, hg.m, T, per),
                                         ctx=gpu)
    svd = L.MKLRedSVD(csr, 144, seed=777, ctx=gpu).run()
    U = svd.matU.double()
    G = (U.T @ U).cpu().numpy()
    err = np.abs(G - np.eye(144)).max()
    print(f"max |U^T U - I| = {err:.2e}, sigma_0 / sigma_143 = {svd.S[0] / svd.S[143]:.3g}")
    assert err < 1e-4
