"""lightne_dense_gram_noalloc through the C-ABI: the same Gram as
lightne_dense_gram from a fixed device array instead of device allocations."""
import ctypes as C

import numpy as np
import pytest
import torch

import lightne_amd as L

pytestmark = pytest.mark.gpu


def _gram(ctx, fn, X, Y, sym, n=None):
    k1, k2 = X.shape[1], Y.shape[1]
    G = np.full((k1, k2), np.nan)
    rc = fn(ctx.ptr, X.data_ptr(), Y.data_ptr(), X.shape[0] if n is None else n, k1, k2, int(sym), G.ctypes.data)
    return rc, G


def test_gram_noalloc_matches_allocating_gram(gpu):
    lib = L._lib.load()
    ctx = L.Context(0)  # a fresh context: no cached buffers of its own
    try:
        gen = torch.Generator(device="cuda").manual_seed(11)
        X1 = torch.randn(300_000, 144, generator=gen, device="cuda")
        X2 = torch.randn(5_000, 64, generator=gen, device="cuda")
        Y2 = torch.randn(5_000, 48, generator=gen, device="cuda")
        X3 = torch.randn(70_000, 256, generator=gen, device="cuda")
        torch.cuda.synchronize()
        # first launches of each kernel variant (the runtime sizes its
        # per-queue private-segment memory when a kernel first needs it)
        for A, B, sym in [(X1, X1, True), (X2, Y2, False), (X3, X3, True)]:
            rc, _ = _gram(ctx, lib.lightne_dense_gram_noalloc, A, B, sym)
            assert rc == 0, lib.lightne_last_error(ctx.ptr)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info(0)[0]
        # two calls in a row, the second with a smaller input
        rc1, G1 = _gram(ctx, lib.lightne_dense_gram_noalloc, X1, X1, True)
        rc2, G2 = _gram(ctx, lib.lightne_dense_gram_noalloc, X2, Y2, False)
        rc3, G3 = _gram(ctx, lib.lightne_dense_gram_noalloc, X3, X3, True)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info(0)[0]
        assert (rc1, rc2, rc3) == (0, 0, 0), lib.lightne_last_error(ctx.ptr)
        assert free1 == free0, (free0, free1)  # nothing was allocated on the device
        # the allocating entry point on the shared context
        _, R1 = _gram(gpu, lib.lightne_dense_gram, X1, X1, True)
        _, R2 = _gram(gpu, lib.lightne_dense_gram, X2, Y2, False)
        _, R3 = _gram(gpu, lib.lightne_dense_gram, X3, X3, True)
        np.testing.assert_array_equal(G1, R1)  # same blocks, same summation order
        np.testing.assert_array_equal(G2, R2)
        # 256 wide: fewer blocks in the scratch, FP64 sums in another order
        assert np.abs(G3 - R3).max() <= 1e-12 * np.abs(R3).max()
        ref = (X2.double().T @ Y2.double()).cpu().numpy()
        assert np.abs(G2 - ref).max() <= 1e-12 * np.abs(ref).max()
        # widths beyond the scratch's layout are refused with a message
        X4 = torch.randn(100, 272, device="cuda")
        rc4, _ = _gram(ctx, lib.lightne_dense_gram_noalloc, X4, X4, True)
        assert rc4 == L._lib.ERR_BAD_ARG
        assert b"multiples of 16" in lib.lightne_last_error(ctx.ptr)
        # no rows: G = 0
        rc5, G5 = _gram(ctx, lib.lightne_dense_gram_noalloc, X2, Y2, False, n=0)
        assert rc5 == 0, lib.lightne_last_error(ctx.ptr)
        assert not G5.any()
        # a NULL panel is refused
        G6 = np.zeros((64, 48))
        assert lib.lightne_dense_gram_noalloc(ctx.ptr, None, Y2.data_ptr(), 10, 64, 48, 0, G6.ctypes.data) == \
            L._lib.ERR_BAD_ARG
    finally:
        ctx.close()
