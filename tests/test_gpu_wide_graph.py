"""Graphs with 2^32 or more directed edges (the reference's 64-bit edge
offsets, ligra/macros.h:44-69): the walks read the graph's 64-bit row pointers
and the sampler launches its lanes in slices; the SpMM over the graph splits
its waves over several launches.

The 64-bit walk is pinned bit for bit against the oracle on every golden case
(forced with wide_offsets = 1).  At full size -- a circulant graph of 2^26
vertices and degree 66, 4.43e9 directed edges -- the checks are what the graph
fixes: every sample's endpoints lie within 33 r ring steps of each other (a
walk that read a truncated row pointer lands anywhere), the sample count is
the binomial count of all 2m edges (a lost or repeated slice of 2^31 lanes
moves it by half), and the propagation maps each Fourier mode cos(2 pi j v / n)
to a known multiple of itself."""
import math

import numpy as np
import pytest

import lightne_amd as L
from lightne_amd import _lib
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

N = 1 << 26
HALF = 33  # neighbours v +- 1 .. v +- 33 (mod n), in that order


@pytest.mark.parametrize("rec", H.manifest(), ids=lambda r: r["case"])
def test_wide_offsets_walks_match_oracle(gpu, rec):
    off, edg = H.golden_graph(rec["graph"])
    ith = 1 if rec.get("compressed") else 0
    if ith:  # the device decode of the encoded file, as `lightne -c` reads it
        import torch
        from lightne_amd import bytepd

        g = L.decode_compressed_graph(torch.from_numpy(bytepd.encode(off, edg)).cuda(), ctx=gpu)
    else:
        g = L.HostGraph(off, edg).to_device()
    t = L.sparsify_counts(g, int(rec["seed"]), rec["walk_len"], bool(rec["sample"]),
                          sample_ratio=rec["sample_ratio"], wide_offsets=1, ctx=gpu)
    keys, cnt, wgh, st = O.sample(off, edg, int(rec["seed"]), rec["walk_len"], rec["sample_ratio"], rec["sample"],
                                  threads=1, ith=ith)
    np.testing.assert_array_equal(t.keys, keys)
    np.testing.assert_array_equal(t.counts.astype(np.uint64), cnt)
    assert t.stats["samples"] == st["samples"] and t.stats["walk_steps"] == st["walk_steps"]


def test_wide_offsets_flag_checked(gpu):
    g = L.chung_lu(500, 2.1, 6.0, 40.0, 1).to_device()
    with pytest.raises(L.LightNEError) as e:
        L.sparsify_counts(g, 1, 5, sample_ratio=0.1, wide_offsets=3, ctx=gpu)
    assert e.value.code == _lib.ERR_BAD_ARG


@pytest.fixture(scope="module")
def circulant():
    import torch

    deg = 2 * HALF
    m = N * deg
    assert m >= 1 << 32
    print(f"circulant graph: n = {N}, degree {deg}, {m} directed edges", flush=True)
    off = torch.arange(N + 1, device="cuda", dtype=torch.int64) * deg
    edg = torch.empty(m, device="cuda", dtype=torch.int32)
    steps = torch.cat([torch.arange(1, HALF + 1), -torch.arange(1, HALF + 1)]).to("cuda")
    rows = 1 << 22
    for r0 in range(0, N, rows):
        v = torch.arange(r0, r0 + rows, device="cuda", dtype=torch.int64)
        edg[r0 * deg:(r0 + rows) * deg] = ((v[:, None] + steps[None, :]) % N).reshape(-1).to(torch.int32)
    torch.cuda.synchronize()
    yield L.DeviceGraph(off, edg, N, m)
    del off, edg
    torch.cuda.empty_cache()


def test_sampler_above_2_32_edges(gpu, circulant):
    g = circulant
    T = 10
    ratio = H.c1_sample_ratio(N, g.m, T, 0.05)  # about 0.05 samples per directed edge
    t = L.sparsify_counts(g, 7, T, sample_ratio=ratio, ctx=gpu)
    st = t.stats
    per_edge = st["exp_rv_rounded"] + st["fractional_bits"]
    mean = g.m * per_edge
    sd = math.sqrt(g.m * st["fractional_bits"] * (1 - st["fractional_bits"]))
    print(f"samples {st['samples']} (expected {mean:.0f} +- {sd:.0f}), keys {len(t.keys)}, "
          f"steps {st['walk_steps']}", flush=True)
    assert abs(st["samples"] - mean) <= 8 * sd + 1
    assert int(t.counts.astype(np.uint64).sum()) == st["samples"] == st["attempts"]
    # r is uniform in [1, T]: r - 1 steps per sample, (T - 1) / 2 on average
    assert abs(st["walk_steps"] / st["samples"] - (T - 1) / 2) < 0.01
    x = (t.keys >> np.uint64(32)).astype(np.int64)
    z = (t.keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert z.max() < N
    d = np.abs(z - x)
    d = np.minimum(d, N - d)
    assert d.max() <= HALF * T, d.max()
    # samples come from every slice of lanes: the vertices past 2^32 / 66
    # (their row pointers exceed 32 bits) own their share of the keys
    hi = int(np.sum(x >= (1 << 32) // (2 * HALF)))
    share = 1 - ((1 << 32) // (2 * HALF)) / N
    assert abs(hi / len(x) - share) < 0.01, (hi / len(x), share)


def test_propagation_above_2_32_edges(gpu, circulant):
    """M = (1 - mu) I - D~^-1 A~ maps the mode cos(2 pi j v / n) to
    lam_j = 1 - mu - (1 + 2 sum_k cos(2 pi j k / n)) / 67; the recurrence then
    gives y = f(lam_j) a and z = (1 + 2 sum_k cos(...)) y."""
    import torch

    g = circulant
    js = [0, 1, 2, 5, 1000, 12345, N // 3, N // 2 - 7, N // 2]
    d = 16
    v = torch.arange(N, device="cuda", dtype=torch.float64)
    a = torch.zeros(N, d, device="cuda", dtype=torch.float32)
    for c, j in enumerate(js):
        a[:, c] = torch.cos((2 * math.pi / N) * ((v * j) % N)).float()
    del v
    mu, theta, K = 0.2, 0.5, 10
    y, z, _ = L.propagate(a, g, order=K, mu=mu, theta=theta, ctx=gpu)
    coef = O.cheb_coeffs(K, theta).astype(np.float32).astype(np.float64)
    for c, j in enumerate(js):
        s = 1 + 2 * sum(math.cos(2 * math.pi * ((j * k) % N) / N) for k in range(1, HALF + 1))
        lam = 1 - mu - s / (2 * HALF + 1)
        l0, l1 = 1.0, 0.5 * lam * lam - 1.0
        conv = coef[0] * l0 + coef[1] * l1
        for i in range(2, K):
            l0, l1 = l1, (lam * lam - 2.0) * l1 - l0
            conv += coef[i] * l1
        f = 1.0 - conv
        ey = (y[:, c].double() - f * a[:, c].double()).abs().max().item()
        ez = (z[:, c].double() - s * f * a[:, c].double()).abs().max().item()
        print(f"mode j={j}: lam={lam:.6f} f={f:.6f} |y - f a| {ey:.2e} |z - s f a| {ez:.2e}", flush=True)
        # FP32 sums of 67 terms over 19 products: errors of order 1e-6; a
        # wrong row pointer or a lost launch of waves gives errors of order 1
        assert ey <= 5e-5
        assert ez <= 5e-5 * max(1.0, abs(s))
    assert torch.count_nonzero(y[:, len(js):]) == 0
