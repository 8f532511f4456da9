"""Differentiation form at full size: the Chung-Lu graph of 1.14M vertices,
d = 128, K = 10 (the adjoint workload the bench times).  No FP64 restatement
fits here, so the checks are size-independent properties: the adjoint identity
<g, J delta> = <J^T g, delta> (the propagation is linear in a), checkpoint
intervals that leave the gradients unchanged, and central differences of the
FP32 forward in mu and theta."""
import numpy as np
import pytest

import lightne_amd as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2_graph():
    hg = L.chung_lu(1_140_000, 2.0, 5.25, 2500.0, 2)
    return hg, hg.to_device()


def _dot(x, y):
    return float((x.double() * y.double()).sum().item())


@pytest.mark.parametrize("aggregate", [False, True])
def test_adjoint_identity_full_size(gpu, c2_graph, aggregate):
    import torch

    hg, g = c2_graph
    gen = torch.Generator(device="cuda").manual_seed(3)
    delta = torch.randn(hg.n, 128, device="cuda", generator=gen)
    w = torch.randn(hg.n, 128, device="cuda", generator=gen)
    y, z, _ = L.propagate(delta, g, order=10, ctx=gpu)
    fwd = z if aggregate else y
    ga, _, _ = L.propagate_backward(delta, w, g, order=10, aggregate=aggregate, checkpoint_every=3, ctx=gpu)
    lhs, rhs = _dot(w, fwd), _dot(ga, delta)
    # both sides are sums of 1.5e8 products of FP32 values with relative
    # errors of order 1e-6 and random signs, so their difference is of order
    # 1e-6 of the terms' root-sum-square; with random w and delta the sums
    # themselves are of the order of that scale, and so is the change a wrong
    # adjoint term makes
    rss = max(float(torch.linalg.vector_norm(w.double() * fwd.double())),
              float(torch.linalg.vector_norm(ga.double() * delta.double())))
    print(f"aggregate={aggregate}: <w, J d> = {lhs:.9g}, <J^T w, d> = {rhs:.9g}, rss {rss:.4g}")
    assert abs(lhs - rhs) <= 1e-4 * rss


def test_checkpoint_interval_full_size(gpu, c2_graph):
    import torch

    hg, g = c2_graph
    gen = torch.Generator(device="cuda").manual_seed(4)
    a = torch.randn(hg.n, 128, device="cuda", generator=gen)
    w = torch.randn(hg.n, 128, device="cuda", generator=gen)
    ref = L.propagate_backward(a, w, g, order=10, aggregate=True, checkpoint_every=0, ctx=gpu)
    for ck in (3, 4):
        got = L.propagate_backward(a, w, g, order=10, aggregate=True, checkpoint_every=ck, ctx=gpu)
        assert torch.equal(got[0], ref[0])
        assert got[1] == ref[1] and got[2] == ref[2]


def test_mu_theta_gradients_vs_central_differences_full_size(gpu, c2_graph):
    """L = <w, z(mu, theta)>; central differences of the FP32 forward with
    step 1e-3 (truncation ~1e-7 relative; FP32 rounding of z adds about
    1e-4 relative to the difference quotient at this size)."""
    import torch

    hg, g = c2_graph
    gen = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randn(hg.n, 128, device="cuda", generator=gen)
    w = torch.randn(hg.n, 128, device="cuda", generator=gen)
    mu0, th0, eps = 0.2, 0.5, 1e-3

    def Lf(mu, th):
        _, z, _ = L.propagate(a, g, order=10, mu=mu, theta=th, ctx=gpu)
        return _dot(w, z)

    _, gm, gt = L.propagate_backward(a, w, g, order=10, mu=mu0, theta=th0, aggregate=True, checkpoint_every=3,
                                     ctx=gpu)
    # the device takes mu and theta as float: differences of the float values
    fm_p, fm_m = float(np.float32(mu0 + eps)), float(np.float32(mu0 - eps))
    ft_p, ft_m = float(np.float32(th0 + eps)), float(np.float32(th0 - eps))
    fd_mu = (Lf(fm_p, th0) - Lf(fm_m, th0)) / (fm_p - fm_m)
    fd_th = (Lf(mu0, ft_p) - Lf(mu0, ft_m)) / (ft_p - ft_m)
    print(f"dL/dmu adjoint {gm:.9g} fd {fd_mu:.9g}; dL/dtheta adjoint {gt:.9g} fd {fd_th:.9g}")
    assert abs(gm - fd_mu) <= 5e-3 * abs(fd_mu)
    assert abs(gt - fd_th) <= 5e-3 * abs(fd_th)
