"""Differentiation form: the hand-written adjoint of the Chebyshev propagation
vs torch float64 autograd through the oracle, and vs central finite
differences in float64 (config 4's check: 2,000-node Chung-Lu graph, d = 16)."""
import numpy as np
import pytest

import lightne_amd as L
from lightne_amd.autograd import chebyshev_propagation
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(seed=0, n=2000, d=16):
    import torch

    hg = L.chung_lu(n, 2.1, 10.0, 150.0, 11)
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((hg.n, d)).astype(np.float32)
    w = rng.standard_normal((hg.n, d)).astype(np.float32)  # L = <w, y>
    return hg, a, w, torch


def _fp64_grads(hg, a, w, mu, theta, order, aggregate):
    import torch

    at = torch.tensor(a, dtype=torch.float64, requires_grad=True)
    mt = torch.tensor(float(np.float32(mu)), dtype=torch.float64, requires_grad=True)
    tt = torch.tensor(float(np.float32(theta)), dtype=torch.float64, requires_grad=True)
    y = O.propagate_torch(hg.offsets, hg.edges, at, mt, tt, order, aggregate)
    (y * torch.tensor(w, dtype=torch.float64)).sum().backward()
    return at.grad.numpy(), mt.grad.item(), tt.grad.item(), y.detach().numpy()


@pytest.mark.parametrize("aggregate", [False, True])
@pytest.mark.parametrize("order", [2, 3, 10])
def test_adjoint_matches_fp64_autograd(gpu, order, aggregate):
    hg, a, w, torch = _setup(order)
    g = hg.to_device()
    ga, gm, gt = L.propagate_backward(torch.from_numpy(a).cuda(), torch.from_numpy(w).cuda(), g, order=order,
                                      aggregate=aggregate, checkpoint_every=3, ctx=gpu)
    ra, rm, rt, y64 = _fp64_grads(hg, a, w, 0.2, 0.5, order, aggregate)
    # the forward of the FP64 restatement agrees with the oracle's numpy one
    y_np, z_np = O.propagate(hg.offsets, hg.edges, a, order=order)
    np.testing.assert_allclose(y64, z_np if aggregate else y_np, rtol=1e-10, atol=1e-10)
    ea = np.abs(ga.cpu().numpy() - ra).max() / np.abs(ra).max()
    print(f"order={order} agg={aggregate}: rel err grad_a={ea:.2e} mu {gm:.6g} vs {rm:.6g} theta {gt:.6g} vs {rt:.6g}")
    assert ea < 2e-5
    assert abs(gm - rm) <= 2e-5 * max(1.0, abs(rm))
    assert abs(gt - rt) <= 2e-5 * max(1.0, abs(rt))


def test_fp64_autograd_matches_finite_differences(gpu):
    """Pins the FP64 reference gradients themselves (central differences of the
    oracle's numpy forward in float64) and the device adjoint against them."""
    hg, a, w, torch = _setup(5)
    order = 10
    ra, rm, rt, _ = _fp64_grads(hg, a, w, 0.2, 0.5, order, False)
    a64 = a.astype(np.float64)
    mu0, th0 = float(np.float32(0.2)), float(np.float32(0.5))

    def Lf(a_, mu_, th_):
        M = O.prop_matrix(hg.offsets, hg.edges, mu_)
        c = O.cheb_coeffs(order, th_)
        Lx0, Lx1 = a_, 0.5 * (M @ (M @ a_)) - a_
        conv = c[0] * Lx0 + c[1] * Lx1
        for i in range(2, order):
            Lx2 = M @ (M @ Lx1) - 2.0 * Lx1 - Lx0
            conv = conv + c[i] * Lx2
            Lx0, Lx1 = Lx1, Lx2
        return float(np.sum(w * (a_ - conv)))

    eps = 1e-5
    dirn = np.random.default_rng(1).standard_normal(a.shape)
    fd_a = (Lf(a64 + eps * dirn, mu0, th0) - Lf(a64 - eps * dirn, mu0, th0)) / (2 * eps)
    fd_mu = (Lf(a64, mu0 + eps, th0) - Lf(a64, mu0 - eps, th0)) / (2 * eps)
    fd_th = (Lf(a64, mu0, th0 + eps) - Lf(a64, mu0, th0 - eps)) / (2 * eps)
    assert abs(np.sum(ra * dirn) - fd_a) <= 1e-7 * max(1.0, abs(fd_a))
    assert abs(rm - fd_mu) <= 1e-7 * max(1.0, abs(fd_mu))
    assert abs(rt - fd_th) <= 1e-7 * max(1.0, abs(fd_th))
    ga, gm, gt = L.propagate_backward(torch.from_numpy(a).cuda(), torch.from_numpy(w).cuda(), hg.to_device(),
                                      order=order, checkpoint_every=3, ctx=gpu)
    assert abs(float(np.sum(ga.cpu().numpy().astype(np.float64) * dirn)) - fd_a) <= 2e-5 * abs(fd_a)
    assert abs(gm - fd_mu) <= 2e-5 * max(1.0, abs(fd_mu))
    assert abs(gt - fd_th) <= 2e-5 * max(1.0, abs(fd_th))


@pytest.mark.parametrize("ckpt", [0, 2, 3, 4, 9])
def test_checkpointing_does_not_change_gradients(gpu, ckpt):
    hg, a, w, torch = _setup(2)
    g = hg.to_device()
    at, wt = torch.from_numpy(a).cuda(), torch.from_numpy(w).cuda()
    ref = L.propagate_backward(at, wt, g, order=10, checkpoint_every=0, ctx=gpu)
    got = L.propagate_backward(at, wt, g, order=10, checkpoint_every=ckpt, ctx=gpu)
    np.testing.assert_array_equal(got[0].cpu().numpy(), ref[0].cpu().numpy())
    assert got[1] == ref[1] and got[2] == ref[2]


def test_autograd_function(gpu):
    hg, a, w, torch = _setup(3)
    g = hg.to_device()
    at = torch.from_numpy(a).cuda().requires_grad_(True)
    mu = torch.tensor(0.2, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    z = chebyshev_propagation(at, mu, th, g, 10, True, 3)
    y_ref, z_ref, _ = L.propagate(at.detach(), g, ctx=gpu)
    np.testing.assert_array_equal(z.detach().cpu().numpy(), z_ref.cpu().numpy())
    (z * torch.from_numpy(w).cuda()).sum().backward()
    ra, rm, rt, _ = _fp64_grads(hg, a, w, 0.2, 0.5, 10, True)
    assert np.abs(at.grad.cpu().numpy() - ra).max() / np.abs(ra).max() < 2e-5
    assert abs(mu.grad.item() - rm) <= 2e-5 * max(1.0, abs(rm))
    assert abs(th.grad.item() - rt) <= 2e-5 * max(1.0, abs(rt))
