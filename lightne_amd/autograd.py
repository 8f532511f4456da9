"""torch.autograd.Function over the C-ABI: Chebyshev propagation with its
hand-written adjoint (lightne_propagate / lightne_propagate_backward).

    y = ChebyshevPropagation.apply(a, mu, theta, graph, order, aggregate, ckpt)

``a`` is an n x d float32 CUDA tensor; ``mu`` and ``theta`` are 0-d tensors
(requires_grad allowed).  The forward computes y = a - sum_i c_i(theta) T_i a
(z = (A + I) y when ``aggregate``) exactly as lightne_propagate does; the
backward returns dL/da, dL/dmu, dL/dtheta from the reverse recurrence with
T_i recomputed from checkpoints every ``ckpt`` steps.
"""
from __future__ import annotations

import torch

from . import api


class ChebyshevPropagation(torch.autograd.Function):
    @staticmethod
    def forward(fctx, a, mu, theta, graph, order=10, aggregate=False, checkpoint_every=3):
        mu_f, th_f = float(mu), float(theta)
        y, z, _ = api.propagate(a, graph, order=order, mu=mu_f, theta=th_f, aggregate=aggregate)
        fctx.save_for_backward(a)
        fctx.graph, fctx.order, fctx.aggregate, fctx.ckpt = graph, order, aggregate, checkpoint_every
        fctx.mu, fctx.theta = mu_f, th_f
        fctx.param_like = (mu, theta)
        return z if aggregate else y

    @staticmethod
    def backward(fctx, grad_out):
        (a,) = fctx.saved_tensors
        ga, gmu, gth = api.propagate_backward(a, grad_out, fctx.graph, order=fctx.order, mu=fctx.mu,
                                              theta=fctx.theta, aggregate=fctx.aggregate,
                                              checkpoint_every=fctx.ckpt)
        mu, theta = fctx.param_like
        gmu_t = torch.tensor(gmu, dtype=mu.dtype, device=mu.device) if torch.is_tensor(mu) else None
        gth_t = torch.tensor(gth, dtype=theta.dtype, device=theta.device) if torch.is_tensor(theta) else None
        return ga, gmu_t, gth_t, None, None, None, None


def chebyshev_propagation(a, mu, theta, graph, order=10, aggregate=False, checkpoint_every=3):
    return ChebyshevPropagation.apply(a, mu, theta, graph, order, aggregate, checkpoint_every)
