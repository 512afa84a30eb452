"""Reverse-mode assembly exposed to torch.autograd.

J(c, x) = u^T K(c, x) u - F(x)^T u for a fixed state u, where K is the assembled bilinear
form and F the assembled linear form. The forward pass runs the adjoint kernel once: it
returns J and, in the same sweep over the cells, dJ/dc (nodal coefficient gradients) and
dJ/dx (vertex coordinate gradients), recomputing element geometry instead of storing it.
backward() only scales the stored gradients by the incoming grad_output.

DOLFIN's counterpart is assembling UFL's derivative(J, kappa) and derivative(J, X) forms
(python/test/unit/fem/test_coordinate_derivative.py:29-63).
"""
import torch

from . import _lib
from .fem import engine_for


def _run(V, form_kind, x, c0, c1, f, f3, u):
    eng = engine_for(V)
    J = torch.zeros(1, dtype=torch.float64, device=x.device)
    dc0 = torch.empty_like(c0)
    dc1 = torch.empty_like(c1) if c1 is not None else None
    dx = torch.empty_like(x)
    _lib.check(_lib.load().dfa_adjoint(eng.handle, form_kind, _lib.ptr(x), _lib.ptr(c0),
                                       _lib.ptr(c1), _lib.ptr(f), _lib.ptr(f3), _lib.ptr(u),
                                       _lib.ptr(J), _lib.ptr(dc0), _lib.ptr(dc1), _lib.ptr(dx),
                                       _lib.stream()))
    return J, dc0, dc1, dx


class _PoissonEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa, x, V, f, u):
        J, dk, _, dx = _run(V, _lib.FORM_P1_POISSON, x.contiguous(), kappa.contiguous(), None,
                            f.contiguous(), None, u.contiguous())
        ctx.save_for_backward(dk, dx)
        return J[0]

    @staticmethod
    def backward(ctx, g):
        dk, dx = ctx.saved_tensors
        return g * dk, g * dx, None, None, None


class _ElasticEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mu, lmbda, x, V, f3, u):
        import numpy as np
        f3a = np.ascontiguousarray(f3, dtype=np.float64)
        J, dmu, dla, dx = _run(V, _lib.FORM_VP2_ELASTICITY, x.contiguous(), mu.contiguous(),
                               lmbda.contiguous(), None, f3a, u.contiguous())
        ctx.save_for_backward(dmu, dla, dx)
        return J[0]

    @staticmethod
    def backward(ctx, g):
        dmu, dla, dx = ctx.saved_tensors
        return g * dmu, g * dla, g * dx, None, None, None


def poisson_energy(V, kappa, x, f, u):
    """J = u^T K(kappa, x) u - F(f, x)^T u for P1 kappa-Poisson; differentiable in kappa and
    x (torch float64 tensors on the device; x has shape [num_vertices, gdim])."""
    return _PoissonEnergy.apply(kappa, x, V, f, u)


def elastic_energy(V, mu, lmbda, x, f3, u):
    """J = u^T K(mu, lmbda, x) u - F(x)^T u for vector P2 elasticity with body force f3;
    differentiable in mu, lmbda (P1 nodal) and x."""
    return _ElasticEnergy.apply(mu, lmbda, x, V, f3, u)
