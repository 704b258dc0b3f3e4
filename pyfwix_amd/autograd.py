"""torch.autograd.Function wrappers over the hand-written adjoint kernels.

* ``BornFunction``: d = B δs, the linearized Born modeling of the
  down-scattering channel (``propagator/src/ExtendedBorn.cpp:152-178``); its
  backward is the migration adjoint B^H (``ExtendedBorn.cpp:180-208``), run
  by the HIP adjoint kernels, with background wavefields taken from the
  operator's HBM checkpoints (every depth, or every k-th with recompute).
* ``OneWayFunction``: u_out = P_{nz-1} ... P_0 u (``OneWay.cpp:6-79``); its
  backward is the back-propagation P_0^H ... P_{nz-1}^H g.

torch's complex convention: for y = A x the vector-Jacobian product is
A^H g, and for a real loss L the gradient is dL/d(conj x) times 2, i.e. the
steepest-ascent direction, so ``x.grad`` of 0.5*||B x - d||^2 is
B^H (B x - d).
"""
from __future__ import annotations

import torch

from .ops import Born, OneWay

__all__ = ["BornFunction", "OneWayFunction", "born", "oneway", "born_misfit_gradient"]


class BornFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dslow: torch.Tensor, op: Born) -> torch.Tensor:
        data = torch.empty(op.trace_shape(True), dtype=torch.complex64, device=dslow.device)
        op.forward(False, dslow.contiguous(), data)
        ctx.op = op
        return data

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        op = ctx.op
        grad = torch.empty(op.model_shape(), dtype=torch.complex64, device=g.device)
        op.adjoint(False, grad, g.contiguous())
        return grad, None


class OneWayFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u: torch.Tensor, op: OneWay) -> torch.Tensor:
        out = torch.empty_like(u)
        op.forward(False, u.contiguous(), out)
        ctx.op = op
        return out

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        grad = torch.empty_like(g)
        ctx.op.adjoint(False, grad, g.contiguous())
        return grad, None


def born(op: Born, dslow: torch.Tensor) -> torch.Tensor:
    """Differentiable B δs.  The operator's background must be modeled (``op.model()``)."""
    return BornFunction.apply(dslow, op)


def oneway(op: OneWay, u: torch.Tensor) -> torch.Tensor:
    """Differentiable depth recursion of a Downward / Upward operator."""
    return OneWayFunction.apply(u, op)


def born_misfit_gradient(op: Born, dslow: torch.Tensor, observed: torch.Tensor):
    """0.5*||B δs - d||^2 and its gradient B^H (B δs - d), through autograd."""
    x = dslow.detach().requires_grad_(True)
    r = born(op, x) - observed
    loss = 0.5 * torch.sum(r.real.double() ** 2 + r.imag.double() ** 2)
    loss.backward()
    return loss.detach(), x.grad
