"""torch.autograd wrapper of the advection apply with a per-point field.

y = adv_coef * A(beta) u is linear in u and (piecewise) linear in beta; the
backward pass calls the hand-written adjoint kernels through the C-ABI:
  dL/du    = adv_coef * A(beta)^T r            (gd_apply_T_f64)
  dL/dbeta = d<r, adv_coef A(beta) u>/dbeta     (gd_adv_grad_beta_f64)
with r = dL/dy.  beta is given at every cell quadrature point
([E, dim, q^dim]) and, as its normal component, at every face quadrature
point ([F_a, q^(dim-1)] per axis a); both sides of a face share that value.
Boundary faces follow the operator's boundary mode; with inflow data g the
map is affine in u and the adjoint is that of its linear part.
No GenDiL counterpart: GenDiL has no operator adjoint.
"""
from __future__ import annotations

import torch

from .operators import AdvectionOperator


class AdvectionApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, op: AdvectionOperator, u, beta_cell, *beta_faces):
        plan = op.plan
        plan.set_beta_pointwise(beta_cell, beta_faces)
        y = torch.empty_like(u)
        plan.apply(0.0, op.adv_coef, u, y)
        ctx.op = op
        ctx.save_for_backward(u, beta_cell, *beta_faces)
        return y

    @staticmethod
    def backward(ctx, gy):
        op = ctx.op
        u, beta_cell, *beta_faces = ctx.saved_tensors
        plan = op.plan
        plan.set_beta_pointwise(beta_cell, beta_faces)
        r = gy.contiguous()
        du = dcell = None
        dfaces = [None] * len(beta_faces)
        if ctx.needs_input_grad[1]:
            du = torch.empty_like(u)
            plan.apply_transpose(0.0, op.adv_coef, r, du)
        if any(ctx.needs_input_grad[2:]):
            dcell = torch.empty_like(beta_cell)
            dfaces = [torch.empty_like(f) for f in beta_faces]
            plan.grad_beta(op.adv_coef, u, r, dcell, dfaces)
        return (None, du, dcell, *dfaces)


def advection_apply(op: AdvectionOperator, u, beta_cell, beta_faces):
    """Differentiable y = A(beta) u for an operator built with a PointwiseField."""
    return AdvectionApply.apply(op, u, beta_cell, *beta_faces)
