"""Adjoint of the upwind advection apply through torch.autograd: the
hand-written A^T and d/dbeta kernels against the oracle, the dot-product
identity, and central finite differences (eps = 1e-6) of
L = 0.5 ||A(beta) u - y*||^2 on random directions, at a small size and at the
48^3, p=4, q=5 adjoint workload."""
import numpy as np
import pytest
import torch

import gendil_amd as G
import oracle
from cases import rel_l2
from gendil_amd import workloads as W
from gendil_amd.autograd import advection_apply

pytestmark = pytest.mark.gpu


def setup(n, p, q, seed=7):
    mesh = G.CartesianMesh.unit(3, n, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(p))
    ir = G.MakeIntegrationRule(q)
    cell, faces = W.noisy_pointwise_beta(mesh, q, seed)
    bc = torch.as_tensor(cell, device="cuda")
    bf = [torch.as_tensor(f, device="cuda") for f in faces]
    op = G.MakeAdvectionOperator(fes, ir, G.PointwiseField(bc, bf))
    return mesh, fes, ir, op, cell, faces, bc, bf


def loss_fn(op, u, bc, bf, ystar):
    y = advection_apply(op, u, bc, bf)
    return 0.5 * torch.sum((y - ystar) ** 2)


def test_autograd_matches_oracle_small():
    mesh, fes, ir, op, cell, faces, bc, bf = setup(4, 4, 5)
    N = op.num_dofs
    u_np = G.fill_random(N, 11)
    ys_np = G.fill_random(N, 12)
    u = torch.as_tensor(u_np, device="cuda").requires_grad_(True)
    bc.requires_grad_(True)
    for f in bf:
        f.requires_grad_(True)
    L = loss_fn(op, u, bc, bf, torch.as_tensor(ys_np, device="cuda"))
    L.backward()
    P = oracle.Problem(fes, ir, adv_coef=1.0, adv=(cell, faces))
    r = P.apply(u_np) - ys_np
    assert rel_l2(u.grad.cpu().numpy(), P.apply_transpose(r)) <= 1e-12
    oc, of = P.grad_beta(u_np, r, G.face_counts(mesh))
    assert rel_l2(bc.grad.cpu().numpy().reshape(-1), oc) <= 1e-12
    for g, o in zip(bf, of):
        assert rel_l2(g.grad.cpu().numpy().reshape(-1), o) <= 1e-12


def fd_check(n, p, q, ndir, eps=1e-6, tol=1e-8):
    mesh, fes, ir, op, cell, faces, bc, bf = setup(n, p, q)
    N = op.num_dofs
    print(f"FD check {n}^3 p={p} q={q}: {N} dofs, {ndir} directions", flush=True)
    u = torch.as_tensor(G.fill_random(N, 21), device="cuda").requires_grad_(True)
    ystar = torch.as_tensor(G.fill_random(N, 22), device="cuda")
    bc.requires_grad_(True)
    for f in bf:
        f.requires_grad_(True)
    L = loss_fn(op, u, bc, bf, ystar)
    L.backward()
    grads = [u.grad, bc.grad, *[f.grad for f in bf]]
    params = [u.detach(), bc.detach(), *[f.detach() for f in bf]]
    gen = torch.Generator(device="cuda").manual_seed(5)
    gnorm = float(torch.sqrt(sum(torch.sum(g * g) for g in grads)))
    worst = 0.0
    worst_rel = 0.0
    with torch.no_grad():
        for k in range(ndir):
            dirs = [torch.randn(t.shape, dtype=torch.float64, device="cuda", generator=gen) for t in params]
            nrm = torch.sqrt(sum(torch.sum(d * d) for d in dirs))
            dirs = [d / nrm for d in dirs]
            plus = [t + eps * d for t, d in zip(params, dirs)]
            minus = [t - eps * d for t, d in zip(params, dirs)]
            # central difference of L in a cancellation-free form:
            # L(+) - L(-) = 0.5 sum (y+ - y-) (y+ + y- - 2 y*)
            yp = advection_apply(op, plus[0], plus[1], plus[2:])
            ym = advection_apply(op, minus[0], minus[1], minus[2:])
            fd = float(0.5 * torch.sum((yp - ym) * (yp + ym - 2.0 * ystar)) / (2 * eps))
            an = float(sum(torch.sum(g * d) for g, d in zip(grads, dirs)))
            # gradient-check criterion: error relative to ||grad L|| (unit directions)
            err = abs(fd - an) / gnorm
            worst = max(worst, err)
            worst_rel = max(worst_rel, abs(fd - an) / max(abs(an), 1e-300))
    print(f"{ndir} directions: max |fd - <grad, d>| / ||grad|| = {worst:.3e}, "
          f"max relative to |<grad, d>| = {worst_rel:.3e}")
    assert worst <= tol
    # dot-product identity on the same operator: <A^T r, v> = <r, A v>
    with torch.no_grad():
        op.plan.set_beta_pointwise(bc.detach(), [f.detach() for f in bf])
        r = torch.as_tensor(G.fill_random(N, 31), device="cuda")
        v = torch.as_tensor(G.fill_random(N, 32), device="cuda")
        Atr = torch.empty_like(r)
        op.transpose(r, Atr)
        Av = torch.empty_like(v)
        op(v, Av)
        lhs, rhs = float(torch.dot(Atr, v)), float(torch.dot(r, Av))
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_finite_differences_small():
    fd_check(5, 2, 3, 16)


def test_finite_differences_adjoint_workload_48_cubed():
    # 48^3 elements, p=4, q=5 (13.8M DoFs), 64 random directions, eps = 1e-6
    fd_check(48, 4, 5, 64)
