"""SIPDG diffusion apply on the GPU (gd_diff_apply_f64 / _f32) against GenDiL's
own DiffusionOperator (tests/golden/diff*, written by oracle/_ref/gendil_ref)
and, on meshes with many workgroups, against the oracle's restatement
(gdo_diffusion_apply, pinned to the same fixtures)."""
import numpy as np
import pytest
import torch

import gendil_amd as G
import oracle
from cases import DIFF_CASES, diffusivity_field, rel_l2
from gendil_amd import _lib as L

pytestmark = pytest.mark.gpu

F64_TOL, F32_TOL = 1e-12, 1e-5


@pytest.mark.parametrize("case", DIFF_CASES, ids=[c.name for c in DIFF_CASES])
def test_diffusion_matches_gendil(case):
    g = np.load(f"{oracle.HERE.parent}/tests/golden/{case.name}.npz")
    fes, ir = case.space()
    op = G.MakeDiffusionOperator(fes, ir, case.mu(), case.sigma, case.kappa)
    x = G.fill_random(int(g["n"]), case.seed)
    for dtype, tol in ((torch.float64, F64_TOL), (torch.float32, F32_TOL)):
        print(f"diffusion {case.name} {dtype}", flush=True)
        y = torch.full((op.num_dofs,), float("nan"), dtype=dtype, device="cuda")
        op(torch.as_tensor(x, dtype=dtype, device="cuda"), y)
        torch.cuda.synchronize()
        err = rel_l2(y.double().cpu().numpy(), g["y"])
        print(f"  rel l2 vs GenDiL = {err:.3e}")
        assert err <= tol


MESHES = [
    # (dim, sizes, periodic, p, q, diffusivity, sigma, kappa)
    (3, (12, 10, 7), (False, True, False), 3, 5, "trig", -1.0, 16.0),
    (3, (16, 16, 16), (True, True, True), 3, 4, "trig", -1.0, 16.0),
    (3, (9, 8, 11), (False, False, False), 2, 4, "poly", 1.0, 9.0),
    (2, (40, 33), (True, False), 3, 5, "trig", 0.0, 16.0),
    (1, (97,), (False,), 4, 6, "poly", -1.0, 25.0),
]


@pytest.mark.parametrize("m", MESHES, ids=[f"d{m[0]}_{'x'.join(map(str, m[1]))}_p{m[3]}q{m[4]}" for m in MESHES])
def test_diffusion_matches_the_oracle_on_larger_meshes(m):
    D, sizes, periodic, p, q, kind, sigma, kappa = m
    mesh = G.CartesianMesh(list(sizes), [1.0 / s for s in sizes], [0.0] * D, list(periodic))
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(p))
    ir = G.MakeIntegrationRule(q)
    mu = diffusivity_field(kind, D)
    op = G.MakeDiffusionOperator(fes, ir, mu, sigma, kappa)
    x = G.fill_random(op.num_dofs, 31)
    ref = oracle.Problem(fes, ir).diffusion(x, mu, sigma, kappa)
    for dtype, tol in ((torch.float64, F64_TOL), (torch.float32, F32_TOL)):
        print(f"diffusion {m} {dtype}", flush=True)
        y = torch.empty(op.num_dofs, dtype=dtype, device="cuda")
        op(torch.as_tensor(x, dtype=dtype, device="cuda"), y)
        torch.cuda.synchronize()
        err = rel_l2(y.double().cpu().numpy(), ref)
        print(f"  rel l2 vs oracle = {err:.3e}")
        assert err <= tol


def test_symmetric_form_is_symmetric_on_the_gpu():
    mesh = G.CartesianMesh([10, 9, 8], [0.1, 1 / 9, 0.125], [0.0] * 3, [True, False, False])
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    op = G.MakeDiffusionOperator(fes, G.MakeIntegrationRule(5), diffusivity_field("trig", 3), -1.0, 16.0)
    u = torch.as_tensor(G.fill_random(op.num_dofs, 1), device="cuda")
    v = torch.as_tensor(G.fill_random(op.num_dofs, 2), device="cuda")
    au, av = torch.empty_like(u), torch.empty_like(v)
    op(u, au)
    op(v, av)
    a, b = float(torch.dot(au, v)), float(torch.dot(u, av))
    assert abs(a - b) <= 1e-12 * float(au.norm() * v.norm())


def test_diffusion_rejects_what_it_cannot_apply():
    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, 4, True), G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    plan = G.Plan(fes, ir)
    x = torch.zeros(plan.num_dofs, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    # coefficients not set
    rc = L.lib().gd_diff_apply_f64(plan.handle, x.data_ptr(), y.data_ptr(), None)
    assert rc == L.GD_ERR_INVALID
    op = G.MakeDiffusionOperator(fes, ir, 1.0, -1.0, 16.0)
    # overlapping storage
    rc = L.lib().gd_diff_apply_f64(op.plan.handle, x.data_ptr(), x.data_ptr() + 8, None)
    assert rc == L.GD_ERR_INVALID
    # no instance for this (nodes, points)
    fes7 = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, 2, True), G.MakeLobattoFiniteElement(6))
    op7 = G.MakeDiffusionOperator(fes7, G.MakeIntegrationRule(7), 1.0, -1.0, 49.0)
    x7 = torch.zeros(op7.num_dofs, dtype=torch.float64, device="cuda")
    rc = L.lib().gd_diff_apply_f64(op7.plan.handle, x7.data_ptr(), torch.zeros_like(x7).data_ptr(), None)
    assert rc == L.GD_ERR_UNSUPPORTED
    assert b"no diffusion kernel" in L.lib().gd_last_error()
