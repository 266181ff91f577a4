"""p-adaptive apply (per-element orders 1..3, q = p + 1, one launch) against
the oracle in FP64 and FP32, plus the invariants that pin it independently:
conservation on a periodic mesh, the uniform-order special case against the
uniform kernels, and agreement of the autotuner's variants."""
import numpy as np
import pytest
import torch

import gendil_amd as G
import oracle
from cases import rel_l2
from gendil_amd import workloads as W
from gendil_amd.fields import AffineField

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 1e-5

A3 = np.array([[0, 1.0, 0], [0, 0, 1.0], [1.0, 0, 0]])
C3 = np.array([0.3, -0.2, 0.1])


def space3d(n, periodic, seed):
    mesh = G.CartesianMesh.unit(3, n, periodic)
    orders = W.padaptive_orders(mesh.num_cells, seed)
    return G.MakePAdaptiveSpace(mesh, orders, orders + 1)


def run(fes, adv, mass_coef, adv_coef, bc, dtype, seed=5, variant=None):
    K = G.MakeMassAdvectionOperator(fes, None, adv, bc=bc, mass_coef=mass_coef, adv_coef=adv_coef)
    if variant is not None:
        K.plan.set_variant(dtype, variant)
    x64 = G.fill_random(K.num_dofs, seed)
    x = torch.as_tensor(x64, device="cuda").to(dtype)
    y = torch.empty_like(x)
    K(x, y)
    torch.cuda.synchronize()
    return K, x64, y.double().cpu().numpy()


def oracle_apply(fes, adv, mass_coef, adv_coef, bc, x):
    P = oracle.Problem(fes, None, mass_coef=mass_coef, adv_coef=adv_coef, adv=adv, bc=bc)
    return P.apply(x)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, F64_TOL), (torch.float32, F32_TOL)])
def test_padapt_3d_periodic_vs_oracle(dtype, tol):
    print(f"p-adaptive 3D 4^3 periodic {dtype}", flush=True)
    fes = space3d(4, True, 11)
    adv = AffineField(A3, C3)
    _, x, y = run(fes, adv, 0.0, 1.0, None, dtype)
    ref = oracle_apply(fes, adv, 0.0, 1.0, None, x)
    err = rel_l2(y, ref)
    print(f"rel l2 = {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("dtype,tol", [(torch.float64, F64_TOL), (torch.float32, F32_TOL)])
def test_padapt_3d_inflow_mass_advection_vs_oracle(dtype, tol):
    print(f"p-adaptive 3D 5^3 inflow bc, M + 0.5 A, {dtype}", flush=True)
    fes = space3d(5, False, 12)
    adv = AffineField(A3, C3)
    _, x, y = run(fes, adv, 1.0, 0.5, 0.75, dtype)
    ref = oracle_apply(fes, adv, 1.0, 0.5, 0.75, x)
    err = rel_l2(y, ref)
    print(f"rel l2 = {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("dtype,tol", [(torch.float64, F64_TOL), (torch.float32, F32_TOL)])
def test_padapt_6d_phase_space_vs_oracle(dtype, tol):
    print(f"p-adaptive 6D 3^6 phase space {dtype}", flush=True)
    fes, adv = W.phase_space_6d_padaptive(3, seed=13)
    _, x, y = run(fes, adv, 0.0, 1.0, None, dtype)
    ref = oracle_apply(fes, adv, 0.0, 1.0, None, x)
    err = rel_l2(y, ref)
    print(f"rel l2 = {err:.3e}")
    assert err <= tol


def test_padapt_6d_variants_agree_bitwise():
    fes, adv = W.phase_space_6d_padaptive(3, seed=14)
    outs = []
    for v in (0, 1):
        _, _, y = run(fes, adv, 0.0, 1.0, None, torch.float32, variant=v)
        outs.append(y)
    assert np.array_equal(outs[0], outs[1])


def test_padapt_conservation_periodic():
    # every face flux enters two elements with opposite signs (shared face
    # rule q = max of the two cell rules): sum_i (A u)_i = 0
    mesh = G.CartesianMesh([3] * 6, [1 / 3] * 6, [0.0] * 6, True)
    orders = W.padaptive_orders(mesh.num_cells, 15)
    fes = G.MakePAdaptiveSpace(mesh, orders, orders + 1)
    A = np.zeros((6, 6))
    A[0, 3] = A[1, 4] = A[2, 5] = 1.0
    A[3, 1] = -0.5
    adv = AffineField(A, np.array([0.1, 0.2, -0.3, 0.4, -0.5, 0.25]))
    _, x, y = run(fes, adv, 0.0, 1.0, None, torch.float64)
    s = abs(y.sum()) / np.abs(y).sum()
    print(f"|sum A u| / sum |A u| = {s:.3e}")
    assert s <= 1e-13


def test_padapt_uniform_orders_match_uniform_kernels():
    mesh = G.CartesianMesh.unit(3, 4, True)
    E = mesh.num_cells
    adv = AffineField(A3, C3)
    pfes = G.MakePAdaptiveSpace(mesh, np.full(E, 3), np.full(E, 4))
    _, x, y = run(pfes, adv, 1.0, 1.0, None, torch.float64)
    ufes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    K = G.MakeMassAdvectionOperator(ufes, G.MakeIntegrationRule(4), adv, bc=None, mass_coef=1.0, adv_coef=1.0)
    xt = torch.as_tensor(x, device="cuda")
    yu = torch.empty_like(xt)
    K(xt, yu)
    assert rel_l2(y, yu.cpu().numpy()) <= 1e-13


def test_padapt_autotune_keeps_results():
    fes, adv = W.phase_space_6d_padaptive(3, seed=16)
    K, x, y0 = run(fes, adv, 0.0, 1.0, None, torch.float32)
    chosen, ms = K.plan.autotune(torch.float32, 0.0, 1.0, reps=3)
    print(f"autotune chose variant {chosen} ({ms:.3f} ms)")
    xt = torch.as_tensor(x, device="cuda").float()
    y = torch.empty_like(xt)
    K(xt, y)
    assert np.array_equal(y.double().cpu().numpy(), y0)
