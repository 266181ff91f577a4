"""p-adaptive apply (per-element orders 1..3, q = p + 1, one launch) against
the oracle in FP64 and FP32, plus the invariants that pin it independently:
conservation on a periodic mesh, the uniform-order special case against the
uniform kernels, and agreement of the autotuner's variants."""
import numpy as np
import pytest
import torch

import gendil_amd as G
import oracle
from cases import PADAPT_CASES, rel_l2
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
    # the quadrature layouts (0, 1) agree bitwise; the Kronecker kernel (2)
    # regroups the same exact integrals, so it agrees to rounding
    fes, adv = W.phase_space_6d_padaptive(3, seed=16)
    K, x, y0 = run(fes, adv, 0.0, 1.0, None, torch.float32)
    chosen, ms = K.plan.autotune(torch.float32, 0.0, 1.0, reps=3)
    print(f"autotune chose variant {chosen} ({ms:.3f} ms)")
    assert chosen in (0, 1, 2)
    xt = torch.as_tensor(x, device="cuda").float()
    y = torch.empty_like(xt)
    K(xt, y)
    assert rel_l2(y.double().cpu().numpy(), y0) <= F32_TOL


def paired_field(signs=(1, -1, 1, -1, 1, 0)):
    """beta_c = a_c X_{(c+3) mod 6} + b_c with the given signs over [0,1]^6
    (the structure the Kronecker 6D kernel serves); 0 = zero component."""
    A = np.zeros((6, 6))
    c0 = np.zeros(6)
    slopes = [0.5, 1.0, 0.0, 0.0, 0.25, 0.0]
    offsets = [0.2, 0.1, 0.3, 0.4, 0.1, 0.0]
    for c, sg in enumerate(signs):
        A[c, (c + 3) % 6] = sg * slopes[c]
        c0[c] = sg * offsets[c]
    return AffineField(A, c0)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, F64_TOL), (torch.float32, F32_TOL)])
def test_padapt_6d_kronecker_is_default_and_matches_oracle(dtype, tol):
    print(f"Kronecker 6D 3^6 phase space {dtype}", flush=True)
    fes, adv = W.phase_space_6d_padaptive(3, seed=13)
    K = G.MakeAdvectionOperator(fes, None, adv)
    assert K.plan.query_variant(dtype, 0.0, 1.0) == 2
    _, x, y = run(fes, adv, 0.0, 1.0, None, dtype)
    ref = oracle_apply(fes, adv, 0.0, 1.0, None, x)
    err = rel_l2(y, ref)
    print(f"rel l2 = {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("dtype,tol", [(torch.float64, F64_TOL), (torch.float32, F32_TOL)])
@pytest.mark.parametrize("periodic", [True, False])
def test_padapt_6d_kronecker_mixed_signs_mass_inflow(dtype, tol, periodic):
    mesh = G.CartesianMesh([3] * 6, [1 / 3] * 6, [0.0] * 6, periodic)
    orders = W.padaptive_orders(mesh.num_cells, 17)
    fes = G.MakePAdaptiveSpace(mesh, orders, orders + 1)
    adv = paired_field()
    bc = None if periodic else 0.75
    print(f"Kronecker 6D mixed signs periodic={periodic} {dtype}", flush=True)
    K = G.MakeMassAdvectionOperator(fes, None, adv, bc=bc, mass_coef=1.0, adv_coef=0.5)
    assert K.plan.query_variant(dtype, 1.0, 0.5) == 2
    _, x, y = run(fes, adv, 1.0, 0.5, bc, dtype, seed=18)
    ref = oracle_apply(fes, adv, 1.0, 0.5, bc, x)
    err = rel_l2(y, ref)
    print(f"rel l2 = {err:.3e}")
    assert err <= tol


def test_padapt_6d_kronecker_mass_only():
    fes, adv = W.phase_space_6d_padaptive(3, seed=19)
    _, x, y = run(fes, adv, 1.0, 0.0, None, torch.float64)
    ref = oracle_apply(fes, adv, 1.0, 0.0, None, x)
    assert rel_l2(y, ref) <= F64_TOL


def test_padapt_6d_unpaired_field_uses_quadrature_kernel():
    mesh = G.CartesianMesh([3] * 6, [1 / 3] * 6, [0.0] * 6, True)
    orders = W.padaptive_orders(mesh.num_cells, 15)
    fes = G.MakePAdaptiveSpace(mesh, orders, orders + 1)
    A = np.zeros((6, 6))
    A[0, 3] = A[1, 4] = A[2, 5] = 1.0
    A[3, 1] = -0.5  # beta_3 depends on X1, not on its partner X0
    K = G.MakeAdvectionOperator(fes, None, AffineField(A, np.zeros(6)))
    assert K.plan.query_variant(torch.float64, 0.0, 1.0) == 0
    # a sign change inside the mesh is not served either
    K2 = G.MakeAdvectionOperator(fes, None, paired_field(signs=(1, 1, 1, 1, 1, 1)))
    assert K2.plan.query_variant(torch.float64, 0.0, 1.0) == 2
    Ac = np.zeros((6, 6))
    Ac[0, 3] = 1.0
    K3 = G.MakeAdvectionOperator(fes, None, AffineField(Ac, np.array([-0.5, 0, 0, 0, 0, 0])))
    assert K3.plan.query_variant(torch.float64, 0.0, 1.0) == 0


def test_padapt_6d_kronecker_uniform_orders_match_uniform_kernel():
    mesh = G.CartesianMesh([3] * 6, [1 / 3] * 6, [0.0] * 6, True)
    E = mesh.num_cells
    adv = paired_field()
    pfes = G.MakePAdaptiveSpace(mesh, np.full(E, 2), np.full(E, 3))
    K, x, y = run(pfes, adv, 0.5, 1.0, None, torch.float64)
    assert K.plan.query_variant(torch.float64, 0.5, 1.0) == 2
    ufes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(2))
    U = G.MakeMassAdvectionOperator(ufes, G.MakeIntegrationRule(3), adv, bc=None, mass_coef=0.5, adv_coef=1.0)
    xt = torch.as_tensor(x, device="cuda")
    yu = torch.empty_like(xt)
    U(xt, yu)
    assert rel_l2(y, yu.cpu().numpy()) <= 1e-12


def padapt_variants(case):
    # 3D: the quadrature kernel; 6D: its two thread layouts and the Kronecker
    # kernel (the phase-space field is of the form it serves)
    return (0,) if case.D == 3 else (0, 1, 2)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, F64_TOL), (torch.float32, F32_TOL)])
@pytest.mark.parametrize("case", PADAPT_CASES, ids=lambda c: c.name)
def test_padapt_kernels_match_gendil_blocks(case, dtype, tol):
    """The one-launch p-adaptive kernels against GenDiL's own p-adaptivity
    (uniform-order blocks + intermesh face operator, tests/golden/padapt*),
    through the C-ABI, every kernel variant that serves the case."""
    g = np.load(f"{oracle.HERE.parent}/tests/golden/{case.name}.npz")
    xg = G.fill_random(int(g["n"]), case.seed)
    fes, perm = case.layout()
    x = np.empty_like(xg)
    x[perm] = xg
    xt = torch.as_tensor(x, device="cuda").to(dtype)
    for v in padapt_variants(case):
        print(f"{case.name} variant {v} {dtype}", flush=True)
        K = G.MakeAdvectionOperator(fes, None, case.field())
        K.plan.set_variant(dtype, v)
        assert K.plan.query_variant(dtype, 0.0, 1.0) == v
        y = torch.empty_like(xt)
        K(xt, y)
        torch.cuda.synchronize()
        err = rel_l2(y.double().cpu().numpy()[perm], g["y"])
        print(f"rel l2 = {err:.3e}")
        assert err <= tol


def test_padapt_coefficient_change_is_ordered_with_applies_on_a_side_stream():
    # apply, change the field, apply again, all queued on a non-default
    # stream: the first apply must see the old field and the second the new
    # one (the setter drains the device before it replaces the parameters)
    fes = space3d(4, True, 21)
    adv1 = AffineField(A3, C3)
    adv2 = AffineField(-0.5 * A3, np.array([-0.1, 0.4, 0.2]))
    x64 = G.fill_random(fes.num_dofs, 22)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        K = G.MakeAdvectionOperator(fes, None, adv1)
        x = torch.as_tensor(x64, device="cuda")
        y1 = torch.empty_like(x)
        y2 = torch.empty_like(x)
        for _ in range(20):  # keep the stream busy when the setter runs
            K(x, y1)
        K.plan.set_beta(adv2)
        K(x, y2)
    side.synchronize()
    r1 = oracle_apply(fes, adv1, 0.0, 1.0, None, x64)
    r2 = oracle_apply(fes, adv2, 0.0, 1.0, None, x64)
    assert rel_l2(y1.cpu().numpy(), r1) <= F64_TOL
    assert rel_l2(y2.cpu().numpy(), r2) <= F64_TOL
