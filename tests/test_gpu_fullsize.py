"""Parity at the benchmark configurations' full sizes.

  - 64^3 periodic upwind advection, GLL p=3, 4 Gauss points, the benchmark
    field (16.8M DoFs), and the 32^3 mass apply, GLL p=3, 4 points, sigma = 1
    (2.1M DoFs): every kernel variant that serves the plan against GenDiL's
    own CPU apply (oracle/_ref/gendil_ref, GenDiL's SerialKernelConfiguration
    operators compiled from its headers) on x = FillRandom(12345), at <= 1e-12
    relative L2 in FP64; the FP32 error against the same output is stated.
  - the 32^3 implicit advection step (M + dt A) x = M u0 solved by the FP64
    and the FP32-inner / FP64-refined GMRES(30) to 1e-10: the true residual
    recomputed by the oracle's operator.
  - the 8^6 phase-space apply with per-element orders 1..3 (seed 2024, 235M
    DoFs): the GPU apply on a sample of elements of every order against the
    oracle (pinned by GenDiL's p-adaptivity fixtures), and conservation over
    the whole mesh.

GenDiL's binary is built in the build container and travels with the
repository; if it is missing these tests fail rather than skip.
"""
import numpy as np
import pytest
import torch

import gendil_amd as G
import oracle
from cases import rel_l2
from gendil_amd import solvers as S
from gendil_amd import workloads as W

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 1e-5


def gendil_apply(tmp_path, op, D, P, Q, basis, n, periodic, coef, bc, seed):
    if not oracle.ref_available():
        pytest.fail("oracle/_ref/gendil_ref (GenDiL's own CPU apply) is missing: build it with `make -C oracle`")
    return oracle.ref_fixture(op, D, P, Q, basis, n, periodic, coef, bc, seed, tmp_path / "gendil.bin")


def variants_served(op, dtype, mass_coef, adv_coef, candidates):
    """Kernel variants of `candidates` that the plan really runs when pinned."""
    out = []
    for v in candidates:
        op.plan.set_variant(dtype, v)
        if op.plan.query_variant(dtype, mass_coef, adv_coef) == v:
            out.append(v)
    op.plan.set_variant(dtype, -1)
    return out


# generic kernel, register kernel (1, 2, 4 lanes), register Kronecker kernel
# (17, 18, 20), the LDS-staged Kronecker kernel (48..55) and its streaming
# forms (56, 57)
CANDIDATES = [0, 1, 2, 4, 17, 18, 20, 48, 49, 50, 51, 52, 53, 56, 57]


def check_variants(op, x, y_ref, mass_coef, adv_coef, label):
    xt = torch.as_tensor(x, device="cuda")
    yt = torch.empty_like(xt)
    default = op.plan.query_variant(torch.float64, mass_coef, adv_coef)
    served = variants_served(op, torch.float64, mass_coef, adv_coef, CANDIDATES)
    assert default in served, (default, served)
    for v in served:
        print(f"{label}: FP64 variant {v}", flush=True)
        op.plan.set_variant(torch.float64, v)
        yt.fill_(np.nan)
        op(xt, yt)
        torch.cuda.synchronize()
        err = rel_l2(yt.cpu().numpy(), y_ref)
        print(f"  rel l2 vs GenDiL = {err:.3e}")
        assert err <= F64_TOL
    op.plan.set_variant(torch.float64, -1)
    del yt
    x32 = xt.float()
    y32 = torch.empty_like(x32)
    op(x32, y32)
    torch.cuda.synchronize()
    err32 = rel_l2(y32.double().cpu().numpy(), y_ref)
    print(f"{label}: FP32 default variant {op.plan.query_variant(torch.float32, mass_coef, adv_coef)}, "
          f"rel l2 vs GenDiL's FP64 = {err32:.3e}")
    assert err32 <= F32_TOL


def test_advection_64cubed_p3_matches_gendil(tmp_path):
    print("64^3 periodic p=3 q=4 advection, benchmark field, vs GenDiL", flush=True)
    x, y = gendil_apply(tmp_path, "adv", 3, 3, 4, "L", 64, True, "c2", "none", 12345)
    assert np.array_equal(x, G.fill_random(x.size, 12345))
    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, 64, True), G.MakeLobattoFiniteElement(3))
    op = G.MakeAdvectionOperator(fes, G.MakeIntegrationRule(4), W.advection_c2())
    assert op.num_dofs == x.size == 16_777_216
    check_variants(op, x, y, 0.0, 1.0, "advection 64^3")


def test_mass_32cubed_p3_matches_gendil(tmp_path):
    print("32^3 p=3 q=4 mass, sigma = 1, vs GenDiL", flush=True)
    x, y = gendil_apply(tmp_path, "mass", 3, 3, 4, "L", 32, False, "one", "none", 12345)
    assert np.array_equal(x, G.fill_random(x.size, 12345))
    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, 32, False), G.MakeLobattoFiniteElement(3))
    op = G.MakeMassFiniteElementOperator(fes, G.MakeIntegrationRule(4))
    assert op.num_dofs == x.size == 2_097_152
    check_variants(op, x, y, 1.0, 0.0, "mass 32^3")


@pytest.mark.parametrize("solver", ["fp64", "mixed"])
def test_implicit_step_32cubed_solve(solver):
    print(f"32^3 p=3 implicit step, GMRES(30) {solver} to 1e-10", flush=True)
    n = 32
    h = 1.0 / n
    mesh = G.CartesianMesh.unit(3, n, False)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    dt = 0.5 * h / W.advection_c2_max_norm()
    K = G.MakeMassAdvectionOperator(fes, ir, W.advection_c2(), bc=0.0, mass_coef=1.0, adv_coef=dt)
    M = G.MakeMassFiniteElementOperator(fes, ir)
    u0 = W.gaussian_bump(fes, (0.5, 0.5, 0.5), 0.1) + 1e-3 * G.fill_random(K.num_dofs, 2024)
    u0t = torch.as_tensor(u0, device="cuda")
    b = torch.empty_like(u0t)
    M(u0t, b)
    x = torch.zeros_like(b)
    fn = S.gmres if solver == "fp64" else S.gmres_mixed
    res = fn(K, b, x, restart=30, rel_tol=1e-10)
    assert res.success, res
    Ko = oracle.Problem(fes, ir, mass_coef=1.0, adv_coef=dt, adv=W.advection_c2(), bc=0.0)
    Mo = oracle.Problem(fes, ir, mass_coef=1.0)
    bo = Mo.apply(u0)
    assert rel_l2(b.cpu().numpy(), bo) <= F64_TOL
    rel = np.linalg.norm(bo - Ko.apply(x.cpu().numpy())) / np.linalg.norm(bo)
    print(f"{solver}: {res.iterations} iterations, {res.refinements} refinements, {res.seconds:.3f} s, "
          f"reported {res.final_rel_residual:.3e}, oracle residual {rel:.3e}")
    assert rel <= 2e-10


def sample_elements(fes, per_order, seed):
    rng = np.random.default_rng(seed)
    picks = []
    for p in (1, 2, 3):
        ids = np.flatnonzero(fes.elem_order == p)
        picks.append(rng.choice(ids, size=min(per_order, ids.size), replace=False))
    # the corners of the mesh (periodic wraps, velocity boundaries) as well
    picks.append(np.array([0, fes.mesh.num_cells - 1]))
    return np.unique(np.concatenate(picks))


@pytest.mark.parametrize("dtype,tol", [(torch.float64, F64_TOL), (torch.float32, F32_TOL)])
def test_phase_space_8pow6_padaptive_sampled_vs_oracle(dtype, tol):
    print(f"8^6 phase space p-adaptive {dtype}", flush=True)
    fes, adv = W.phase_space_6d_padaptive(8, seed=2024)
    op = G.MakeAdvectionOperator(fes, None, adv)
    N = op.num_dofs
    gen = torch.Generator(device="cuda").manual_seed(2024)
    xt = (torch.rand(N, generator=gen, device="cuda", dtype=torch.float64) * 2.0 - 1.0).to(dtype)
    yt = torch.empty_like(xt)
    op(xt, yt)
    torch.cuda.synchronize()
    print(f"  variant {op.plan.query_variant(dtype, 0.0, 1.0)}, {N} DoFs", flush=True)
    y64 = yt.double()
    cons = float(y64.sum().abs() / y64.abs().sum())
    print(f"  |sum A u| / sum |A u| = {cons:.3e}")
    assert cons <= (1e-12 if dtype == torch.float64 else 1e-9)
    x = xt.double().cpu().numpy()
    y = y64.cpu().numpy()
    del xt, yt, y64
    elems = sample_elements(fes, 400, 7)
    P = oracle.Problem(fes, None, adv_coef=1.0, adv=adv)
    ref = P.apply_elements(x, elems)
    err = rel_l2(y[P.element_dofs(elems)], ref)
    print(f"  {elems.size} sampled elements: rel l2 vs oracle = {err:.3e}")
    assert err <= tol
