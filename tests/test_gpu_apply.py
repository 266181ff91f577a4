"""GPU parity: the gfx950 kernels (through the C-ABI) against GenDiL's golden
vectors and the CPU oracle on the same seeded inputs.

Tolerances: FP64 <= 1e-12 relative L2 (GenDiL's own cross-implementation bar,
tests/integration/gendil/test-mass-3d-generic.cpp:196); FP32 results are
compared with the FP64 oracle at <= 1e-5 relative L2 (stated, not a claim of
bit-level agreement).
"""
import numpy as np
import pytest
import torch

import gendil_amd as G
import oracle
from cases import CASES, beta_field, bc_field, gpu_operator, oracle_problem, rel_l2
from gendil_amd import workloads as W
from gendil_amd.fields import face_counts

pytestmark = pytest.mark.gpu

DEV = "cuda"
F64_TOL = 1e-12
F32_TOL = 1e-5


def golden(case):
    return np.load(f"{oracle.HERE.parent}/tests/golden/{case.name}.npz")


def to_dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_fp64_apply_matches_gendil_golden(case):
    print(f"case {case.name}", flush=True)
    g = golden(case)
    x = G.fill_random(int(g["n"]), case.seed)
    op = gpu_operator(case)
    y = torch.empty(op.num_dofs, dtype=torch.float64, device=DEV)
    op(to_dev(x), y)
    torch.cuda.synchronize()
    yg = y.cpu().numpy()
    assert rel_l2(yg, g["y"]) <= F64_TOL
    assert rel_l2(yg, oracle_problem(case).apply(x)) <= F64_TOL


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_fp32_apply_close_to_fp64_oracle(case):
    print(f"case {case.name}", flush=True)
    g = golden(case)
    x = G.fill_random(int(g["n"]), case.seed)
    op = gpu_operator(case)
    y = torch.empty(op.num_dofs, dtype=torch.float32, device=DEV)
    op(to_dev(x, torch.float32), y)
    torch.cuda.synchronize()
    err = rel_l2(y.double().cpu().numpy(), g["y"])
    print(f"fp32 rel err vs fp64 reference = {err:.3e}")
    assert err <= F32_TOL


@pytest.mark.parametrize("name", ["adv3d_c2_p3q4_bc0", "adv3d_trig_p3q4_bcpoly", "adv2d_trig_p3q4_per",
                                  "adv4d_trig_p2q3_bcpoly", "adv3d_bubble_p2q4", "adv6d_phase_p1q2_per"])
def test_combined_mass_plus_advection(name):
    # K = M + dt A in one fused kernel vs the oracle's M and A applied separately
    case = [c for c in CASES if c.name == name][0]
    print(f"case {name}", flush=True)
    fes, ir = case.space()
    dt = 0.37
    op = G.MakeMassAdvectionOperator(fes, ir, beta_field(case.coef, case.D), sigma=None,
                                     bc=bc_field(case.bc, case.D), mass_coef=1.0, adv_coef=dt)
    x = G.fill_random(op.num_dofs, 99)
    y = torch.empty(op.num_dofs, dtype=torch.float64, device=DEV)
    op(to_dev(x), y)
    ref = oracle.Problem(fes, ir, mass_coef=1.0, adv_coef=dt, adv=beta_field(case.coef, case.D),
                         bc=bc_field(case.bc, case.D)).apply(x)
    assert rel_l2(y.cpu().numpy(), ref) <= F64_TOL


@pytest.mark.parametrize("name", ["adv3d_c2_p3q4_per", "adv3d_c2_p3q4_bc0", "adv3d_trig_gl_p3q4_bcpoly",
                                  "adv2d_trig_p2q3_bcpoly", "adv4d_trig_p1q2_per", "adv3d_trig_p2q3_m1per",
                                  "adv1d_const_p3q4_bcpoly", "adv6d_phase_p2q3_bc0"])
def test_transpose_matches_oracle(name):
    case = [c for c in CASES if c.name == name][0]
    print(f"case {name}", flush=True)
    fes, ir = case.space()
    op = gpu_operator(case)
    r = G.fill_random(op.num_dofs, 4242)
    z = torch.empty(op.num_dofs, dtype=torch.float64, device=DEV)
    op.transpose(to_dev(r), z)
    ref = oracle_problem(case).apply_transpose(r)
    assert rel_l2(z.cpu().numpy(), ref) <= F64_TOL
    # dot-product identity <A^T r, v> = <r, A v> on the GPU (g = 0 linear part)
    if case.bc in ("none", "zero"):
        v = G.fill_random(op.num_dofs, 77)
        Av = torch.empty_like(z)
        op(to_dev(v), Av)
        lhs = float(torch.dot(z, to_dev(v)))
        rhs = float(torch.dot(to_dev(r), Av))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_fp64_full_size_c2_properties():
    # 64^3 periodic p=3 (the benchmark workload): conservation and linearity
    mesh = G.CartesianMesh.unit(3, 64, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    op = G.MakeAdvectionOperator(fes, G.MakeIntegrationRule(4), W.advection_c2())
    N = op.num_dofs
    print(f"64^3 p=3 periodic, {N} dofs", flush=True)
    u = to_dev(G.fill_random(N, 12345))
    v = to_dev(G.fill_random(N, 54321))
    y = torch.empty_like(u)
    op(u, y)
    assert abs(float(y.sum())) <= 1e-12 * float(y.abs().sum())
    yv = torch.empty_like(u)
    op(v, yv)
    w = torch.empty_like(u)
    op(2.0 * u - 3.0 * v, w)
    assert float((w - (2.0 * y - 3.0 * yv)).norm()) <= 1e-13 * float(w.norm())


def test_fp64_matches_oracle_at_32_cubed():
    mesh = G.CartesianMesh.unit(3, 32, False)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    dt = 0.5 * (1.0 / 32) / W.advection_c2_max_norm()
    op = G.MakeMassAdvectionOperator(fes, ir, W.advection_c2(), bc=0.0, mass_coef=1.0, adv_coef=dt)
    x = G.fill_random(op.num_dofs, 12345)
    y = torch.empty(op.num_dofs, dtype=torch.float64, device=DEV)
    op(to_dev(x), y)
    ref = oracle.Problem(fes, ir, mass_coef=1.0, adv_coef=dt, adv=W.advection_c2(), bc=0.0).apply(x)
    assert rel_l2(y.cpu().numpy(), ref) <= F64_TOL


def test_grad_beta_matches_oracle_and_fd():
    rng = np.random.default_rng(3)
    mesh = G.CartesianMesh.unit(3, 3, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(2))
    ir = G.MakeIntegrationRule(3)
    D, q = 3, 3
    cell = rng.standard_normal(mesh.num_cells * D * q ** D)
    faces = [rng.standard_normal(fc * q ** (D - 1)) for fc in face_counts(mesh)]
    dcell_t = to_dev(cell)
    dfaces_t = [to_dev(f) for f in faces]
    op = G.MakeAdvectionOperator(fes, ir, G.PointwiseField(dcell_t, dfaces_t))
    N = op.num_dofs
    u = rng.uniform(-1, 1, N)
    r = rng.uniform(-1, 1, N)
    gc = torch.empty_like(dcell_t)
    gf = [torch.empty_like(f) for f in dfaces_t]
    op.plan.grad_beta(1.0, to_dev(u), to_dev(r), gc, gf)
    oc, of = oracle.Problem(fes, ir, adv_coef=1.0, adv=(cell, faces)).grad_beta(u, r, face_counts(mesh))
    assert rel_l2(gc.cpu().numpy(), oc) <= F64_TOL
    for a, b in zip(gf, of):
        assert rel_l2(a.cpu().numpy(), b) <= F64_TOL
    y = torch.empty(N, dtype=torch.float64, device=DEV)
    op(to_dev(u), y)
    yo = oracle.Problem(fes, ir, adv_coef=1.0, adv=(cell, faces)).apply(u)
    assert rel_l2(y.cpu().numpy(), yo) <= F64_TOL


def test_apply_runs_on_the_current_torch_stream():
    """The apply launches on torch's current stream, read through the raw
    stream query: it matches torch.cuda.current_stream() on the default
    stream and inside a side stream, and an apply queued on a side stream
    behind a slow producer sees the producer's values."""
    from gendil_amd import operators as O
    # (a null stream handle reads back as None from a c_void_p)
    assert (O._stream().value or 0) == torch.cuda.current_stream().cuda_stream
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        assert (O._stream().value or 0) == side.cuda_stream != 0
    mesh = G.CartesianMesh.unit(3, 8, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    op = G.MakeAdvectionOperator(fes, ir, W.advection_c2())
    x = G.fill_random(op.num_dofs, 11)
    ref = oracle.Problem(fes, ir, adv_coef=1.0, adv=W.advection_c2()).apply(x)
    xd = torch.as_tensor(x, device="cuda")
    big = torch.rand(1 << 24, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        xs = torch.zeros_like(xd)
        for _ in range(4):  # a slow producer ahead of the copy on the same stream
            big.mul_(1.0000001)
        xs.copy_(xd)
        y = torch.empty_like(xd)
        op(xs, y)
    side.synchronize()
    assert rel_l2(y.cpu().numpy(), ref) <= 1e-12
