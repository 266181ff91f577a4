"""Every kernel variant (generic LDS-staged kernel, register-resident 3D
kernel with 1, 2 or 4 lanes per element) against GenDiL's golden vectors,
and the autotuner's choice."""
import numpy as np
import pytest
import torch

import gendil_amd as G
import oracle
from cases import BY_NAME, beta_field, bc_field, gpu_operator, rel_l2
from gendil_amd import workloads as W

pytestmark = pytest.mark.gpu

ELIGIBLE = ["adv3d_c2_p3q4_per", "adv3d_c2_p3q4_bc0", "adv3d_c2_p3q4_nobc", "adv3d_trig_p3q4_bcpoly",
            "mass3d_p3q4_one", "adv3d_trig_p2q3_m1per"]


def variants(case):
    return [0, 1, 2, 4] if case.P == 3 else [0, 1]


@pytest.mark.parametrize("name", ELIGIBLE)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_all_variants_match_golden(name, dtype):
    case = BY_NAME[name]
    g = np.load(f"{oracle.HERE.parent}/tests/golden/{name}.npz")
    x = G.fill_random(int(g["n"]), case.seed)
    op = gpu_operator(case)
    for lanes in variants(case):
        print(f"case {name} {dtype} lanes={lanes}", flush=True)
        op.plan.set_variant(dtype, lanes)
        y = torch.empty(op.num_dofs, dtype=dtype, device="cuda")
        op(torch.as_tensor(x, dtype=dtype, device="cuda"), y)
        err = rel_l2(y.double().cpu().numpy(), g["y"])
        assert err <= (1e-12 if dtype == torch.float64 else 1e-5), (lanes, err)


@pytest.mark.parametrize("lanes", [0, 1, 2, 4])
def test_variants_combined_operator_at_16_cubed(lanes):
    mesh = G.CartesianMesh.unit(3, 16, False)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    dt = 0.5 * (1.0 / 16) / W.advection_c2_max_norm()
    op = G.MakeMassAdvectionOperator(fes, ir, W.advection_c2(), bc=0.0, mass_coef=1.0, adv_coef=dt)
    op.plan.set_variant(torch.float64, lanes)
    x = G.fill_random(op.num_dofs, 5)
    y = torch.empty(op.num_dofs, dtype=torch.float64, device="cuda")
    op(torch.as_tensor(x, device="cuda"), y)
    ref = oracle.Problem(fes, ir, mass_coef=1.0, adv_coef=dt, adv=W.advection_c2(), bc=0.0).apply(x)
    assert rel_l2(y.cpu().numpy(), ref) <= 1e-12


def test_autotune_picks_a_correct_variant():
    mesh = G.CartesianMesh.unit(3, 24, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    op = G.MakeAdvectionOperator(fes, ir, W.advection_c2())
    for dtype in (torch.float64, torch.float32):
        lanes, ms = op.plan.autotune(dtype, 0.0, 1.0, reps=3)
        print(f"autotune {dtype}: lanes={lanes} {ms:.4f} ms", flush=True)
        assert lanes in (0, 1, 2, 4) and ms > 0
    x = G.fill_random(op.num_dofs, 1)
    y = torch.empty(op.num_dofs, dtype=torch.float64, device="cuda")
    op(torch.as_tensor(x, device="cuda"), y)
    ref = oracle.Problem(fes, ir, adv_coef=1.0, adv=W.advection_c2()).apply(x)
    assert rel_l2(y.cpu().numpy(), ref) <= 1e-12
