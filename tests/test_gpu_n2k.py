"""GPU parity of the two-node Kronecker kernel (variant 64: 4D-6D meshes,
order 1, constant beta and sigma, no boundary data or inflow data 0) against
the CPU oracle and against the generic kernel (variant 0) on the same seeded
inputs: both bases, two rules, beta components of both signs, periodic and
bounded axes, mass, advection and their sum, FP64 and FP32.  GenDiL's own 6D
benchmark operator (product mesh, Legendre p = 1, three Gauss points,
beta = 1) is checked against its golden vector by test_gpu_apply's case
adv6dprod_ones_p1q3_gl, which runs this kernel by default.

Tolerances: FP64 <= 1e-12 relative L2 (the generic kernel and the oracle
evaluate the same integrals by quadrature; this kernel by exact 1D factor
matrices, so agreement is to rounding), FP32 <= 1e-5 against the FP64 oracle.
"""
import pytest
import torch

import gendil_amd as G
import oracle
from cases import rel_l2
from gendil_amd.fields import SeparableField, VectorField

pytestmark = pytest.mark.gpu

N2K = 64

# (dim, sizes, periodic axes, basis, q, beta per axis, bc, mass_coef, adv_coef)
CONFIGS = [
    (6, [3, 2, 3, 2, 3, 2], [1, 0, 1, 0, 0, 1], "G", 3, [1.0, -0.5, 0.25, 1.0, -1.0, 0.75], None, 0.0, 1.0),
    (6, [2, 3, 2, 3, 2, 2], [0, 0, 0, 0, 0, 0], "L", 2, [0.5, 1.0, -1.0, 0.2, 0.3, -0.4], "zero", 0.7, 0.9),
    (6, [3, 3, 3, 1, 2, 2], [1, 1, 1, 0, 0, 0], "L", 3, [1.0] * 6, None, 1.0, 0.0),
    (5, [3, 4, 2, 3, 2], [0, 1, 0, 1, 0], "G", 2, [-1.0, 0.5, 2.0, -0.25, 1.0], "zero", 0.0, 1.0),
    (5, [2, 2, 3, 3, 4], [1, 1, 1, 1, 1], "L", 3, [0.3, -0.6, 0.9, 1.2, -1.5], None, 0.4, 1.0),
    (4, [5, 3, 4, 2], [0, 0, 1, 1], "L", 2, [1.0, -2.0, 0.5, -0.5], None, 1.0, 1.0),
    (4, [3, 3, 3, 3], [0, 0, 0, 0], "G", 3, [-1.0, -1.0, 1.0, 1.0], "zero", 0.0, 1.0),
]


def build(cfg):
    D, sizes, per, basis, q, beta, bc, mc, ac = cfg
    h = [1.0 / s * (1.0 + 0.1 * d) for d, s in enumerate(sizes)]
    mesh = G.CartesianMesh(sizes, h, [0.0] * D, [bool(p) for p in per])
    fe = G.MakeLobattoFiniteElement(1) if basis == "L" else G.MakeLegendreFiniteElement(1)
    fes = G.MakeFiniteElementSpace(mesh, fe)
    ir = G.MakeIntegrationRule(q)
    adv = VectorField([SeparableField.constant(b) for b in beta])
    g = SeparableField.constant(0.0) if bc == "zero" else None
    sigma = SeparableField.constant(1.5)
    op = G.MakeMassAdvectionOperator(fes, ir, adv, sigma=sigma, bc=g, mass_coef=mc, adv_coef=ac)
    ref = oracle.Problem(fes, ir, mass_coef=mc, adv_coef=ac, sigma=sigma, adv=adv, bc=g)
    return op, ref


@pytest.mark.parametrize("cfg", CONFIGS, ids=[f"{c[0]}d_{i}" for i, c in enumerate(CONFIGS)])
def test_two_node_kronecker_matches_oracle_and_generic(cfg):
    print(f"n2k config {cfg}", flush=True)
    op, ref = build(cfg)
    mc, ac = cfg[7], cfg[8]
    x = G.fill_random(op.num_dofs, 31)
    want = ref.apply(x)
    xd = torch.as_tensor(x, device="cuda")
    assert op.plan.query_variant(torch.float64, mc, ac) == N2K
    y = torch.empty_like(xd)
    op(xd, y)
    torch.cuda.synchronize()
    assert rel_l2(y.cpu().numpy(), want) <= 1e-12
    # the generic kernel on the same plan and input
    op.plan.set_variant(torch.float64, 0)
    assert op.plan.query_variant(torch.float64, mc, ac) == 0
    yg = torch.empty_like(xd)
    op(xd, yg)
    torch.cuda.synchronize()
    assert rel_l2(y.cpu().numpy(), yg.cpu().numpy()) <= 1e-12
    op.plan.set_variant(torch.float64, -1)
    # FP32
    assert op.plan.query_variant(torch.float32, mc, ac) == N2K
    x32 = xd.float()
    y32 = torch.empty_like(x32)
    op(x32, y32)
    torch.cuda.synchronize()
    assert rel_l2(y32.double().cpu().numpy(), want) <= 1e-5


def test_two_node_kronecker_refuses_what_it_does_not_serve():
    # a varying beta component, order 2, and inflow data other than 0 keep the generic kernel
    mesh = G.CartesianMesh([2] * 6, [0.5] * 6, [0.0] * 6, [False] * 6)
    ir = G.MakeIntegrationRule(2)
    fes1 = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(1))
    varying = VectorField([SeparableField([(1.0, {0: lambda s: s})])] + [SeparableField.constant(1.0)] * 5)
    op = G.MakeAdvectionOperator(fes1, ir, varying)
    assert op.plan.query_variant(torch.float64, 0.0, 1.0) == 0
    const = VectorField([SeparableField.constant(1.0)] * 6)
    op = G.MakeAdvectionOperator(fes1, ir, const, bc=SeparableField.constant(2.0))
    assert op.plan.query_variant(torch.float64, 0.0, 1.0) == 0
    op = G.MakeAdvectionOperator(G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(2)),
                                 G.MakeIntegrationRule(3), const)
    assert op.plan.query_variant(torch.float64, 0.0, 1.0) == 0
