"""Every kernel variant (generic LDS-staged kernel, register-resident 3D
kernel with 1, 2 or 4 lanes per element, Kronecker-factored 3D kernel with
16 + lanes, its LDS-staged forms 48..55 and streaming form 56) against GenDiL's golden vectors
and the oracle, and the autotuner's choice.  A pinned variant that does not
serve a plan falls back to one that does, so every case must still be
correct."""
import numpy as np
import pytest
import torch

import gendil_amd as G
import oracle
from cases import BY_NAME, beta_field, bc_field, gpu_operator, rel_l2
from gendil_amd import workloads as W

pytestmark = pytest.mark.gpu

ELIGIBLE = ["adv3d_c2_p3q4_per", "adv3d_c2_p3q4_bc0", "adv3d_c2_p3q4_nobc", "adv3d_trig_p3q4_bcpoly",
            "mass3d_p3q4_one", "adv3d_trig_p2q3_m1per", "adv3d_const_p2q4_per", "adv3d_bubble_p1q3",
            "adv3d_bubble_p2q4", "adv3d_bubble_p3q5"]
STAGED = list(range(48, 56))
KRON = {1: [17, 18] + STAGED, 2: [17], 3: [18, 20] + STAGED}
ALL_VARIANTS = (0, 1, 2, 4, 17, 18, 20, *STAGED, 56, 57)


def variants(case):
    base = [0, 1, 2, 4] if case.P == 3 else [0, 1]
    return base + KRON.get(case.P, [])


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


@pytest.mark.parametrize("lanes", [0, 1, 2, 4, 18, 20, *STAGED, 56, 57])
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
        assert lanes in ALL_VARIANTS and ms > 0
    x = G.fill_random(op.num_dofs, 1)
    y = torch.empty(op.num_dofs, dtype=torch.float64, device="cuda")
    op(torch.as_tensor(x, device="cuda"), y)
    ref = oracle.Problem(fes, ir, adv_coef=1.0, adv=W.advection_c2()).apply(x)
    assert rel_l2(y.cpu().numpy(), ref) <= 1e-12


def test_kronecker_kernel_is_the_default_for_the_benchmark_field():
    mesh = G.CartesianMesh.unit(3, 8, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    op = G.MakeAdvectionOperator(fes, G.MakeIntegrationRule(4), W.advection_c2())
    # staged Kronecker kernel, outputs staged through LDS (16 + 32 + 1),
    # FP32 in 128-thread blocks (+ 2)
    assert op.plan.query_variant(torch.float64, 0.0, 1.0) == 49
    assert op.plan.query_variant(torch.float32, 0.0, 1.0) == 51
    # rows of a multiple of 16 elements: the streaming kernel
    fes16 = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, 16, True), G.MakeLobattoFiniteElement(3))
    op16 = G.MakeAdvectionOperator(fes16, G.MakeIntegrationRule(4), W.advection_c2())
    assert op16.plan.query_variant(torch.float64, 0.0, 1.0) == 56
    assert op16.plan.query_variant(torch.float32, 0.0, 1.0) == 56
    # a field whose sign changes along a face is served by the register kernel
    bad = G.MakeAdvectionOperator(fes, G.MakeIntegrationRule(4), beta_field("trig", 3))
    assert bad.plan.query_variant(torch.float64, 0.0, 1.0) in (1, 2, 4)


KRON_CASES = [
    # (p, n, periodic, bc, beta kind, mass_coef, adv_coef)
    (1, 5, True, None, "c2", 0.0, 1.0),
    (2, 4, False, 0.0, "c2", 0.0, 1.0),
    (2, 4, False, None, "c2", 1.0, 0.3),
    (3, 3, False, 0.0, "const", 0.0, 1.0),
    (3, 4, True, None, "negc2", 0.0, 1.0),
    (3, 3, False, None, "negc2", 0.7, 1.0),
]


def negc2():
    # the benchmark field with every component negated (upwind sides swap)
    from gendil_amd.fields import SeparableField, VectorField
    comps = W.advection_c2().components
    return VectorField([SeparableField([(-c, f) for c, f in comp.terms]) for comp in comps])


@pytest.mark.parametrize("case", KRON_CASES, ids=[f"p{c[0]}_n{c[1]}_{c[4]}_bc{c[3]}_m{c[5]}" for c in KRON_CASES])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_kronecker_kernel_vs_oracle(case, dtype):
    p, n, periodic, bc, kind, mc, ac = case
    mesh = G.CartesianMesh.unit(3, n, periodic)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(p))
    ir = G.MakeIntegrationRule(p + 1)
    adv = negc2() if kind == "negc2" else beta_field(kind, 3)
    op = G.MakeMassAdvectionOperator(fes, ir, adv, bc=bc, mass_coef=mc, adv_coef=ac)
    chosen = op.plan.query_variant(dtype, mc, ac)
    print(f"kronecker case {case} {dtype}: variant {chosen}", flush=True)
    assert chosen >= 16
    x = G.fill_random(op.num_dofs, 9)
    y = torch.empty(op.num_dofs, dtype=dtype, device="cuda")
    op(torch.as_tensor(x, dtype=dtype, device="cuda"), y)
    ref = oracle.Problem(fes, ir, mass_coef=mc, adv_coef=ac, adv=adv, bc=bc).apply(x)
    err = rel_l2(y.double().cpu().numpy(), ref)
    print(f"rel l2 = {err:.3e}")
    assert err <= (1e-12 if dtype == torch.float64 else 1e-5)


STAGED_MESHES = [
    # (sizes, periodic): rows longer than a block, rows shorter than a block,
    # a partial last block, a single-element axis
    ((70, 3, 2), (True, False, True)),
    ((5, 7, 3), (False, True, True)),
    ((9, 1, 4), (True, True, False)),
]


@pytest.mark.parametrize("mesh_case", STAGED_MESHES, ids=["rows70", "rows5", "axis1"])
@pytest.mark.parametrize("p", [1, 3])
@pytest.mark.parametrize("kind", ["c2", "negc2", "const"])
def test_staged_kronecker_variants_on_uneven_meshes(mesh_case, p, kind):
    sizes, periodic = mesh_case
    mesh = G.CartesianMesh(sizes, [1.0 / s for s in sizes], [0.0] * 3, periodic)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(p))
    ir = G.MakeIntegrationRule(p + 1)
    adv = negc2() if kind == "negc2" else beta_field(kind, 3)
    for mc, ac, bc in ((0.0, 1.0, None), (0.6, 0.8, 0.0)):
        op = G.MakeMassAdvectionOperator(fes, ir, adv, bc=bc, mass_coef=mc, adv_coef=ac)
        x = G.fill_random(op.num_dofs, 21)
        ref = oracle.Problem(fes, ir, mass_coef=mc, adv_coef=ac, adv=adv, bc=bc).apply(x)
        for dtype in (torch.float64, torch.float32):
            for v in STAGED:
                print(f"staged {sizes} p={p} {kind} m={mc} {dtype} variant {v}", flush=True)
                op.plan.set_variant(dtype, v)
                assert op.plan.query_variant(dtype, mc, ac) == v
                y = torch.empty(op.num_dofs, dtype=dtype, device="cuda")
                op(torch.as_tensor(x, dtype=dtype, device="cuda"), y)
                err = rel_l2(y.double().cpu().numpy(), ref)
                assert err <= (1e-12 if dtype == torch.float64 else 1e-5), (v, err)


def test_staged_kernel_unaligned_vectors_fall_back():
    """The staged kernels move 16-byte chunks; vectors that are not 16-byte
    aligned are served by the register Kronecker kernel instead."""
    mesh = G.CartesianMesh.unit(3, 6, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    op = G.MakeAdvectionOperator(fes, ir, W.advection_c2())
    N = op.num_dofs
    x = G.fill_random(N, 4)
    ref = oracle.Problem(fes, ir, adv_coef=1.0, adv=W.advection_c2()).apply(x)
    xb = torch.zeros(N + 1, dtype=torch.float64, device="cuda")
    yb = torch.zeros(N + 1, dtype=torch.float64, device="cuda")
    xb[1:] = torch.as_tensor(x, device="cuda")
    op.plan.set_variant(torch.float64, 53)
    op(xb[1:], yb[1:])
    assert rel_l2(yb[1:].cpu().numpy(), ref) <= 1e-12
    assert float(yb[0]) == 0.0


STREAM_MESHES = [
    # (sizes, periodic): a grid that is not a multiple of the 8 XCDs (13
    # tiles), single-element y / z axes, and a mesh of 4096 tiles, several
    # per wave, so every wave runs the double-buffered loop
    ((16, 13, 1), (False, True, True)),
    ((32, 1, 5), (True, False, False)),
    ((48, 3, 2), (True, True, False)),
    ((64, 64, 16), (False, False, True)),
]


@pytest.mark.parametrize("mesh_case", STREAM_MESHES, ids=["x16y13", "y1", "x48", "tiles4096"])
@pytest.mark.parametrize("kind", ["c2", "negc2", "const"])
@pytest.mark.parametrize("variant", [56, 57])
def test_streaming_kernel_vs_oracle(mesh_case, kind, variant):
    """Variants 56 and 57 (one wave per block, persistent over tiles of 16
    elements of one x-row; 56: one tile buffer, refilled as soon as the
    tile's outputs are in registers, two waves per SIMD; 57: two tile
    buffers, the next tile's copies in flight while a tile computes)."""
    sizes, periodic = mesh_case
    mesh = G.CartesianMesh(sizes, [1.0 / s for s in sizes], [0.0] * 3, periodic)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    ir = G.MakeIntegrationRule(4)
    adv = negc2() if kind == "negc2" else beta_field(kind, 3)
    for mc, ac, bc in ((0.0, 1.0, None), (0.6, 0.8, 0.0), (1.0, 0.0, None)):
        op = G.MakeMassAdvectionOperator(fes, ir, adv, bc=bc, mass_coef=mc, adv_coef=ac)
        x = G.fill_random(op.num_dofs, 23)
        ref = oracle.Problem(fes, ir, mass_coef=mc, adv_coef=ac, adv=adv, bc=bc).apply(x)
        for dtype in (torch.float64, torch.float32):
            print(f"stream {variant} {sizes} {kind} m={mc} a={ac} {dtype}", flush=True)
            op.plan.set_variant(dtype, variant)
            assert op.plan.query_variant(dtype, mc, ac) == variant
            y = torch.full((op.num_dofs,), float("nan"), dtype=dtype, device="cuda")
            op(torch.as_tensor(x, dtype=dtype, device="cuda"), y)
            err = rel_l2(y.double().cpu().numpy(), ref)
            assert err <= (1e-12 if dtype == torch.float64 else 1e-5), (sizes, kind, mc, dtype, err)


def test_streaming_kernel_needs_whole_tiles():
    """Rows that are not a multiple of 16 elements are served by another variant."""
    mesh = G.CartesianMesh([20, 4, 4], [0.05, 0.25, 0.25], [0.0] * 3, True)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    op = G.MakeAdvectionOperator(fes, G.MakeIntegrationRule(4), W.advection_c2())
    for v in (56, 57):
        op.plan.set_variant(torch.float64, v)
        assert op.plan.query_variant(torch.float64, 0.0, 1.0) != v
    x = G.fill_random(op.num_dofs, 2)
    y = torch.empty(op.num_dofs, dtype=torch.float64, device="cuda")
    op(torch.as_tensor(x, device="cuda"), y)
    ref = oracle.Problem(fes, G.MakeIntegrationRule(4), adv_coef=1.0, adv=W.advection_c2()).apply(x)
    assert rel_l2(y.cpu().numpy(), ref) <= 1e-12


def one_axis_field(kind):
    """One-axis fields whose axis signatures have no compiled instance of the
    register kernel: the run-time-axis instance serves them."""
    from gendil_amd.fields import SeparableField, VectorField
    if kind == "rotating":  # (0.5 - y, x - 0.5, 0.25): signs change inside the mesh
        return VectorField([SeparableField([(1.0, {1: lambda s: 0.5 - s})]),
                            SeparableField([(1.0, {0: lambda s: s - 0.5})]), SeparableField.constant(0.25)])
    if kind == "own_axis":  # (sin 2 pi x, -0.5, 0.3 + cos 2 pi z): beta_d varies along d
        return VectorField([SeparableField([(1.0, {0: lambda s: np.sin(2 * np.pi * s)})]),
                            SeparableField.constant(-0.5),
                            SeparableField([(1.0, {2: lambda s: 0.3 + np.cos(2 * np.pi * s)})])])
    raise ValueError(kind)


RUNTIME_AXIS_CASES = [
    # (p, n, periodic, kind, sigma y-dependent, bc, mass_coef, adv_coef, variants)
    (3, 5, True, "rotating", False, None, 0.0, 1.0, (2, 4)),
    (3, 4, False, "own_axis", True, "poly", 0.7, 1.0, (2, 4)),
    (3, 4, False, "rotating", False, "zero", 0.0, 1.0, (2, 4)),
    (2, 5, False, "own_axis", False, "zero", 1.0, 0.5, (1,)),
    (1, 6, True, "rotating", True, None, 0.3, 1.0, (1,)),
]


@pytest.mark.parametrize("case", RUNTIME_AXIS_CASES,
                         ids=[f"p{c[0]}_n{c[1]}_{c[3]}_bc{c[5]}_m{c[6]}" for c in RUNTIME_AXIS_CASES])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_register_kernel_runtime_axes_vs_oracle(case, dtype):
    """Fields with any one-axis signature (rotating, varying along their own
    axis, mixed signs) run on the register kernel, not the generic one, and
    match the oracle."""
    from gendil_amd.fields import SeparableField
    p, n, periodic, kind, sig_y, bc, mc, ac, lanes_list = case
    mesh = G.CartesianMesh.unit(3, n, periodic)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(p))
    ir = G.MakeIntegrationRule(p + 1)
    beta = one_axis_field(kind)
    sigma = SeparableField([(1.0, {1: lambda s: 1.0 + s})]) if sig_y else None
    g = bc_field(bc, 3) if bc else None
    op = G.MakeMassAdvectionOperator(fes, ir, beta, sigma=sigma, bc=g, mass_coef=mc, adv_coef=ac)
    x = G.fill_random(op.num_dofs, 3)
    ref = oracle.Problem(fes, ir, mass_coef=mc, adv_coef=ac, sigma=sigma, adv=beta, bc=g).apply(x)
    if dtype == torch.float32:
        lanes_list = (1,) if p < 3 else (2, 4)
    for lanes in (-1, *lanes_list):
        op.plan.set_variant(dtype, lanes)
        v = op.plan.query_variant(dtype, mc, ac)
        print(f"runtime axes {case} {dtype} lanes={lanes} -> variant {v}", flush=True)
        assert v in (1, 2, 4), v
        y = torch.empty(op.num_dofs, dtype=dtype, device="cuda")
        op(torch.as_tensor(x, dtype=dtype, device="cuda"), y)
        err = rel_l2(y.double().cpu().numpy(), ref)
        assert err <= (1e-12 if dtype == torch.float64 else 1e-5), (lanes, err)
