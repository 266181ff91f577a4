"""Face-centric advection (GenDiL's AdvectionFaceOperator and its
nonconforming, intermesh variant).

CPU: the numpy oracle against GenDiL's own operators (golden vectors written
by oracle/_ref/gendil_ref, tests/golden/make_golden.py), conservation, and the
split-mesh identity of tests/integration/gendil/test-face-advection-split-
mesh.cpp.  GPU (marked): the HIP kernels through the C-ABI against the same
golden vectors at <= 1e-12 relative L2 (FP64), accumulation into a nonzero y,
and the size-independent identities at larger sizes.
"""
import hashlib

import numpy as np
import pytest

import gendil_amd as G
import oracle
from cases import FACE_CASES, beta_numpy, rel_l2

F64_TOL = 1e-12


def golden(case):
    return np.load(f"{oracle.HERE.parent}/tests/golden/{case.name}.npz")


def split_pieces(nxl, nxr, ny, nz, p, q, kind):
    """Oracle pieces of the split-mesh test: the full mesh's interior faces,
    the left and right meshes' interior faces and the -x interface of the
    right mesh (minus) with the left mesh (plus), and the permutation from the
    full layout to [left | right]."""
    h = 1.0 / nxl
    n = p + 1
    nd = n ** 3
    beta = beta_numpy(kind, 3)
    full = dict(sizes=[nxl + nxr, ny, nz], h=[h] * 3, origin=[0.0] * 3, n=n, basis=0)
    left = dict(sizes=[nxl, ny, nz], h=[h] * 3, origin=[0.0] * 3, n=n, basis=0)
    right = dict(sizes=[nxr, ny, nz], h=[h] * 3, origin=[1.0, 0.0, 0.0], n=n, basis=0)
    perm = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nxl + nxr):
                if i < nxl:
                    e = i + nxl * (j + ny * k)
                else:
                    e = nxl * ny * nz + (i - nxl) + nxr * (j + ny * k)
                perm.extend(range(e * nd, (e + 1) * nd))
    return full, left, right, beta, np.asarray(perm)


@pytest.mark.parametrize("case", FACE_CASES, ids=lambda c: c.name)
def test_face_oracle_matches_gendil_golden(case):
    g = golden(case)
    x = G.fill_random(int(g["n"]), case.seed)
    assert hashlib.sha256(x.tobytes()).hexdigest() == str(g["x_sha256"])
    assert rel_l2(case.oracle(x), g["y"]) <= 1e-13


@pytest.mark.parametrize("case", FACE_CASES, ids=lambda c: c.name)
def test_face_oracle_conserves(case):
    # every face flux enters once with + and once with -, and the GLL/GL
    # Lagrange bases sum to one: sum(y) = 0 for any u
    # (test-advection-p-adaptivity.cpp:92-106 checks it for u = 1)
    x = G.fill_random(int(golden(case)["n"]), 99)
    y = case.oracle(x)
    assert abs(y.sum()) <= 1e-13 * np.abs(y).sum()


def test_face_oracle_split_mesh_identity():
    # test-face-advection-split-mesh.cpp: the full mesh's interior faces equal
    # the two halves' interior faces plus their interface, up to the layout
    full, left, right, beta, perm = split_pieces(3, 3, 3, 3, 2, 4, "const")
    nl = 27 * 27
    u_full = G.fill_random(2 * nl, 5)
    y_full = np.zeros(2 * nl)
    oracle.face_operator(full["sizes"], full["h"], full["origin"], 3, 0, 4, beta, u_full, y_full)
    u = np.empty(2 * nl)
    u[perm] = u_full
    y = np.zeros(2 * nl)
    oracle.face_operator(left["sizes"], left["h"], left["origin"], 3, 0, 4, beta, u[:nl], y[:nl])
    oracle.face_operator(right["sizes"], right["h"], right["origin"], 3, 0, 4, beta, u[nl:], y[nl:])
    oracle.intermesh_operator(0, 4, right, left, beta, u[nl:], u[:nl], y[nl:], y[:nl])
    assert rel_l2(y[perm], y_full) <= 1e-14


def test_face_oracle_planted_sign_error_is_caught():
    # a plausible mistake (the normal's sign flipped, i.e. the wrong side
    # taken as upwind) must fail the golden comparison
    case = FACE_CASES[0]
    g = golden(case)
    x = G.fill_random(int(g["n"]), case.seed)
    m, _ = case.sides()
    y = np.zeros_like(x)
    beta = beta_numpy(case.coef, case.D)
    for axis in range(case.D):
        oracle.advection_faces(axis, +1, case.Q, m, m, oracle.interior_face_pairs(m["sizes"], axis), beta, x, x, y, y)
    assert rel_l2(y, g["y"]) > 1e-3


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def _dev(a):
    import torch

    return torch.as_tensor(np.asarray(a, dtype=np.float64), device="cuda")


@pytest.mark.gpu
@pytest.mark.parametrize("case", FACE_CASES, ids=lambda c: c.name)
def test_face_gpu_matches_gendil_golden(case):
    g = golden(case)
    n = int(g["n"])
    op = case.gpu_operator()
    x = _dev(G.fill_random(n, case.seed))
    y0 = G.fill_random(n, 4242)
    y = _dev(y0)
    op(x, y)  # accumulates, as GenDiL's read-write view
    got = y.cpu().numpy() - y0
    assert rel_l2(got, g["y"]) <= F64_TOL, case.name


@pytest.mark.gpu
@pytest.mark.parametrize("p,n", [(3, 16), (2, 12)])
def test_face_gpu_large_conservation_and_oracle(p, n):
    # 3D interior faces at a size with several blocks per colour; checked
    # against the oracle and for conservation
    import torch

    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, n, False), G.MakeLobattoFiniteElement(p))
    from cases import beta_field

    op = G.MakeAdvectionFaceOperator(fes, G.MakeCartesianInteriorFaceConnectivity(fes.mesh.sizes),
                                     G.MakeIntegrationRule(p + 1), beta_field("c2", 3))
    x = G.fill_random(fes.num_dofs, 17)
    y = torch.zeros(fes.num_dofs, dtype=torch.float64, device="cuda")
    op(_dev(x), y)
    got = y.cpu().numpy()
    ref = np.zeros_like(x)
    oracle.face_operator([n] * 3, [1.0 / n] * 3, [0.0] * 3, p + 1, 0, p + 1, beta_numpy("c2", 3), x, ref)
    assert rel_l2(got, ref) <= F64_TOL
    assert abs(got.sum()) <= 1e-13 * np.abs(got).sum()


@pytest.mark.gpu
def test_face_gpu_split_mesh_identity():
    # test-face-advection-split-mesh.cpp on the GPU at 16 x 8 x 8 per half, p = 2
    import torch

    nxl, ny, nz, p = 16, 8, 8, 2
    _, _, _, _, perm = split_pieces(nxl, nxl, ny, nz, p, p + 2, "const")
    h = 1.0 / nxl
    fe = G.MakeLobattoFiniteElement(p)
    ir = G.MakeIntegrationRule(p + 2)
    from cases import beta_field

    beta = beta_field("const", 3)
    mf = G.Cartesian3DMesh(h, 2 * nxl, ny, nz)
    ml = G.Cartesian3DMesh(h, nxl, ny, nz)
    mr = G.Cartesian3DMesh(h, nxl, ny, nz, origin=(1.0, 0.0, 0.0))
    ff = G.MakeFiniteElementSpace(mf, fe)
    fl = G.MakeFiniteElementSpace(ml, fe)
    fr = G.MakeFiniteElementSpace(mr, fe, G.L2Restriction(fl.num_dofs))
    op_full = G.MakeAdvectionFaceOperator(ff, G.MakeCartesianInteriorFaceConnectivity(mf.sizes), ir, beta)
    op_l = G.MakeAdvectionFaceOperator(fl, G.MakeCartesianInteriorFaceConnectivity(ml.sizes), ir, beta)
    op_r = G.MakeAdvectionFaceOperator(fr, G.MakeCartesianInteriorFaceConnectivity(mr.sizes), ir, beta)
    op_if = G.MakeAdvectionFaceOperator(fr, fl, G.CartesianIntermeshFaceConnectivity(3, 0, mr.sizes, ml.sizes), ir,
                                        beta)
    N = ff.num_dofs
    u_full = G.fill_random(N, 5)
    u = np.empty(N)
    u[perm] = u_full
    y_full = torch.zeros(N, dtype=torch.float64, device="cuda")
    op_full(_dev(u_full), y_full)
    y = torch.zeros(N, dtype=torch.float64, device="cuda")
    ud = _dev(u)
    op_l(ud, y)
    op_r(ud, y)
    op_if(ud, y)
    assert rel_l2(y.cpu().numpy()[perm], y_full.cpu().numpy()) <= 1e-13


@pytest.mark.gpu
def test_face_gpu_rejects_unsupported_inputs():
    import torch

    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, 2, False), G.MakeLobattoFiniteElement(1))
    op = G.MakeAdvectionFaceOperator(fes, G.MakeCartesianInteriorFaceConnectivity(fes.mesh.sizes),
                                     G.MakeIntegrationRule(2), G.AffineField(np.eye(3), np.zeros(3)))
    x = torch.zeros(fes.num_dofs, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        op(x, x)  # in place would race
    with pytest.raises(ValueError):
        G.MakeAdvectionFaceOperator(fes, G.MakeCartesianInteriorFaceConnectivity((2, 2, 3)), G.MakeIntegrationRule(2),
                                    G.AffineField(np.eye(3), np.zeros(3)))
    # rules that differ between the two sides
    m = G.Cartesian3DMesh(0.5, 2, 2, 2)
    fm = G.MakeFiniteElementSpace(m, G.MakeLobattoFiniteElement(1))
    fp = G.MakeFiniteElementSpace(m, G.MakeLobattoFiniteElement(2), G.L2Restriction(fm.num_dofs))
    from gendil_amd.operators import AdvectionNonconformingFaceOperator, Plan

    pm, pp = Plan(fm, G.MakeIntegrationRule(2)), Plan(fp, G.MakeIntegrationRule(3))
    pm.set_beta(G.AffineField(np.eye(3), np.zeros(3)))
    bad = AdvectionNonconformingFaceOperator(pm, pp, 3)
    v = torch.zeros(fm.num_dofs + fp.num_dofs, dtype=torch.float64, device="cuda")
    with pytest.raises(G._lib.GdError, match="same rule"):
        bad(v, v.clone())


@pytest.mark.gpu
@pytest.mark.parametrize(
    "D,p,q,n,coef",
    [
        (3, 1, 2, 4, "c2"),    # one launch, sizes fixed at compile time
        (3, 4, 5, 3, "trig"),  # one launch, sizes read at run time
        (6, 2, 3, 2, "phase"),  # faces too large for one launch: colours per axis
        (1, 3, 4, 7, "const"),
    ],
)
def test_face_gpu_kernel_paths_against_oracle(D, p, q, n, coef):
    import torch
    from cases import beta_field

    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(D, n, False), G.MakeLobattoFiniteElement(p))
    op = G.MakeAdvectionFaceOperator(fes, G.MakeCartesianInteriorFaceConnectivity(fes.mesh.sizes),
                                     G.MakeIntegrationRule(q), beta_field(coef, D))
    x = G.fill_random(fes.num_dofs, 23)
    y0 = G.fill_random(fes.num_dofs, 24)
    y = _dev(y0)
    op(_dev(x), y)
    ref = np.zeros_like(x)
    oracle.face_operator([n] * D, [1.0 / n] * D, [0.0] * D, p + 1, 0, q, beta_numpy(coef, D), x, ref)
    assert rel_l2(y.cpu().numpy() - y0, ref) <= F64_TOL


@pytest.mark.gpu
def test_hmesh_gpu_ratio_one_is_the_conforming_interface_and_conserves():
    # an h-interface with one fine cell per coarse cell is the conforming
    # interface; with two per axis the fluxes still cancel (sum y = 0)
    import torch
    from cases import beta_field

    n, p = 12, 3
    h = 1.0 / n
    fe = G.MakeLobattoFiniteElement(p)
    ir = G.MakeIntegrationRule(p + 1)
    beta = beta_field("c2", 3)
    mc = G.Cartesian3DMesh(h, n, n, n)
    mf = G.Cartesian3DMesh(h, n, n, n, origin=(1.0, 0.0, 0.0))
    fc = G.MakeFiniteElementSpace(mc, fe)
    ff = G.MakeFiniteElementSpace(mf, fe, G.L2Restriction(fc.num_dofs))
    op_h = G.MakeAdvectionFaceOperator(fc, ff, G.NonconformingCartesianIntermeshFaceConnectivity(3, 3, mc.sizes, mf.sizes),
                                       ir, beta)
    op_c = G.MakeAdvectionFaceOperator(fc, ff, G.CartesianIntermeshFaceConnectivity(3, 3, mc.sizes, mf.sizes), ir, beta)
    N = fc.num_dofs + ff.num_dofs
    x = _dev(G.fill_random(N, 31))
    y_h = torch.zeros(N, dtype=torch.float64, device="cuda")
    y_c = torch.zeros(N, dtype=torch.float64, device="cuda")
    op_h(x, y_h)
    op_c(x, y_c)
    assert rel_l2(y_h.cpu().numpy(), y_c.cpu().numpy()) <= 1e-14
    mf2 = G.CartesianMesh([n, 2 * n, 2 * n], [h, h / 2, h / 2], [1.0, 0.0, 0.0], False)
    ff2 = G.MakeFiniteElementSpace(mf2, fe, G.L2Restriction(fc.num_dofs))
    op_2 = G.MakeAdvectionFaceOperator(fc, ff2, G.NonconformingCartesianIntermeshFaceConnectivity(3, 3, mc.sizes,
                                                                                                   mf2.sizes), ir, beta)
    N2 = fc.num_dofs + ff2.num_dofs
    x2 = G.fill_random(N2, 32)
    y2 = torch.zeros(N2, dtype=torch.float64, device="cuda")
    op_2(_dev(x2), y2)
    got = y2.cpu().numpy()
    assert abs(got.sum()) <= 1e-13 * np.abs(got).sum()
    ref = np.zeros(N2)
    m, f = (dict(sizes=[n] * 3, h=[h] * 3, origin=[0.0] * 3, n=p + 1, basis=0),
            dict(sizes=[n, 2 * n, 2 * n], h=[h, h / 2, h / 2], origin=[1.0, 0.0, 0.0], n=p + 1, basis=0))
    nc = fc.num_dofs
    oracle.hmesh_operator(3, p + 1, m, f, beta_numpy("c2", 3), x2[:nc], x2[nc:], ref[:nc], ref[nc:])
    assert rel_l2(got, ref) <= F64_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("sizes", [(1, 4, 3), (1, 1, 1), (5, 1, 2)])
def test_face_gpu_degenerate_meshes(sizes):
    # axes with a single cell have no interior faces; a single cell has none at all
    import torch
    from cases import beta_field

    h = [1.0 / s for s in sizes]
    mesh = G.CartesianMesh(list(sizes), h, [0.0] * 3, False)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(2))
    op = G.MakeAdvectionFaceOperator(fes, G.MakeCartesianInteriorFaceConnectivity(mesh.sizes), G.MakeIntegrationRule(3),
                                     beta_field("c2", 3))
    x = G.fill_random(fes.num_dofs, 41)
    y0 = G.fill_random(fes.num_dofs, 42)
    y = _dev(y0)
    op(_dev(x), y)
    ref = np.zeros_like(x)
    oracle.face_operator(list(sizes), h, [0.0] * 3, 3, 0, 3, beta_numpy("c2", 3), x, ref)
    got = y.cpu().numpy() - y0
    if np.abs(ref).max() == 0.0:
        assert np.abs(got).max() == 0.0
    else:
        assert rel_l2(got, ref) <= F64_TOL


@pytest.mark.gpu
@pytest.mark.parametrize(
    "sizes,periodic,coef",
    [((16, 4, 3), False, "c2"), ((32, 2, 2), True, "c2"), ((16, 3, 2), False, "const"), ((48, 1, 3), False, "c2")],
)
def test_face_gpu_streaming_path_matches_face_kernel_and_oracle(sizes, periodic, coef):
    # 3D GLL p=3, q=4, rows of whole 16-element tiles: the interior faces run
    # on the streaming Kronecker kernel (cell terms off, y accumulated); the
    # same operator with variant 0 pinned runs the face kernel
    import torch
    from cases import beta_field

    print(f"face streaming path {sizes} periodic={periodic} {coef}", flush=True)
    h = [1.0 / s for s in sizes]
    mesh = G.CartesianMesh(list(sizes), h, [0.0] * 3, [periodic] * 3)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(3))
    op = G.MakeAdvectionFaceOperator(fes, G.MakeCartesianInteriorFaceConnectivity(fes.mesh.sizes),
                                     G.MakeIntegrationRule(4), beta_field(coef, 3))
    x = G.fill_random(fes.num_dofs, 29)
    y0 = G.fill_random(fes.num_dofs, 30)
    y = _dev(y0)
    op(_dev(x), y)
    op.plan.set_variant(torch.float64, 0)
    yk = _dev(y0)
    op(_dev(x), yk)
    op.plan.set_variant(torch.float64, -1)
    ref = np.zeros_like(x)
    oracle.face_operator(list(sizes), h, [0.0] * 3, 4, 0, 4, beta_numpy(coef, 3), x, ref)
    got = y.cpu().numpy() - y0
    assert rel_l2(got, ref) <= F64_TOL
    assert rel_l2(got, yk.cpu().numpy() - y0) <= F64_TOL
    assert abs(got.sum()) <= 1e-12 * np.abs(got).sum()
