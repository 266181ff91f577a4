"""Catalogue of parity cases shared by the golden-fixture generator and tests.

Each case mirrors one invocation of oracle/_ref/gendil_ref (GenDiL's own
operators): same mesh, element, rule and coefficient, described here with the
fields the C-ABI accepts.  Coefficient names match ref_harness.cpp.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

import gendil_amd as G
from gendil_amd import workloads as W
from gendil_amd.fields import SeparableField, VectorField

TWO_PI = 2.0 * math.pi


@dataclass(frozen=True)
class Case:
    name: str
    op: str  # mass | adv
    D: int
    P: int
    Q: int
    basis: str  # L | G
    n: int
    periodic: bool
    coef: str
    bc: str  # none | zero | poly
    seed: int = 12345

    def space(self):
        if self.op == "adv6dprod":
            # the mesh of GenDiL's 6D benchmark (benchmark-advection-6d.cpp:14-38)
            mesh = G.MakeCartesianProductMesh(
                G.Cartesian3DMesh(1.0, self.n, self.n, self.n), G.Cartesian2DMesh(1.2345, self.n, self.n),
                G.Cartesian1DMesh(1.0, self.n),
            )
        else:
            mesh = G.CartesianMesh.unit(self.D, self.n, self.periodic)
        fe = G.MakeLobattoFiniteElement(self.P) if self.basis == "L" else G.MakeLegendreFiniteElement(self.P)
        return G.MakeFiniteElementSpace(mesh, fe), G.MakeIntegrationRule(self.Q)


def sigma_field(kind: str, D: int):
    if kind == "one":
        return SeparableField.constant(1.0)
    if kind == "poly":  # 1 + x y + z^2  (1 + x^2 in 1D)
        if D >= 2:
            return SeparableField(
                [(1.0, {}), (1.0, {0: lambda s: s, 1: lambda s: s}), (1.0, {D - 1: lambda s: s * s})]
            )
        return SeparableField([(1.0, {}), (1.0, {0: lambda s: s * s})])
    raise ValueError(kind)


def beta_field(kind: str, D: int):
    if kind == "c2":
        assert D == 3
        return W.advection_c2()
    if kind == "bubble":
        return VectorField([SeparableField([(1.0, {d: lambda s: -s * (1.0 - s)})]) for d in range(D)])
    if kind == "const":
        return VectorField([SeparableField.constant((-0.5) ** d) for d in range(D)])
    if kind == "trig":
        return VectorField(
            [SeparableField([(1.0, {(d + 1) % D: lambda s: 0.5 + np.sin(TWO_PI * s)})]) for d in range(D)]
        )
    if kind == "phase":
        assert D == 6
        return W.phase_space_6d_separable()
    if kind == "ones":
        return VectorField([SeparableField.constant(1.0) for _ in range(D)])
    if kind == "yzx":  # test-advection-p-adaptivity.cpp:62-71
        assert D == 3
        return G.AffineField(np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]]), np.zeros(3))
    raise ValueError(kind)


def beta_numpy(kind: str, D: int):
    """The same fields as functions of points [..., D] (for the numpy face oracle)."""
    if kind == "c2":
        return lambda X: np.stack([1 + 0.5 * np.sin(TWO_PI * X[..., 1]), 0.5 + 0.5 * np.cos(TWO_PI * X[..., 2]),
                                   0.25 + 0.25 * np.sin(TWO_PI * X[..., 0])], -1)
    if kind == "trig":
        return lambda X: np.stack([0.5 + np.sin(TWO_PI * X[..., (d + 1) % D]) for d in range(D)], -1)
    if kind == "const":
        return lambda X: np.stack([np.full(X.shape[:-1], (-0.5) ** d) for d in range(D)], -1)
    if kind == "phase":
        return lambda X: np.stack([X[..., 3], X[..., 4], X[..., 5], np.full(X.shape[:-1], 0.1),
                                   np.zeros(X.shape[:-1]), np.zeros(X.shape[:-1])], -1)
    if kind == "yzx":
        return lambda X: np.stack([X[..., 1], X[..., 2], X[..., 0]], -1)
    raise ValueError(kind)


def bc_field(kind: str, D: int):
    if kind == "none":
        return None
    if kind == "zero":
        return SeparableField.constant(0.0)
    if kind == "poly":  # 1 + x - 0.5 y  (1 + x in 1D)
        if D >= 2:
            return SeparableField([(1.0, {}), (1.0, {0: lambda s: s}), (-0.5, {1: lambda s: s})])
        return SeparableField([(1.0, {}), (1.0, {0: lambda s: s})])
    raise ValueError(kind)


CASES = [
    # mass: test-mass-3d-generic.cpp pattern (n=6, GLL p, q=p+3, sigma = 1+xy+z^2)
    Case("mass3d_p1q3_poly", "mass", 3, 1, 3, "L", 6, False, "poly", "none"),
    Case("mass3d_p2q5_poly", "mass", 3, 2, 5, "L", 6, False, "poly", "none"),
    Case("mass3d_p3q6_poly", "mass", 3, 3, 6, "L", 6, False, "poly", "none"),
    Case("mass3d_p3q4_one", "mass", 3, 3, 4, "L", 4, False, "one", "none"),
    Case("mass3d_gl_p3q4_poly", "mass", 3, 3, 4, "G", 4, False, "poly", "none"),
    Case("mass2d_p2q3_poly", "mass", 2, 2, 3, "L", 5, False, "poly", "none"),
    Case("mass1d_p3q4_poly", "mass", 1, 3, 4, "L", 7, False, "poly", "none"),
    Case("mass4d_p1q2_poly", "mass", 4, 1, 2, "L", 3, False, "poly", "none"),
    Case("mass6d_p1q2_one", "mass", 6, 1, 2, "L", 2, False, "one", "none"),
    # advection: the benchmark field on periodic / inflow / skipped boundaries
    Case("adv3d_c2_p3q4_per", "adv", 3, 3, 4, "L", 4, True, "c2", "none"),
    Case("adv3d_c2_p3q4_bc0", "adv", 3, 3, 4, "L", 4, False, "c2", "zero"),
    Case("adv3d_c2_p3q4_nobc", "adv", 3, 3, 4, "L", 4, False, "c2", "none"),
    Case("adv3d_c2_p4q5_per", "adv", 3, 4, 5, "L", 3, True, "c2", "none"),
    Case("adv3d_trig_p3q4_bcpoly", "adv", 3, 3, 4, "L", 3, False, "trig", "poly"),
    Case("adv3d_trig_gl_p3q4_bcpoly", "adv", 3, 3, 4, "G", 3, False, "trig", "poly"),
    Case("adv3d_const_p2q4_per", "adv", 3, 2, 4, "L", 4, True, "const", "none"),
    Case("adv3d_trig_p2q3_m1per", "adv", 3, 2, 3, "L", 1, True, "trig", "none"),
    # test-advection-3d-generic.cpp pattern (n=2, GLL p=0..3, q=p+2, bubble field, g=0)
    Case("adv3d_bubble_p0q2", "adv", 3, 0, 2, "L", 2, False, "bubble", "zero"),
    Case("adv3d_bubble_p1q3", "adv", 3, 1, 3, "L", 2, False, "bubble", "zero"),
    Case("adv3d_bubble_p2q4", "adv", 3, 2, 4, "L", 2, False, "bubble", "zero"),
    Case("adv3d_bubble_p3q5", "adv", 3, 3, 5, "L", 2, False, "bubble", "zero"),
    Case("adv2d_trig_p2q3_bcpoly", "adv", 2, 2, 3, "L", 5, False, "trig", "poly"),
    Case("adv2d_trig_p3q4_per", "adv", 2, 3, 4, "L", 4, True, "trig", "none"),
    Case("adv1d_const_p3q4_bcpoly", "adv", 1, 3, 4, "L", 7, False, "const", "poly"),
    Case("adv4d_trig_p1q2_per", "adv", 4, 1, 2, "L", 3, True, "trig", "none"),
    Case("adv4d_trig_p2q3_bcpoly", "adv", 4, 2, 3, "L", 3, False, "trig", "poly"),
    Case("adv6d_phase_p1q2_per", "adv", 6, 1, 2, "L", 2, True, "phase", "none"),
    Case("adv6d_phase_p2q3_bc0", "adv", 6, 2, 3, "L", 2, False, "phase", "zero"),
    # GenDiL's 6D benchmark operator: 3D x 2D x 1D product mesh (h 1, 1.2345, 1),
    # Legendre p=1, 3 points, beta = 1, no boundary data
    Case("adv6dprod_ones_p1q3_gl", "adv6dprod", 6, 1, 3, "G", 2, False, "ones", "none"),
    # 5D, among them the (order, points) pairs of GenDiL's 5D range benchmarks
    # (Legendre, beta = 1, boundary faces skipped; advection q = p + 1, mass q = p + 2)
    Case("adv5d_ones_gl_p1q2", "adv", 5, 1, 2, "G", 2, False, "ones", "none"),
    Case("adv5d_trig_p2q3_bcpoly", "adv", 5, 2, 3, "L", 2, False, "trig", "poly"),
    Case("adv5d_const_p3q4_per", "adv", 5, 3, 4, "L", 2, True, "const", "none"),
    Case("mass5d_gl_p2q4_poly", "mass", 5, 2, 4, "G", 2, False, "poly", "none"),
    # the highest (order, points) pairs of GenDiL's 3D, 4D and 6D range benchmarks
    Case("adv3d_ones_gl_p5q7", "adv", 3, 5, 7, "G", 2, False, "ones", "none"),
    Case("mass3d_gl_p6q8_poly", "mass", 3, 6, 8, "G", 2, False, "poly", "none"),
    Case("adv4d_ones_gl_p4q5", "adv", 4, 4, 5, "G", 2, False, "ones", "none"),
    Case("mass4d_gl_p4q6_poly", "mass", 4, 4, 6, "G", 2, False, "poly", "none"),
    Case("adv6d_ones_gl_p2q3", "adv", 6, 2, 3, "G", 2, False, "ones", "none"),
    Case("mass6d_gl_p2q4_one", "mass", 6, 2, 4, "G", 2, False, "one", "none"),
]

BY_NAME = {c.name: c for c in CASES}


@dataclass(frozen=True)
class FaceCase:
    """A face-centric operator of ref_harness.cpp: `interior` = advface on the
    unit n^D mesh (non-periodic); `intermesh` = face F of an n^3 GLL mesh of
    order P at the origin glued to an n^3 mesh of order PP shifted by one unit
    along the face normal, q shared, vector [minus | plus]."""

    name: str
    kind: str  # interior | intermesh | hmesh
    D: int
    P: int
    Q: int
    basis: str
    n: int
    coef: str
    PP: int = 0
    face: int = 0
    seed: int = 12345
    R: int = 1  # hmesh: fine cells per coarse cell along the tangential axes

    @property
    def harness_op(self):
        if self.kind == "hmesh":
            return f"hmesh_{self.face}_{self.R}"
        return "advface" if self.kind == "interior" else f"intermesh_{self.face}_{self.PP}"

    def sides(self):
        """(minus, plus) oracle side descriptions."""
        h = 1.0 / self.n
        b = 0 if self.basis == "L" else 1
        m = dict(sizes=[self.n] * self.D, h=[h] * self.D, origin=[0.0] * self.D, n=self.P + 1, basis=b)
        if self.kind == "interior":
            return m, m
        if self.kind == "hmesh":
            ax = self.face % self.D
            fo = [0.0] * self.D
            fo[ax] = 1.0 if self.face >= self.D else -1.0
            return m, dict(sizes=[self.n if d == ax else self.R * self.n for d in range(self.D)],
                           h=[h if d == ax else h / self.R for d in range(self.D)], origin=fo, n=self.P + 1, basis=b)
        po = [0.0] * self.D
        po[self.face % self.D] = 1.0 if self.face >= self.D else -1.0
        return m, dict(sizes=[self.n] * self.D, h=[h] * self.D, origin=po, n=self.PP + 1, basis=b)

    def oracle(self, x):
        import oracle

        beta = beta_numpy(self.coef, self.D)
        m, p = self.sides()
        y = np.zeros_like(x)
        if self.kind == "interior":
            oracle.face_operator(m["sizes"], m["h"], m["origin"], m["n"], m["basis"], self.Q, beta, x, y)
        else:
            nm = self.n ** self.D * m["n"] ** self.D
            fn = oracle.hmesh_operator if self.kind == "hmesh" else oracle.intermesh_operator
            fn(self.face, self.Q, m, p, beta, x[:nm], x[nm:], y[:nm], y[nm:])
        return y

    def gpu_operator(self):
        """The operator built through gendil_amd's GenDiL-shaped factories."""
        fe = G.MakeLobattoFiniteElement if self.basis == "L" else G.MakeLegendreFiniteElement
        ir = G.MakeIntegrationRule(self.Q)
        beta = beta_field(self.coef, self.D)
        if self.kind == "interior":
            fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(self.D, self.n, False), fe(self.P))
            return G.MakeAdvectionFaceOperator(fes, G.MakeCartesianInteriorFaceConnectivity(fes.mesh.sizes), ir, beta)
        m, p = self.sides()
        if self.kind == "hmesh":
            mesh_c = G.CartesianMesh(m["sizes"], m["h"], m["origin"], False)
            mesh_f = G.CartesianMesh(p["sizes"], p["h"], p["origin"], False)
            fc = G.MakeFiniteElementSpace(mesh_c, fe(self.P))
            ff = G.MakeFiniteElementSpace(mesh_f, fe(self.P), G.L2Restriction(fc.num_dofs))
            conn = G.NonconformingCartesianIntermeshFaceConnectivity(self.D, self.face, mesh_c.sizes, mesh_f.sizes)
            return G.MakeAdvectionFaceOperator(fc, ff, conn, ir, beta)
        mesh_m = G.Cartesian3DMesh(1.0 / self.n, self.n, self.n, self.n, origin=tuple(m["origin"]))
        mesh_p = G.Cartesian3DMesh(1.0 / self.n, self.n, self.n, self.n, origin=tuple(p["origin"]))
        fm = G.MakeFiniteElementSpace(mesh_m, fe(self.P))
        fp = G.MakeFiniteElementSpace(mesh_p, fe(self.PP), G.L2Restriction(fm.num_dofs))
        conn = G.CartesianIntermeshFaceConnectivity(3, self.face, mesh_m.sizes, mesh_p.sizes)
        return G.MakeAdvectionFaceOperator(fm, fp, conn, ir, beta)


FACE_CASES = [
    FaceCase("face3d_c2_p3q4", "interior", 3, 3, 4, "L", 4, "c2"),
    FaceCase("face3d_trig_gl_p2q3", "interior", 3, 2, 3, "G", 3, "trig"),
    FaceCase("face2d_trig_p2q3", "interior", 2, 2, 3, "L", 4, "trig"),
    FaceCase("face6d_phase_p1q2", "interior", 6, 1, 2, "L", 2, "phase"),
    # test-advection-p-adaptivity.cpp: +x face, p = 1 | p = 3, q = 5, beta = (y, z, x)
    FaceCase("imesh3d_yzx_f3_p1p3q5", "intermesh", 3, 1, 5, "L", 3, "yzx", PP=3, face=3),
    FaceCase("imesh3d_yzx_f0_p3p1q5", "intermesh", 3, 3, 5, "L", 3, "yzx", PP=1, face=0),
    # test-face-advection-split-mesh.cpp: -x face, p = 2 both sides, q = 4
    FaceCase("imesh3d_const_f0_p2p2q4", "intermesh", 3, 2, 4, "L", 3, "const", PP=2, face=0),
    FaceCase("imesh3d_c2_f4_p2p3q4", "intermesh", 3, 2, 4, "L", 3, "c2", PP=3, face=4),
    FaceCase("imesh3d_c2_f2_p3p2q4", "intermesh", 3, 3, 4, "L", 3, "c2", PP=2, face=2),
    # h-nonconforming interfaces: test-advection-h-adaptivity-2d.cpp (2D, +x face,
    # p = 2, q = 4, twice the cells along y on the fine side), then 3D ones
    FaceCase("hmesh2d_const_f2_p2q4_r2", "hmesh", 2, 2, 4, "L", 2, "const", face=2, R=2),
    FaceCase("hmesh3d_c2_f3_p2q3_r2", "hmesh", 3, 2, 3, "L", 2, "c2", face=3, R=2),
    FaceCase("hmesh3d_c2_f0_p1q2_r2", "hmesh", 3, 1, 2, "L", 2, "c2", face=0, R=2),
    FaceCase("hmesh3d_trig_f4_p3q4_r3", "hmesh", 3, 3, 4, "L", 2, "trig", face=4, R=3),
]

FACE_BY_NAME = {c.name: c for c in FACE_CASES}


def oracle_problem(case: Case, **kw):
    import oracle

    fes, ir = case.space()
    if case.op == "mass":
        return oracle.Problem(fes, ir, mass_coef=1.0, sigma=sigma_field(case.coef, case.D), **kw)
    return oracle.Problem(
        fes, ir, adv_coef=1.0, adv=beta_field(case.coef, case.D), bc=bc_field(case.bc, case.D), **kw
    )


def gpu_operator(case: Case):
    fes, ir = case.space()
    if case.op == "mass":
        return G.MakeMassFiniteElementOperator(fes, ir, sigma_field(case.coef, case.D))
    return G.MakeAdvectionOperator(fes, ir, beta_field(case.coef, case.D), bc_field(case.bc, case.D))


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return float(d / nb) if nb > 0 else float(d)


def padapt_field(kind: str, D: int):
    """Fields of the p-adaptive composites in the descriptor forms the
    p-adaptive kernels take (affine or constant components)."""
    if kind == "phase":
        assert D == 6
        return W.phase_space_6d_affine()
    if kind == "yzx":
        return beta_field("yzx", D)
    if kind == "const":
        return beta_field("const", D)
    raise ValueError(kind)


@dataclass(frozen=True)
class PadaptCase:
    """A p-adaptive block composite of oracle/ref_padapt.cpp (GenDiL's own
    p-adaptivity: uniform-order blocks, each with its element-centric
    operator on q = p + 1 points, coupled by the intermesh face operator on
    q = max + 1 points).  The minus block (order PM) has n cells per axis at
    the origin; the plus block (order PP) has n + 1 cells along the interface
    axis, beyond face `face` of the minus block; h = 1/n; periodic along the
    tangential axes when `periodic`.  GenDiL's vector is [minus | plus].

    The same operator is one Cartesian mesh over both blocks with per-element
    orders, which `layout()` builds together with the permutation between
    the two DoF orderings."""

    name: str
    D: int
    PM: int
    PP: int
    face: int
    n: int
    periodic: bool
    coef: str
    seed: int = 12345

    @property
    def harness_op(self):
        return f"padapt_{self.face}_{self.PP}"

    @property
    def axis(self):
        return self.face % self.D

    def layout(self):
        """(space, perm): the per-element-order space over the union of the
        blocks, and perm with ours[perm[i]] = GenDiL's entry i."""
        D, n, ax = self.D, self.n, self.axis
        h = 1.0 / n
        plus_first = self.face < D
        gsizes = [2 * n + 1 if d == ax else n for d in range(D)]
        origin = [0.0] * D
        if plus_first:
            origin[ax] = -(n + 1) * h
        periodic = [bool(self.periodic) and d != ax for d in range(D)]
        mesh = G.CartesianMesh(gsizes, [h] * D, origin, periodic)
        sub = np.array(np.unravel_index(np.arange(mesh.num_cells), gsizes[::-1]))[::-1]  # x fastest
        in_minus = sub[ax] >= n + 1 if plus_first else sub[ax] < n
        orders = np.where(in_minus, self.PM, self.PP).astype(np.int8)
        fes = G.MakePAdaptiveSpace(mesh, orders, orders + 1)
        offs = fes.dof_offsets()
        strides = np.cumprod([1] + gsizes[:-1])
        parts = []
        blocks = (
            ([n] * D, (n + 1) if plus_first else 0, self.PM),
            ([n + 1 if d == ax else n for d in range(D)], 0 if plus_first else n, self.PP),
        )
        for bsizes, shift, p in blocks:
            nb = int(np.prod(bsizes))
            bsub = np.array(np.unravel_index(np.arange(nb), bsizes[::-1]))[::-1]
            bsub[ax] += shift
            g = (bsub * strides[:, None]).sum(0)
            nd = (p + 1) ** D
            parts.append((offs[g][:, None] + np.arange(nd)[None, :]).reshape(-1))
        return fes, np.concatenate(parts)

    def field(self):
        return padapt_field(self.coef, self.D)


PADAPT_CASES = [
    # +x interface, every ordered pair of orders 1..3, beta = (y, z, x) (the
    # p-adaptivity test's field), tangential axes periodic
    *[PadaptCase(f"padapt3d_yzx_f3_p{a}p{b}", 3, a, b, 3, 3, True, "yzx")
      for a, b in ((1, 2), (2, 1), (1, 3), (3, 1), (2, 3), (3, 2))],
    # -y interface, every ordered pair, no periodic axes
    *[PadaptCase(f"padapt3d_yzx_f1_p{a}p{b}", 3, a, b, 1, 3, False, "yzx")
      for a, b in ((1, 2), (2, 1), (1, 3), (3, 1), (2, 3), (3, 2))],
    # the other orientation of each axis, and the z axis; a constant field of
    # mixed signs
    PadaptCase("padapt3d_yzx_f0_p1p3", 3, 1, 3, 0, 3, True, "yzx"),
    PadaptCase("padapt3d_const_f0_p3p2", 3, 3, 2, 0, 3, False, "const"),
    PadaptCase("padapt3d_const_f4_p2p1", 3, 2, 1, 4, 3, True, "const"),
    PadaptCase("padapt3d_yzx_f4_p3p1", 3, 3, 1, 4, 3, False, "yzx"),
    PadaptCase("padapt3d_yzx_f5_p2p3", 3, 2, 3, 5, 3, True, "yzx"),
    PadaptCase("padapt3d_const_f2_p1p2", 3, 1, 2, 2, 3, True, "const"),
    # phase space (beta = (X3, X4, X5, 0.1, 0, 0)): split along a velocity
    # axis and along a space axis
    PadaptCase("padapt6d_phase_f9_p1p2", 6, 1, 2, 9, 2, True, "phase"),
    PadaptCase("padapt6d_phase_f0_p2p1", 6, 2, 1, 0, 2, True, "phase"),
]

PADAPT_BY_NAME = {c.name: c for c in PADAPT_CASES}


def diffusivity_field(kind: str, D: int):
    """The diffusivity fields of oracle/ref_diffusion.cpp as separable fields."""
    if kind == "one":
        return SeparableField.constant(1.0)
    if kind == "poly":  # 1 + x_0 x_{D-1} + 0.5 x_0^2
        if D == 1:
            return SeparableField([(1.0, {}), (1.5, {0: lambda s: s * s})])
        return SeparableField(
            [(1.0, {}), (1.0, {0: lambda s: s, D - 1: lambda s: s}), (0.5, {0: lambda s: s * s})]
        )
    if kind == "trig":
        if D == 3:
            return W.diffusivity_trig_3d()
        last = lambda s: np.cos(TWO_PI * (2.0 * s + 0.13))  # noqa: E731
        if D == 1:
            return SeparableField([(2.0, {}), (0.25, {0: lambda s: np.sin(TWO_PI * s)}), (0.17, {0: last})])
        return SeparableField([(2.0, {}), (0.25, {0: lambda s: np.sin(TWO_PI * s)}), (0.17, {D - 1: last})])
    raise ValueError(kind)


@dataclass(frozen=True)
class DiffCase:
    """One SIPDG diffusion fixture: GenDiL's DiffusionOperator on an n^D mesh
    (oracle/ref_diffusion.cpp), sigma the symmetry parameter, kappa the penalty."""
    name: str
    D: int
    P: int
    Q: int
    basis: str
    n: int
    periodic: bool
    coef: str
    sigma: float
    kappa: float
    seed: int = 12345

    @property
    def harness_bc(self):
        return f"s{self.sigma:g}_k{self.kappa:g}"

    def space(self):
        mesh = G.CartesianMesh.unit(self.D, self.n, self.periodic)
        fe = G.MakeLobattoFiniteElement(self.P) if self.basis == "L" else G.MakeLegendreFiniteElement(self.P)
        return G.MakeFiniteElementSpace(mesh, fe), G.MakeIntegrationRule(self.Q)

    def mu(self):
        return diffusivity_field(self.coef, self.D)


DIFF_CASES = [
    # test-diffusion-3d-generic.cpp:78-167: n = 2 periodic, GLL, q = p + 3,
    # sigma = -1, kappa = (p + 1)^2
    DiffCase("diff3d_trig_p1q4_per", 3, 1, 4, "L", 2, True, "trig", -1.0, 4.0),
    DiffCase("diff3d_trig_p2q5_per", 3, 2, 5, "L", 2, True, "trig", -1.0, 9.0),
    DiffCase("diff3d_trig_p3q6_per", 3, 3, 6, "L", 2, True, "trig", -1.0, 16.0),
    # boundary faces (weak homogeneous Dirichlet), the other two symmetry parameters
    DiffCase("diff3d_trig_p3q6_bnd", 3, 3, 6, "L", 3, False, "trig", -1.0, 16.0),
    DiffCase("diff3d_poly_p3q4_nip", 3, 3, 4, "L", 3, False, "poly", 1.0, 10.0),
    DiffCase("diff3d_one_p3q5_iip", 3, 3, 5, "L", 4, True, "one", 0.0, 16.0),
    DiffCase("diff3d_gl_poly_p2q4", 3, 2, 4, "G", 3, False, "poly", -1.0, 9.0),
    # test-diffusion-symmetry-1d-minimal.cpp:10-40 (sigma = 1, kappa = 0)
    DiffCase("diff1d_one_p1q4_per", 1, 1, 4, "L", 2, True, "one", 1.0, 0.0),
    # poisson-convergence-{1d,2d,6d}.cpp: Legendre, q = p + 2, sigma = -1, kappa = (p + 1)^2
    DiffCase("diff1d_gl_poly_p3q5", 1, 3, 5, "G", 5, False, "poly", -1.0, 16.0),
    DiffCase("diff2d_gl_trig_p2q4", 2, 2, 4, "G", 4, False, "trig", -1.0, 9.0),
    DiffCase("diff2d_trig_p3q4_per", 2, 3, 4, "L", 4, True, "trig", -1.0, 16.0),
    DiffCase("diff6d_gl_one_p1q3", 6, 1, 3, "G", 2, False, "one", -1.0, 4.0),
]

DIFF_BY_NAME = {c.name: c for c in DIFF_CASES}
