"""The oracle's SIPDG diffusion apply (gdo_diffusion_apply) against GenDiL's own
DiffusionOperator (diffusionoperator.hpp:44-331), through the fixtures that
oracle/_ref/gendil_ref wrote (tests/golden/diff*.npz), plus the properties the
bilinear form fixes: symmetry for sigma = -1 and the constant null space of a
periodic mesh.  Planted mistakes must fail the fixtures."""
import hashlib

import numpy as np
import pytest

import gendil_amd as G
import oracle
from cases import DIFF_BY_NAME, DIFF_CASES, rel_l2


def load(case):
    g = np.load(f"{oracle.HERE.parent}/tests/golden/{case.name}.npz")
    x = G.fill_random(int(g["n"]), case.seed)
    assert hashlib.sha256(x.tobytes()).hexdigest() == str(g["x_sha256"])
    return x, g["y"]


@pytest.mark.parametrize("case", DIFF_CASES, ids=[c.name for c in DIFF_CASES])
def test_oracle_diffusion_matches_gendil(case):
    x, y_ref = load(case)
    fes, ir = case.space()
    y = oracle.Problem(fes, ir).diffusion(x, case.mu(), case.sigma, case.kappa)
    assert rel_l2(y, y_ref) <= 1e-13


# (case, plant): a case where the planted mistake changes the result (a p = 1
# element has the same normal derivative at both ends; sigma = 0 has no
# symmetry term)
PLANTED = [("diff3d_trig_p2q5_per", 3), ("diff3d_trig_p2q5_per", 4), ("diff3d_trig_p3q6_bnd", 3),
           ("diff3d_trig_p3q6_bnd", 4), ("diff2d_gl_trig_p2q4", 4), ("diff1d_gl_poly_p3q5", 3)]


@pytest.mark.parametrize("name,plant", PLANTED)
def test_planted_mistakes_fail_the_fixtures(name, plant):
    case = DIFF_BY_NAME[name]
    x, y_ref = load(case)
    fes, ir = case.space()
    y = oracle.Problem(fes, ir, plant=plant).diffusion(x, case.mu(), case.sigma, case.kappa)
    assert rel_l2(y, y_ref) > 1e-3


def dense(fes, ir, mu, sigma, kappa):
    P = oracle.Problem(fes, ir)
    n = P.num_dofs
    return np.stack([P.diffusion(np.eye(n)[i], mu, sigma, kappa) for i in range(n)], axis=1)


@pytest.mark.parametrize("periodic", [True, False])
def test_symmetric_interior_penalty_is_symmetric(periodic):
    case = DIFF_BY_NAME["diff2d_gl_trig_p2q4"]
    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(2, 3, periodic), G.MakeLegendreFiniteElement(2))
    A = dense(fes, G.MakeIntegrationRule(4), case.mu(), -1.0, 9.0)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    # positive definite with zero Dirichlet data, semidefinite on a torus
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert ev.min() >= -1e-10 * ev.max()
    if not periodic:
        assert ev.min() > 1e-8 * ev.max()


def test_constants_are_the_null_space_on_a_periodic_mesh():
    case = DIFF_BY_NAME["diff3d_trig_p3q6_per"]
    fes, ir = case.space()
    P = oracle.Problem(fes, ir)
    for sigma in (-1.0, 0.0, 1.0):
        y = P.diffusion(np.ones(P.num_dofs), case.mu(), sigma, case.kappa)
        assert np.abs(y).max() <= 1e-12


def test_one_dimensional_p1_stencil():
    """1D, p = 1 GLL, unit diffusivity on a uniform mesh: with sigma = -1 and
    kappa = 2 the SIPDG matrix of one interior node pair is the textbook
    element matrix; checked against a direct assembly of the bilinear form
    a(u, v) = sum_K int u'v' - sum_F ({u'} [v] + {v'} [u]) + kappa/h [u][v]."""
    n, h = 4, 0.25
    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(1, n, True), G.MakeLobattoFiniteElement(1))
    A = dense(fes, G.MakeIntegrationRule(2), 1.0, -1.0, 2.0)
    # local basis on element k: l0 = 1 - xi, l1 = xi; u' = (u1 - u0)/h
    ref = np.zeros((2 * n, 2 * n))
    for k in range(n):
        i = [2 * k, 2 * k + 1]
        ref[np.ix_(i, i)] += np.array([[1.0, -1.0], [-1.0, 1.0]]) / h
    for k in range(n):  # face between element k (right node) and k+1 (left node)
        kp = (k + 1) % n
        um, up = 2 * k + 1, 2 * kp  # trace dofs of the two sides
        dm = {2 * k: -1.0 / h, 2 * k + 1: 1.0 / h}    # u' on the minus element
        dp = {2 * kp: -1.0 / h, 2 * kp + 1: 1.0 / h}  # u' on the plus element
        jump = {um: 1.0, up: -1.0}                    # [u] = u^- - u^+, normal +x
        avg = {}
        for dct in (dm, dp):
            for j, v in dct.items():
                avg[j] = avg.get(j, 0.0) + 0.5 * v
        for a_, va in jump.items():
            for b_, vb in avg.items():
                ref[a_, b_] -= va * vb   # -{u'}[v]
                ref[b_, a_] -= va * vb   # -{v'}[u]
            for b_, vb in jump.items():
                ref[a_, b_] += 2.0 / h * va * vb
    assert np.abs(A - ref).max() <= 1e-12 * np.abs(ref).max()
