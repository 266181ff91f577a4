"""The p-adaptive oracle against GenDiL's own p-adaptivity.

GenDiL has no per-element order: its p-adaptive discretisation is blocks of
uniform compile-time order in one vector (L2Restriction shifts), each block
applied by its own element-centric advection operator, and the interfaces by
the intermesh face operator on one shared rule (the pattern of
tests/integration/gendil/test-advection-p-adaptivity.cpp:13-100;
advectionoperator.hpp:596-676, 1210-1396).  oracle/ref_padapt.cpp runs that
composite through GenDiL's own operators; tests/golden/padapt*.npz hold its
outputs.  Here the oracle's per-element-order apply, on one mesh over the
blocks with the face rule q = max(q-, q+), is compared with them after the
permutation between the two DoF orderings, and two planted mistakes in the
oracle (the smaller face rule, which conservation cannot see; the neighbour
trace with two tangential axes exchanged) are shown to fail them.
"""
import hashlib

import numpy as np
import pytest

import gendil_amd as G
import oracle
from cases import PADAPT_CASES, rel_l2


def golden(case):
    return np.load(f"{oracle.HERE.parent}/tests/golden/{case.name}.npz")


def oracle_apply(case, x_gendil, plant=0):
    fes, perm = case.layout()
    x = np.empty_like(x_gendil)
    x[perm] = x_gendil
    P = oracle.Problem(fes, None, adv_coef=1.0, adv=case.field(), plant=plant)
    return P.apply(x)[perm]


@pytest.mark.parametrize("case", PADAPT_CASES, ids=lambda c: c.name)
def test_padapt_oracle_matches_gendil_blocks(case):
    g = golden(case)
    x = G.fill_random(int(g["n"]), case.seed)
    assert hashlib.sha256(x.tobytes()).hexdigest() == str(g["x_sha256"])
    fes, perm = case.layout()
    assert fes.num_dofs == x.size and np.array_equal(np.sort(perm), np.arange(x.size))
    err = rel_l2(oracle_apply(case, x), g["y"])
    assert err <= 1e-13, err


@pytest.mark.parametrize("plant", [1, 2])
def test_padapt_fixtures_catch_planted_face_mistakes(plant):
    # mistakes on the mixed-order faces that are the same from both sides:
    # each keeps the sum of A u at zero, and the fixtures still reject it
    worst = 0.0
    for case in PADAPT_CASES:
        if case.D != 3:
            continue
        g = golden(case)
        x = G.fill_random(int(g["n"]), case.seed)
        y = oracle_apply(case, x, plant=plant)
        worst = max(worst, rel_l2(y, g["y"]))
    assert worst > 1e-6, worst


def test_padapt_planted_rule_mistake_keeps_conservation():
    # the smaller face rule is the same from both sides of a face, so the sum
    # of A u stays zero: conservation, which pinned the mixed-order faces
    # before the block fixtures existed, cannot see it; the fixtures can
    mesh = G.CartesianMesh.unit(3, 4, True)
    rng = np.random.default_rng(11)
    orders = rng.integers(1, 4, mesh.num_cells)
    pfes = G.MakePAdaptiveSpace(mesh, orders, orders + 1)
    A = np.array([[0, 1.0, 0], [0, 0, 1.0], [1.0, 0, 0]])
    P = oracle.Problem(pfes, None, adv_coef=1.0, adv=G.AffineField(A, np.array([0.3, -0.2, 0.1])), plant=1)
    y = P.apply(rng.uniform(-1, 1, P.num_dofs))
    assert abs(y.sum()) <= 1e-13 * np.abs(y).sum()
