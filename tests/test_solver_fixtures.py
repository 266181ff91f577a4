"""GenDiL's own Krylov solvers run through oracle/_ref/gendil_ref
(BiCGSTAB_no_alloc, GMRES_no_alloc, ConjugateGradient; tests/golden/
{bicgstab,gmres,cg}3d_*.npz) against the oracle's operator: the stored
iterate solves the system to the solver's tolerance, the right-hand side is
GenDiL's FillRandom sequence, and the reported counts are those of converged
runs.  The GPU solvers are compared with the same fixtures in
test_gpu_solvers.py."""
import hashlib

import numpy as np
import pytest

import gendil_amd as G
import oracle
from cases import sigma_field
from gendil_amd import workloads as W

SOLVES = [
    # name, system, n, p, q, seed
    ("bicgstab3d_c2_p1q2_bc0", "madv", 4, 1, 2, 7),
    ("gmres3d_c2_p1q2_bc0", "madv", 4, 1, 2, 7),
    ("cg3d_mass_p2q3_poly", "mass", 5, 2, 3, 5),
]


def system(kind, n, p, q):
    fes = G.MakeFiniteElementSpace(G.CartesianMesh.unit(3, n, False), G.MakeLobattoFiniteElement(p))
    ir = G.MakeIntegrationRule(q)
    if kind == "madv":
        return oracle.Problem(fes, ir, mass_coef=1.0, adv_coef=0.05, adv=W.advection_c2(), bc=0.0)
    return oracle.Problem(fes, ir, mass_coef=1.0, sigma=sigma_field("poly", 3))


@pytest.mark.parametrize("name,kind,n,p,q,seed", SOLVES, ids=[s[0] for s in SOLVES])
def test_gendil_solver_fixture_solves_the_system(name, kind, n, p, q, seed):
    g = np.load(f"{oracle.HERE.parent}/tests/golden/{name}.npz")
    P = system(kind, n, p, q)
    b = G.fill_random(P.num_dofs, seed)
    assert hashlib.sha256(b.tobytes()).hexdigest() == str(g["b_sha256"])
    assert bool(g["success"]) and 0 < int(g["iterations"]) < 500
    r = b - P.apply(g["x"])
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-9
