"""Regenerate tests/golden/*.npz from GenDiL's own CPU operators.

Runs oracle/_ref/gendil_ref (GenDiL's SerialKernelConfiguration apply built
from the reference headers by oracle/Makefile) for every case of
tests/cases.py and stores, per case, the output vector y = op(x) with
x = FillRandom(seed), plus a SHA-256 of x so the tests can check that
gendil_amd.fill_random reproduces GenDiL's input sequence bit for bit.

Usage (needs the reference tree, i.e. only in the build container):
    make -C oracle && python tests/golden/make_golden.py [name-prefix ...]
With prefixes, only the fixtures whose names start with one of them are
written (GenDiL's face operators accumulate from OpenMP threads, so
rewriting the others would only add rounding-level churn).
"""
from __future__ import annotations

import hashlib
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle  # noqa: E402
from cases import CASES, DIFF_CASES, FACE_CASES, PADAPT_CASES  # noqa: E402


def main(only=None):
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/gendil_ref not built (needs the GenDiL sources)")
    with tempfile.TemporaryDirectory() as tmp:
        def wanted(name):
            return not only or any(name.startswith(p) for p in only)

        for c in CASES:
            if not wanted(c.name):
                continue
            path = Path(tmp) / f"{c.name}.bin"
            x, y = oracle.ref_fixture(c.op, c.D, c.P, c.Q, c.basis, c.n, c.periodic, c.coef, c.bc, c.seed, path)
            digest = hashlib.sha256(x.tobytes()).hexdigest()
            np.savez_compressed(HERE / f"{c.name}.npz", y=y, x_sha256=np.array(digest), n=np.array(x.size))
            print(f"{c.name}: {x.size} dofs, |y|={np.linalg.norm(y):.6e}")
        # face-centric operators (GenDiL's AdvectionFaceOperator and its
        # nonconforming variant), y accumulated from zero
        for c in FACE_CASES:
            if not wanted(c.name):
                continue
            path = Path(tmp) / f"{c.name}.bin"
            x, y = oracle.ref_fixture(c.harness_op, c.D, c.P, c.Q, c.basis, c.n, 0, c.coef, "none", c.seed, path)
            digest = hashlib.sha256(x.tobytes()).hexdigest()
            np.savez_compressed(HERE / f"{c.name}.npz", y=y, x_sha256=np.array(digest), n=np.array(x.size))
            print(f"{c.name}: {x.size} dofs, |y|={np.linalg.norm(y):.6e}")
        # GenDiL's p-adaptivity: uniform-order blocks with their own operators
        # coupled by the intermesh face operator (ref_padapt.cpp)
        for c in PADAPT_CASES:
            if not wanted(c.name):
                continue
            path = Path(tmp) / f"{c.name}.bin"
            x, y = oracle.ref_fixture(c.harness_op, c.D, c.PM, 0, "L", c.n, c.periodic, c.coef, "none", c.seed, path)
            digest = hashlib.sha256(x.tobytes()).hexdigest()
            np.savez_compressed(HERE / f"{c.name}.npz", y=y, x_sha256=np.array(digest), n=np.array(x.size))
            print(f"{c.name}: {x.size} dofs, |y|={np.linalg.norm(y):.6e}")
        # GenDiL's SIPDG DiffusionOperator (ref_diffusion.cpp)
        for c in DIFF_CASES:
            if not wanted(c.name):
                continue
            path = Path(tmp) / f"{c.name}.bin"
            x, y = oracle.ref_fixture("diff", c.D, c.P, c.Q, c.basis, c.n, int(c.periodic), c.coef, c.harness_bc,
                                      c.seed, path)
            digest = hashlib.sha256(x.tobytes()).hexdigest()
            np.savez_compressed(HERE / f"{c.name}.npz", y=y, x_sha256=np.array(digest), n=np.array(x.size))
            print(f"{c.name}: {x.size} dofs, |y|={np.linalg.norm(y):.6e}")
        # GenDiL's own solvers: BiCGSTAB_no_alloc and GMRES_no_alloc(30) on
        # M + 0.05 A (3D, n=4, GLL p=1, 2 points, C2 field, zero inflow),
        # b = FillRandom(7); ConjugateGradient on M(1 + xy + z^2) (3D, n=5,
        # GLL p=2, 3 points), b = FillRandom(5); x0 = 0, tol 1e-10
        for name, args in (("bicgstab3d_c2_p1q2_bc0", ("bicgstab", 3, 1, 2, "L", 4, "c2", "zero", 7)),
                           ("gmres3d_c2_p1q2_bc0", ("gmres", 3, 1, 2, "L", 4, "c2", "zero", 7)),
                           ("cg3d_mass_p2q3_poly", ("cg", 3, 2, 3, "L", 5, "poly", "none", 5))):
            if not wanted(name):
                continue
            b, x, it, ok, rel = oracle.ref_solve(*args, Path(tmp) / f"{name}.bin")
            np.savez_compressed(HERE / f"{name}.npz", x=x, iterations=np.array(it), success=np.array(ok),
                                rel_residual=np.array(rel), b_sha256=np.array(hashlib.sha256(b.tobytes()).hexdigest()))
            print(f"{name}: {b.size} dofs, {it} iterations, success {ok}, rel {rel:.3e}")


if __name__ == "__main__":
    main(sys.argv[1:])
