"""Regenerate tests/golden/*.npz from GenDiL's own CPU operators.

Runs oracle/_ref/gendil_ref (GenDiL's SerialKernelConfiguration apply built
from the reference headers by oracle/Makefile) for every case of
tests/cases.py and stores, per case, the output vector y = op(x) with
x = FillRandom(seed), plus a SHA-256 of x so the tests can check that
gendil_amd.fill_random reproduces GenDiL's input sequence bit for bit.

Usage (needs the reference tree, i.e. only in the build container):
    make -C oracle && python tests/golden/make_golden.py
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
from cases import CASES, FACE_CASES  # noqa: E402


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/gendil_ref not built (needs the GenDiL sources)")
    with tempfile.TemporaryDirectory() as tmp:
        for c in CASES:
            path = Path(tmp) / f"{c.name}.bin"
            x, y = oracle.ref_fixture(c.op, c.D, c.P, c.Q, c.basis, c.n, c.periodic, c.coef, c.bc, c.seed, path)
            digest = hashlib.sha256(x.tobytes()).hexdigest()
            np.savez_compressed(HERE / f"{c.name}.npz", y=y, x_sha256=np.array(digest), n=np.array(x.size))
            print(f"{c.name}: {x.size} dofs, |y|={np.linalg.norm(y):.6e}")
        # face-centric operators (GenDiL's AdvectionFaceOperator and its
        # nonconforming variant), y accumulated from zero
        for c in FACE_CASES:
            path = Path(tmp) / f"{c.name}.bin"
            x, y = oracle.ref_fixture(c.harness_op, c.D, c.P, c.Q, c.basis, c.n, 0, c.coef, "none", c.seed, path)
            digest = hashlib.sha256(x.tobytes()).hexdigest()
            np.savez_compressed(HERE / f"{c.name}.npz", y=y, x_sha256=np.array(digest), n=np.array(x.size))
            print(f"{c.name}: {x.size} dofs, |y|={np.linalg.norm(y):.6e}")
        # GenDiL's BiCGSTAB_no_alloc on M + 0.05 A (3D, n=4, GLL p=1, 2 points,
        # C2 field, zero inflow), b = FillRandom(7), x0 = 0, tol 1e-10
        b, x, it, ok, rel = oracle.ref_bicgstab(3, 1, 2, "L", 4, "c2", "zero", 7, Path(tmp) / "bicgstab.bin")
        np.savez_compressed(HERE / "bicgstab3d_c2_p1q2_bc0.npz", x=x, iterations=np.array(it), success=np.array(ok),
                            rel_residual=np.array(rel), b_sha256=np.array(hashlib.sha256(b.tobytes()).hexdigest()))
        print(f"bicgstab3d_c2_p1q2_bc0: {b.size} dofs, {it} iterations, success {ok}, rel {rel:.3e}")


if __name__ == "__main__":
    main()
