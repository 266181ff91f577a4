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
from cases import CASES  # noqa: E402


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


if __name__ == "__main__":
    main()
