"""Named coefficient fields and synthetic inputs of the benchmark workloads.

  advection_c2(mesh):  beta(x) = (1 + 0.5 sin 2pi y, 0.5 + 0.5 cos 2pi z,
                       0.25 + 0.25 sin 2pi x) -- each component depends on one
                       coordinate, so it is a separable field (no HBM traffic).
  phase_space_6d():    beta = (X3, X4, X5, 0.1, 0, 0) on 3D space x 3D velocity.
  gaussian_bump(...):  exp(-|x - c|^2 / (2 sigma^2)) at the element nodes.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib as L
from .discretization import CartesianMesh, FiniteElementSpace, rule_points
from .fields import AffineField, SeparableField, VectorField

TWO_PI = 2.0 * math.pi


def advection_c2() -> VectorField:
    return VectorField(
        [
            SeparableField([(1.0, {1: lambda y: 1.0 + 0.5 * np.sin(TWO_PI * y)})]),
            SeparableField([(1.0, {2: lambda z: 0.5 + 0.5 * np.cos(TWO_PI * z)})]),
            SeparableField([(1.0, {0: lambda x: 0.25 + 0.25 * np.sin(TWO_PI * x)})]),
        ]
    )


def diffusivity_trig_3d() -> SeparableField:
    """2 + 0.25 sin 2 pi x + 0.17 cos 2 pi (2y + 0.13) + 0.11 sin 2 pi (3z + 0.37)
    + 0.07 sin 2 pi (x + 2y + 3z): the diffusivity of GenDiL's SIPDG test
    (test-diffusion-3d-generic.cpp:121-133) without its last term, whose
    expansion would need more separable terms than a field holds.  The last
    sine is expanded as sin(A + B + C) with A = 2 pi x, B = 4 pi y, C = 6 pi z."""
    sA, cA = (lambda s: np.sin(TWO_PI * s)), (lambda s: np.cos(TWO_PI * s))
    sB, cB = (lambda s: np.sin(2 * TWO_PI * s)), (lambda s: np.cos(2 * TWO_PI * s))
    sC, cC = (lambda s: np.sin(3 * TWO_PI * s)), (lambda s: np.cos(3 * TWO_PI * s))
    return SeparableField([
        (2.0, {}),
        (0.25, {0: sA}),
        (0.17, {1: lambda s: np.cos(TWO_PI * (2.0 * s + 0.13))}),
        (0.11, {2: lambda s: np.sin(TWO_PI * (3.0 * s + 0.37))}),
        (0.07, {0: sA, 1: cB, 2: cC}),
        (0.07, {0: cA, 1: sB, 2: cC}),
        (0.07, {0: cA, 1: cB, 2: sC}),
        (-0.07, {0: sA, 1: sB, 2: sC}),
    ])


def advection_c2_max_norm() -> float:
    # |beta|^2 = (1+0.5 s1)^2 + (0.5+0.5 c)^2 + (0.25+0.25 s2)^2, maximised at
    # s1 = 1, c = 1, s2 = 1: 2.25 + 1 + 0.25 = 3.5
    return math.sqrt(3.5)


def phase_space_6d_affine() -> AffineField:
    A = np.zeros((6, 6))
    A[0, 3] = A[1, 4] = A[2, 5] = 1.0
    c = np.zeros(6)
    c[3] = 0.1
    return AffineField(A, c)


def phase_space_6d_separable() -> VectorField:
    ident = lambda s: s  # noqa: E731
    comps = [
        SeparableField([(1.0, {3: ident})]),
        SeparableField([(1.0, {4: ident})]),
        SeparableField([(1.0, {5: ident})]),
        SeparableField.constant(0.1),
        SeparableField.constant(0.0),
        SeparableField.constant(0.0),
    ]
    return VectorField(comps)


def padaptive_orders(num_elements: int, seed: int = 2024, probs=(0.5, 0.35, 0.15)) -> np.ndarray:
    """Per-element orders p in {1, 2, 3} drawn with the given probabilities."""
    rng = np.random.default_rng(seed)
    return rng.choice(np.array([1, 2, 3], dtype=np.int8), size=num_elements, p=np.asarray(probs))


def phase_space_6d_padaptive(n: int = 8, seed: int = 2024):
    """6D phase-space workload on [0,1]^6, n elements per axis, periodic in
    the three space axes (no boundary terms in velocity), per-element orders
    drawn from {1,2,3} with probabilities {0.5, 0.35, 0.15}, q_e = p_e + 1
    Gauss points.  Returns (space, beta) with beta = (X3, X4, X5, 0.1, 0, 0)."""
    mesh = CartesianMesh([n] * 6, [1.0 / n] * 6, [0.0] * 6, [True, True, True, False, False, False])
    orders = padaptive_orders(mesh.num_cells, seed)
    from .discretization import MakePAdaptiveSpace

    return MakePAdaptiveSpace(mesh, orders, orders + 1), phase_space_6d_affine()


def advection_c2_numpy(X: np.ndarray) -> np.ndarray:
    """The benchmark field evaluated at points X[..., 3]."""
    b = np.empty(X.shape[:-1] + (3,))
    b[..., 0] = 1.0 + 0.5 * np.sin(TWO_PI * X[..., 1])
    b[..., 1] = 0.5 + 0.5 * np.cos(TWO_PI * X[..., 2])
    b[..., 2] = 0.25 + 0.25 * np.sin(TWO_PI * X[..., 0])
    return b


def noisy_pointwise_beta(mesh: CartesianMesh, q: int, seed: int = 2024, noise: float = 0.1):
    """Adjoint workload field: the benchmark field at every cell quadrature
    point and (normal component) at every face quadrature point, plus
    noise * N(0,1) drawn from `seed`.  Returns (cell [E, 3, q^3], faces[a]
    [F_a, q^2]) as float64 numpy arrays."""
    from .fields import cell_points, face_points

    rng = np.random.default_rng(seed)
    Xc = cell_points(mesh, q)  # [E, q^3, 3]
    cell = np.ascontiguousarray(np.transpose(advection_c2_numpy(Xc), (0, 2, 1)))
    cell += noise * rng.standard_normal(cell.shape)
    faces = []
    for a in range(3):
        Xf = face_points(mesh, q, a)
        fa = advection_c2_numpy(Xf)[..., a].copy()
        fa += noise * rng.standard_normal(fa.shape)
        faces.append(np.ascontiguousarray(fa))
    return cell, faces


def node_coordinates(fes: FiniteElementSpace) -> np.ndarray:
    """Physical coordinates of every DoF node, [N, dim] (uniform order)."""
    m = fes.mesh
    D = m.dim
    n = fes.fe.nodes
    xn, _ = rule_points(fes.fe.basis, n)
    E = m.num_cells
    e = np.arange(E)
    sub = []
    r = e.copy()
    for a in range(D):
        sub.append(r % m.sizes[a])
        r //= m.sizes[a]
    loc = np.arange(n ** D)
    ls = []
    r = loc.copy()
    for a in range(D):
        ls.append(r % n)
        r //= n
    X = np.empty((E, n ** D, D))
    for a in range(D):
        X[:, :, a] = (m.origin[a] + m.h[a] * sub[a][:, None]) + m.h[a] * xn[ls[a]][None, :]
    return X.reshape(-1, D)


def gaussian_bump(fes: FiniteElementSpace, center, sigma: float) -> np.ndarray:
    X = node_coordinates(fes)
    c = np.asarray(center, dtype=np.float64)
    return np.exp(-np.sum((X - c) ** 2, axis=1) / (2.0 * sigma * sigma))
