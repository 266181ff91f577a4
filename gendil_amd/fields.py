"""Coefficient fields that can cross the C-ABI.

GenDiL takes sigma(X), adv(X, v) and g(X) as device lambdas
(massoperator.hpp:270, advectionoperator.hpp:1416-1528).  Lambdas cannot
cross a C-ABI, so fields are described by one of:

  SeparableField  f(X) = sum_r scale_r * prod_d f_{r,d}(X_d): each 1D factor is
                  tabulated host-side at the exact positions the kernels use
                  (per element: Gauss points, then xi=0, xi=1).
  AffineField     beta(X) = c + A X, evaluated in-kernel.
  PointwiseField  beta given at every cell quadrature point and (normal
                  component) at every face quadrature point, device arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Sequence, Tuple

import numpy as np

from . import _lib as L
from .discretization import CartesianMesh, axis_points


@dataclass
class SeparableField:
    """sum over terms of scale * prod_{axis} fn_axis(X_axis); axes not listed are 1."""

    terms: List[Tuple[float, Dict[int, Callable[[np.ndarray], np.ndarray]]]] = field(default_factory=list)

    @staticmethod
    def constant(c: float) -> "SeparableField":
        return SeparableField([(float(c), {})]) if c != 0.0 else SeparableField([])

    def tabulate(self, mesh: CartesianMesh, q: int):
        """Return (GdSeparable, data array kept alive by the caller)."""
        if len(self.terms) > L.MAX_TERMS:
            raise ValueError(f"at most {L.MAX_TERMS} separable terms")
        s = L.GdSeparable()
        s.rank = len(self.terms)
        chunks = []
        off = 0
        for r, (scale, fns) in enumerate(self.terms):
            s.scale[r] = float(scale)
            for d in range(L.MAX_DIM):
                s.offset[r][d] = -1
            for axis, fn in fns.items():
                if not 0 <= axis < mesh.dim:
                    raise ValueError("axis out of range")
                xs = axis_points(mesh, q, axis)
                vals = np.asarray(fn(xs), dtype=np.float64)
                if vals.shape != xs.shape:
                    vals = np.broadcast_to(vals, xs.shape).astype(np.float64)
                s.offset[r][axis] = off
                chunks.append(vals)
                off += vals.size
        data = np.ascontiguousarray(np.concatenate(chunks)) if chunks else np.zeros(1)
        s.data = data.ctypes.data_as(L._pd)
        s.data_len = data.size
        return s, data


@dataclass
class VectorField:
    """One SeparableField per component."""

    components: Sequence[SeparableField]


@dataclass
class AffineField:
    A: np.ndarray  # dim x dim
    c: np.ndarray  # dim


@dataclass
class PointwiseField:
    """Advection field per quadrature point (FP64 torch device tensors).

    cell: [E, dim, q^dim]; faces[a]: [F_a, q^(dim-1)] normal component, where
    F_a = prod_{d != a} M_d * (M_a + (0 if periodic_a else 1)) and global face
    f_a sits between element f_a - 1 (its + face) and element f_a (its - face).
    """

    cell: object
    faces: Sequence[object]


def face_counts(mesh: CartesianMesh) -> List[int]:
    out = []
    for a in range(mesh.dim):
        n = 1
        for d in range(mesh.dim):
            n *= mesh.sizes[d] if d != a else mesh.sizes[a] + (0 if mesh.periodic[a] else 1)
        out.append(n)
    return out


def cell_points(mesh: CartesianMesh, q: int) -> np.ndarray:
    """Physical coordinates of every cell quadrature point: [E, q^dim, dim]."""
    D = mesh.dim
    per_axis = [axis_points(mesh, q, a).reshape(mesh.sizes[a], q + 2)[:, :q] for a in range(D)]
    E = mesh.num_cells
    e = np.arange(E)
    sub = []
    r = e.copy()
    for a in range(D):
        sub.append(r % mesh.sizes[a])
        r //= mesh.sizes[a]
    k = np.arange(q ** D)
    ks = []
    r = k.copy()
    for a in range(D):
        ks.append(r % q)
        r //= q
    X = np.empty((E, q ** D, D))
    for a in range(D):
        X[:, :, a] = per_axis[a][sub[a][:, None], ks[a][None, :]]
    return X


def face_points(mesh: CartesianMesh, q: int, axis: int) -> np.ndarray:
    """Coordinates of every face quadrature point of axis `axis`, as seen from
    the element on the face's + side of the normal (the element whose - face it
    is), or from element M-1 for the last face of a non-periodic axis:
    [F_a, q^(dim-1), dim]."""
    D = mesh.dim
    ext = [mesh.sizes[d] if d != axis else mesh.sizes[axis] + (0 if mesh.periodic[axis] else 1) for d in range(D)]
    F = int(np.prod(ext))
    f = np.arange(F)
    sub = []
    r = f.copy()
    for d in range(D):
        sub.append(r % ext[d])
        r //= ext[d]
    tang = [d for d in range(D) if d != axis]
    k = np.arange(q ** (D - 1))
    ks = {}
    r = k.copy()
    for d in tang:
        ks[d] = r % q
        r //= q
    X = np.empty((F, q ** (D - 1), D))
    for d in tang:
        pts = axis_points(mesh, q, d).reshape(mesh.sizes[d], q + 2)[:, :q]
        X[:, :, d] = pts[sub[d][:, None], ks[d][None, :]]
    pa = axis_points(mesh, q, axis).reshape(mesh.sizes[axis], q + 2)
    fa = sub[axis]
    last = fa == mesh.sizes[axis]  # non-periodic top boundary: xi = 1 of element M-1
    xa = np.where(last, pa[np.minimum(fa, mesh.sizes[axis] - 1), q + 1], pa[np.minimum(fa, mesh.sizes[axis] - 1), q])
    X[:, :, axis] = xa[:, None]
    return X
