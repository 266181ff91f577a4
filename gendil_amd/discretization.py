"""Mesh / finite element / integration-rule descriptions with GenDiL's names.

These mirror GenDiL's factories so call sites read the same:
  CartesianMesh<D>(sizes, h, origin[, periodic])   cartesianndmesh.hpp:16-80
  Cartesian{1,2,3}DMesh(h, n...)                   cartesian3dmesh.hpp:32-47 (and 1D/2D)
  MakeCartesianProductMesh(meshes...)              productmesh.hpp:154-225
  MakeLobattoFiniteElement / MakeLegendreFiniteElement   finiteelement.hpp:103-109
  MakeFiniteElementSpace(mesh, fe)                 finiteelementspace.hpp:69,84
  MakeIntegrationRule(IntegrationRuleNumPoints)    integrationrule.hpp:80
Only uniform-order Cartesian L2 spaces plus the per-element-order extension.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L


@dataclass
class CartesianMesh:
    """N-D Cartesian mesh, lexicographic elements (x fastest), optionally periodic."""

    sizes: Sequence[int]
    h: Sequence[float]
    origin: Sequence[float]
    periodic: object = False  # bool or per-axis sequence

    def __post_init__(self):
        self.sizes = tuple(int(s) for s in self.sizes)
        self.h = tuple(float(v) for v in self.h)
        self.origin = tuple(float(v) for v in self.origin)
        if isinstance(self.periodic, (bool, int, np.bool_)):
            self.periodic = tuple(bool(self.periodic) for _ in self.sizes)
        else:
            self.periodic = tuple(bool(v) for v in self.periodic)
        if not (len(self.sizes) == len(self.h) == len(self.origin) == len(self.periodic)):
            raise ValueError("sizes, h, origin, periodic must have one entry per dimension")
        if not 1 <= self.dim <= L.MAX_DIM:
            raise ValueError("dimension must be 1..6")

    @property
    def dim(self) -> int:
        return len(self.sizes)

    @property
    def num_cells(self) -> int:
        return int(np.prod(self.sizes))

    @staticmethod
    def unit(dim: int, n: int, periodic=False) -> "CartesianMesh":
        """[0,1]^dim with n elements per axis."""
        return CartesianMesh([n] * dim, [1.0 / n] * dim, [0.0] * dim, periodic)


def Cartesian1DMesh(h: float, n: int, origin: float = 0.0, periodic: bool = False) -> CartesianMesh:
    """n cells of size h from origin (Cartesian1DMesh)."""
    return CartesianMesh([n], [h], [origin], periodic)


def Cartesian2DMesh(h: float, nx: int, ny: int, origin=(0.0, 0.0), periodic=False) -> CartesianMesh:
    """nx x ny square cells of size h (Cartesian2DMesh)."""
    return CartesianMesh([nx, ny], [h, h], origin, periodic)


def Cartesian3DMesh(h: float, nx: int, ny: int, nz: int, origin=(0.0, 0.0, 0.0), periodic=False) -> CartesianMesh:
    """nx x ny x nz cubic cells of size h (Cartesian3DMesh(h, nx, ny, nz),
    cartesian3dmesh.hpp:32-36)."""
    return CartesianMesh([nx, ny, nz], [h, h, h], origin, periodic)


def MakeCartesianProductMesh(*meshes: CartesianMesh) -> CartesianMesh:
    """Tensor product of Cartesian meshes (MakeCartesianProductMesh,
    productmesh.hpp:154-225) as the Cartesian mesh over the concatenated axes.

    The two index the same: the product numbers cell (head, tail) as
    head + cells(head) * tail (productmesh.hpp:91-94), which unrolled over
    the factors is the x-fastest lexicographic order of the concatenated
    axes; a face normal to axis d of the product is the sub-mesh's face
    normal to that axis, with the sub-mesh's neighbour and boundary
    (productmesh.hpp:104-134); cells are products of the factors' cells,
    so coordinates, sizes and periodicity concatenate."""
    if not meshes:
        raise ValueError("MakeCartesianProductMesh needs at least one mesh")
    sizes, h, origin, periodic = [], [], [], []
    for m in meshes:
        sizes += list(m.sizes)
        h += list(m.h)
        origin += list(m.origin)
        periodic += list(m.periodic)
    return CartesianMesh(sizes, h, origin, periodic)


@dataclass(frozen=True)
class FiniteElement:
    order: int
    basis: int  # L.BASIS_GLL or L.BASIS_GL

    @property
    def nodes(self) -> int:
        return self.order + 1


def MakeLobattoFiniteElement(order: int) -> FiniteElement:
    """Lagrange element on Gauss-Lobatto nodes (finiteelement.hpp:109)."""
    return FiniteElement(int(order), L.BASIS_GLL)


def MakeLegendreFiniteElement(order: int) -> FiniteElement:
    """Lagrange element on Gauss-Legendre nodes (finiteelement.hpp:103)."""
    return FiniteElement(int(order), L.BASIS_GL)


@dataclass(frozen=True)
class IntegrationRule:
    num_points: int  # Gauss-Legendre points per dimension


def MakeIntegrationRule(num_points: int) -> IntegrationRule:
    return IntegrationRule(int(num_points))


@dataclass
class FiniteElementSpace:
    mesh: CartesianMesh
    fe: FiniteElement
    # p-adaptive extension: per-element orders / rules (int8 arrays)
    elem_order: Optional[np.ndarray] = None
    elem_quad: Optional[np.ndarray] = None
    # L2Restriction shift: first DoF of this space in a vector shared by
    # several spaces (restriction.hpp:38)
    shift: int = 0

    @property
    def dim(self) -> int:
        return self.mesh.dim

    @property
    def padaptive(self) -> bool:
        return self.elem_order is not None

    def dofs_per_element(self) -> int:
        return self.fe.nodes ** self.dim

    def GetNumberOfFiniteElementDofs(self) -> int:
        if self.padaptive:
            return int(np.sum((self.elem_order.astype(np.int64) + 1) ** self.dim))
        return self.mesh.num_cells * self.dofs_per_element()

    num_dofs = property(GetNumberOfFiniteElementDofs)

    def dof_offsets(self) -> np.ndarray:
        """Start of each element's DoFs (E+1 entries)."""
        if not self.padaptive:
            return np.arange(self.mesh.num_cells + 1, dtype=np.int64) * self.dofs_per_element()
        nd = (self.elem_order.astype(np.int64) + 1) ** self.dim
        return np.concatenate([[0], np.cumsum(nd)]).astype(np.int64)


@dataclass(frozen=True)
class L2Restriction:
    """GenDiL's L2Restriction{shift} (restriction.hpp:38): the space's DoFs
    start at `shift` in a vector shared with other spaces."""

    shift: int = 0


def MakeFiniteElementSpace(mesh: CartesianMesh, fe: FiniteElement, restriction: Optional[L2Restriction] = None
                           ) -> FiniteElementSpace:
    return FiniteElementSpace(mesh, fe, shift=int(restriction.shift) if restriction is not None else 0)


def MakePAdaptiveSpace(mesh: CartesianMesh, elem_order, elem_quad, basis: int = L.BASIS_GLL) -> FiniteElementSpace:
    """Per-element order p_e and Gauss rule q_e (no GenDiL counterpart)."""
    eo = np.ascontiguousarray(np.asarray(elem_order, dtype=np.int8).reshape(-1))
    eq = np.ascontiguousarray(np.asarray(elem_quad, dtype=np.int8).reshape(-1))
    if eo.size != mesh.num_cells or eq.size != mesh.num_cells:
        raise ValueError("one order and one rule per element")
    return FiniteElementSpace(mesh, FiniteElement(int(eo.max()), basis), eo, eq)


def rule_points(basis: int, n: int):
    """Points and weights of the 1D rule on [0,1] used by the kernels."""
    pts = np.zeros(n)
    wts = np.zeros(n)
    L.check(
        L.lib().gd_rule_points(basis, n, pts.ctypes.data_as(L._pd), wts.ctypes.data_as(L._pd)),
        "gd_rule_points",
    )
    return pts, wts


def axis_points(mesh: CartesianMesh, q: int, axis: int) -> np.ndarray:
    """Sample positions of axis `axis`: per element e, the q Gauss points then
    xi = 0 and xi = 1, each X = (origin + h e) + h xi (hypercubecell.hpp:41-52)."""
    xq, _ = rule_points(L.BASIS_GL, q)
    xi = np.concatenate([xq, [0.0, 1.0]])
    e = np.arange(mesh.sizes[axis], dtype=np.float64)
    o = mesh.origin[axis] + mesh.h[axis] * e
    return (o[:, None] + mesh.h[axis] * xi[None, :]).reshape(-1)


def make_desc(fes: FiniteElementSpace, ir: Optional[IntegrationRule]):
    m = fes.mesh
    d = L.GdDesc()
    d.dim = m.dim
    for i in range(m.dim):
        d.sizes[i] = m.sizes[i]
        d.h[i] = m.h[i]
        d.origin[i] = m.origin[i]
        d.periodic[i] = int(m.periodic[i])
    d.order = fes.fe.order
    d.num_quad = ir.num_points if ir is not None else fes.fe.order + 1
    d.basis = fes.fe.basis
    keep = []
    if fes.padaptive:
        d.elem_order = fes.elem_order.ctypes.data_as(C.POINTER(C.c_int8))
        d.elem_quad = fes.elem_quad.ctypes.data_as(C.POINTER(C.c_int8))
        keep = [fes.elem_order, fes.elem_quad]
    return d, keep
