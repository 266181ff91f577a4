"""MI355X (gfx950) matrix-free DG operator apply with GenDiL's operator API.

The compute path is libgendil_amd.so (hand-written HIP kernels behind the
C-ABI of include/gendil_amd.h); this package is the Python host mirror used by
tests, benchmarks and the autograd wrapper.
"""
from . import _lib
from .discretization import (
    Cartesian1DMesh,
    Cartesian2DMesh,
    Cartesian3DMesh,
    CartesianMesh,
    FiniteElement,
    FiniteElementSpace,
    IntegrationRule,
    L2Restriction,
    MakeFiniteElementSpace,
    MakeIntegrationRule,
    MakeLegendreFiniteElement,
    MakeCartesianProductMesh,
    MakeLobattoFiniteElement,
    MakePAdaptiveSpace,
    axis_points,
    rule_points,
)
from .fields import AffineField, PointwiseField, SeparableField, VectorField, cell_points, face_counts, face_points
from .random import fill_random

__all__ = [
    "Cartesian1DMesh",
    "Cartesian2DMesh",
    "Cartesian3DMesh",
    "CartesianMesh",
    "FiniteElement",
    "FiniteElementSpace",
    "IntegrationRule",
    "L2Restriction",
    "MakeFiniteElementSpace",
    "MakeIntegrationRule",
    "MakeLegendreFiniteElement",
    "MakeCartesianProductMesh",
    "MakeLobattoFiniteElement",
    "MakePAdaptiveSpace",
    "axis_points",
    "rule_points",
    "AffineField",
    "PointwiseField",
    "SeparableField",
    "VectorField",
    "cell_points",
    "face_counts",
    "face_points",
    "fill_random",
]


def __getattr__(name):
    # operators / solvers / autograd pull in torch; import on first use
    if name in (
        "Plan",
        "MassOperator",
        "AdvectionOperator",
        "LinearCombinationOperator",
        "MakeMassFiniteElementOperator",
        "MakeAdvectionOperator",
        "MakeMassAdvectionOperator",
        "AdvectionFaceOperator",
        "AdvectionNonconformingFaceOperator",
        "CartesianInteriorFaceConnectivity",
        "CartesianIntermeshFaceConnectivity",
        "NonconformingCartesianIntermeshFaceConnectivity",
        "MakeCartesianInteriorFaceConnectivity",
        "MakeAdvectionFaceOperator",
        "DiffusionOperator",
        "MakeDiffusionOperator",
    ):
        from . import operators

        return getattr(operators, name)
    if name in ("gmres", "gmres_mixed", "cg", "cg_mixed", "bicgstab", "SolveResult"):
        from . import solvers

        return getattr(solvers, name)
    if name in ("AdvectionApply", "advection_apply"):
        from . import autograd

        return getattr(autograd, name)
    raise AttributeError(name)
