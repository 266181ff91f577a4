"""Device-resident Krylov solves over the C-ABI (gd_gmres_f64 / gd_gmres_mixed /
gd_cg_f64 / gd_cg_mixed / gd_bicgstab_f64, implemented by the header-only solvers of
include/gendil_amd/krylov.hpp).

They solve K x = b for K = mass_coef * M + adv_coef * A of one operator plan,
with GenDiL's GMRES_no_alloc contract (gmres.hpp:74-288): x holds x0 on
entry, the relative residual is ||b - K x|| / ||b||.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib as L
from .operators import _OperatorBase, _stream


@dataclass
class SolveResult:
    success: bool
    iterations: int
    refinements: int
    final_rel_residual: float
    seconds: float


def _run(fn, op: _OperatorBase, b, x, restart, max_iters, rel_tol, max_refinements, inner_rel_tol):
    for t, name in ((b, "b"), (x, "x")):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == op.num_dofs):
            raise ValueError(f"{name} must be a contiguous float64 device vector of length {op.num_dofs}")
    p = L.GdSolveParams(restart, max_iters, rel_tol, max_refinements, inner_rel_tol)
    r = L.GdSolveResult()
    rc = fn(op.plan.handle, op.mass_coef, op.adv_coef, b.data_ptr(), x.data_ptr(), C.byref(p), C.byref(r), _stream())
    if rc not in (L.GD_OK, L.GD_ERR_NOT_CONVERGED):
        L.check(rc, "solve")
    return SolveResult(bool(r.success), r.iterations, r.refinements, r.final_rel_residual, r.seconds)


def gmres(op, b, x, restart=30, max_iters=10000, rel_tol=1e-10):
    """All-FP64 restarted GMRES(m)."""
    return _run(L.lib().gd_gmres_f64, op, b, x, restart, max_iters, rel_tol, 100000, 0.0)


def gmres_mixed(op, b, x, restart=30, max_iters=10000, rel_tol=1e-10, inner_rel_tol=1e-6, max_refinements=1000):
    """FP32 operator apply and FP32 basis inside each Arnoldi cycle; FP64
    residual recompute, FP64 inner products/orthogonalisation coefficients and
    FP64 iterative-refinement update of x."""
    return _run(L.lib().gd_gmres_mixed, op, b, x, restart, max_iters, rel_tol, max_refinements, inner_rel_tol)


def cg(op, b, x, max_iters=10000, rel_tol=1e-10):
    """Unpreconditioned CG (conjugategradient.hpp:217) for SPD operators,
    device-resident (no dot product synchronises the stream)."""
    return _run(L.lib().gd_cg_f64, op, b, x, 30, max_iters, rel_tol, 1, 0.0)


def cg_mixed(op, b, x, max_iters=10000, rel_tol=1e-10, inner_rel_tol=1e-6, max_refinements=1000):
    """CG for SPD operators with FP32 inner solves (FP32 vectors and apply)
    refined in FP64 on the true residual (gd_cg_mixed)."""
    return _run(L.lib().gd_cg_mixed, op, b, x, 30, max_iters, rel_tol, max_refinements, inner_rel_tol)


def bicgstab(op, b, x, max_iters=10000, rel_tol=1e-10):
    """Unpreconditioned BiCGSTAB (BiCGSTAB_no_alloc, bicgstab.hpp:44-157),
    device-resident; stops at ||r|| / ||b|| < rel_tol (the recurrence
    residual, as the reference does) and reports the true residual."""
    return _run(L.lib().gd_bicgstab_f64, op, b, x, 30, max_iters, rel_tol, 1, 0.0)
