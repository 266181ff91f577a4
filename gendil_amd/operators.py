"""Operators with GenDiL's call shape, backed by the gfx950 kernels.

  MakeMassFiniteElementOperator(fes, ir, sigma)       massoperator.hpp:270
  MakeAdvectionOperator(fes, ir, adv[, bc])           advectionoperator.hpp:1416-1528
  MakeDiffusionOperator(fes, ir, mu, sigma, kappa)    diffusionoperator.hpp:519-537
  op(x, y)  overwrites y (L2 semantics, massoperator.hpp:209-216,
            advectionoperator.hpp:1070-1076)

Vectors are torch tensors on the GPU (float64 or float32); the apply is
launched on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from .discretization import FiniteElementSpace, IntegrationRule, make_desc
from .fields import AffineField, PointwiseField, SeparableField, VectorField, face_counts


_get_device = getattr(torch._C, "_cuda_getDevice", None)
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _stream():
    """The current HIP stream of the current device (torch's current stream),
    read without building a torch Stream object: applies of small meshes are
    submitted in a few microseconds, and that object costs about as much."""
    if _get_device is not None and _raw_stream is not None:
        return C.c_void_p(_raw_stream(_get_device()))
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_vec(t, n, dtype, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA (HIP) torch tensor")
    if t.dtype != dtype or t.numel() != n or not t.is_contiguous():
        raise ValueError(f"{name}: expected contiguous {dtype} with {n} entries, got {t.dtype} {tuple(t.shape)}")


def _overlap(a: torch.Tensor, b: torch.Tensor) -> bool:
    """True when the device storage ranges of two contiguous tensors intersect."""
    a0, b0 = a.data_ptr(), b.data_ptr()
    a1 = a0 + a.numel() * a.element_size()
    b1 = b0 + b.numel() * b.element_size()
    return a0 < b1 and b0 < a1


def _span(t, plan, name):
    """The space's part of `t`: all of it, or the plan's num_dofs entries from
    its space's L2Restriction shift when `t` is a vector shared by several
    spaces (MakeReadOnlyElementTensorView on a shifted space,
    elementtensorview.hpp:42-64).  Sizes come from the plan (fixed when it was
    created), not recomputed from the space on every apply."""
    n, shift = plan.num_dofs, plan.shift
    if shift == 0 and isinstance(t, torch.Tensor) and t.numel() == n:
        return t
    if not isinstance(t, torch.Tensor) or t.dim() != 1:
        raise TypeError(f"{name} must be a 1-D torch tensor")
    if shift < 0 or shift + n > t.numel():
        raise ValueError(f"{name}: {t.numel()} entries cannot hold {n} DoFs from shift {shift}")
    return t.narrow(0, shift, n)


class Plan:
    """Owns one gd_plan: mesh, basis tables and coefficients on the device."""

    def __init__(self, fes: FiniteElementSpace, ir: Optional[IntegrationRule]):
        self.fes = fes
        self.ir = ir
        self.q = ir.num_points if ir is not None else fes.fe.order + 1
        desc, keep = make_desc(fes, ir)
        h = C.c_void_p()
        L.check(L.lib().gd_plan_create(C.byref(desc), C.byref(h)), "gd_plan_create")
        self.handle = h
        self._keep = [keep]
        self.num_dofs = int(L.lib().gd_plan_num_dofs(self.handle))
        self.shift = int(fes.shift)
        self.beta_kind = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                L.lib().gd_plan_destroy(h)
            except Exception:
                pass
            self.handle = None

    # ---- coefficients ----
    def set_sigma(self, sigma):
        if sigma is None:
            return
        if isinstance(sigma, (int, float)):
            sigma = SeparableField.constant(float(sigma))
        s, data = sigma.tabulate(self.fes.mesh, self.q)
        L.check(L.lib().gd_plan_set_sigma(self.handle, C.byref(s)), "gd_plan_set_sigma")

    def set_beta(self, adv):
        D = self.fes.dim
        if isinstance(adv, VectorField):
            if len(adv.components) != D:
                raise ValueError("one component per dimension")
            arr = (L.GdSeparable * D)()
            keep = []
            for d, comp in enumerate(adv.components):
                s, data = comp.tabulate(self.fes.mesh, self.q)
                arr[d] = s
                keep.append(data)
            L.check(L.lib().gd_plan_set_beta_separable(self.handle, arr), "gd_plan_set_beta_separable")
            self.beta_kind = L.BETA_SEPARABLE
        elif isinstance(adv, AffineField):
            A = np.ascontiguousarray(np.asarray(adv.A, dtype=np.float64).reshape(D, D))
            c = np.ascontiguousarray(np.asarray(adv.c, dtype=np.float64).reshape(D))
            L.check(
                L.lib().gd_plan_set_beta_affine(self.handle, A.ctypes.data_as(L._pd), c.ctypes.data_as(L._pd)),
                "gd_plan_set_beta_affine",
            )
            self.beta_kind = L.BETA_AFFINE
        elif isinstance(adv, PointwiseField):
            self.set_beta_pointwise(adv.cell, adv.faces)
        else:
            raise TypeError("adv must be VectorField, AffineField or PointwiseField")

    def set_beta_pointwise(self, cell: torch.Tensor, faces: Sequence[torch.Tensor]):
        D = self.fes.dim
        faces = list(faces)
        if self.fes.padaptive:
            raise ValueError("pointwise beta needs a uniform order and rule")
        if len(faces) != D:
            raise ValueError(f"pointwise beta needs one face array per axis ({D}), got {len(faces)}")
        # sizes the kernels index: cell [E][D][q^D], face a [F_a][q^(D-1)]
        q = self.q
        want = self.fes.mesh.num_cells * D * q ** D
        if cell.numel() != want:
            raise ValueError(f"pointwise beta: cell array has {cell.numel()} values, expected E*D*q^D = {want}")
        for a, (f, Fa) in enumerate(zip(faces, face_counts(self.fes.mesh))):
            if f.numel() != Fa * q ** (D - 1):
                raise ValueError(f"pointwise beta: face array {a} has {f.numel()} values, "
                                 f"expected F_a*q^(D-1) = {Fa * q ** (D - 1)}")
        for t in [cell, *faces]:
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise ValueError("pointwise beta must be contiguous float64 device tensors")
        arr = (C.c_void_p * D)(*[f.data_ptr() for f in faces])
        L.check(L.lib().gd_plan_set_beta_pointwise_f64(self.handle, cell.data_ptr(), arr), "set_beta_pointwise")
        self._beta_pw = (cell, list(faces))
        self.beta_kind = L.BETA_POINTWISE

    def set_diffusion(self, mu, sigma: float, kappa: float):
        """SIPDG diffusivity mu (separable field or constant), symmetry
        parameter sigma and penalty kappa (gd_plan_set_diffusion)."""
        if mu is None or isinstance(mu, (int, float)):
            mu = SeparableField.constant(1.0 if mu is None else float(mu))
        if not isinstance(mu, SeparableField):
            raise TypeError("the diffusivity must be a SeparableField or a constant")
        s, data = mu.tabulate(self.fes.mesh, self.q)
        L.check(L.lib().gd_plan_set_diffusion(self.handle, C.byref(s), float(sigma), float(kappa)),
                "gd_plan_set_diffusion")

    def set_bc(self, bc):
        if bc is None:
            L.check(L.lib().gd_plan_set_bc(self.handle, L.BC_NONE, None), "gd_plan_set_bc")
            return
        if isinstance(bc, (int, float)):
            bc = SeparableField.constant(float(bc))
        s, data = bc.tabulate(self.fes.mesh, self.q)
        L.check(L.lib().gd_plan_set_bc(self.handle, L.BC_INFLOW, C.byref(s)), "gd_plan_set_bc")

    # ---- kernel variant selection ----
    def autotune(self, dtype=torch.float64, mass_coef=0.0, adv_coef=1.0, reps=5):
        """Time every compiled variant on scratch vectors and keep the fastest.
        Returns (lanes per element, ms per apply); lanes 0 = generic kernel."""
        nbytes = 8 if dtype == torch.float64 else 4
        chosen = C.c_int32()
        best = C.c_double()
        L.check(L.lib().gd_plan_autotune(self.handle, nbytes, mass_coef, adv_coef, reps, C.byref(chosen),
                                         C.byref(best)), "gd_plan_autotune")
        return chosen.value, best.value

    def set_variant(self, dtype, lanes: int):
        """Pin the kernel variant: -1 default, 0 generic, 1/2/4 lanes per element
        (register kernel), 17/18/20 (Kronecker kernel, 16 + lanes), 48..55 (LDS-staged
        Kronecker kernel: +1 outputs staged through LDS, +2 128-thread blocks, +4 one term at
        a time), 56 / 57 (streaming form: one wave per block, persistent; 56 one tile buffer and two
        waves per SIMD, 57 two tile buffers with the next tile's copies in flight), 64 (4D-6D, order 1,
        constant coefficients: two-node Kronecker kernel);
        p-adaptive plans: the layout index."""
        nbytes = 8 if dtype == torch.float64 else 4
        L.check(L.lib().gd_plan_set_variant(self.handle, nbytes, lanes), "gd_plan_set_variant")

    def query_variant(self, dtype, mass_coef: float, adv_coef: float) -> int:
        """The kernel variant an apply with these coefficients would run."""
        nbytes = 8 if dtype == torch.float64 else 4
        v = C.c_int32()
        L.check(L.lib().gd_plan_query_variant(self.handle, nbytes, mass_coef, adv_coef, C.byref(v)),
                "gd_plan_query_variant")
        return v.value

    # ---- applies ----
    def apply(self, mass_coef: float, adv_coef: float, x: torch.Tensor, y: torch.Tensor):
        out = y
        x, y = _span(x, self, "x"), _span(y, self, "y")
        if x.dtype == torch.float64:
            _check_vec(x, self.num_dofs, torch.float64, "x")
            _check_vec(y, self.num_dofs, torch.float64, "y")
            L.check(L.lib().gd_apply_f64(self.handle, mass_coef, adv_coef, x.data_ptr(), y.data_ptr(), _stream()),
                    "gd_apply_f64")
        elif x.dtype == torch.float32:
            _check_vec(x, self.num_dofs, torch.float32, "x")
            _check_vec(y, self.num_dofs, torch.float32, "y")
            L.check(L.lib().gd_apply_f32(self.handle, mass_coef, adv_coef, x.data_ptr(), y.data_ptr(), _stream()),
                    "gd_apply_f32")
        else:
            raise TypeError("x must be float64 or float32")
        return out

    def apply_transpose(self, mass_coef: float, adv_coef: float, r: torch.Tensor, z: torch.Tensor):
        _check_vec(r, self.num_dofs, torch.float64, "r")
        _check_vec(z, self.num_dofs, torch.float64, "z")
        L.check(L.lib().gd_apply_T_f64(self.handle, mass_coef, adv_coef, r.data_ptr(), z.data_ptr(), _stream()),
                "gd_apply_T_f64")
        return z

    def grad_beta(self, adv_coef: float, u: torch.Tensor, r: torch.Tensor, dcell: torch.Tensor, dfaces):
        _check_vec(u, self.num_dofs, torch.float64, "u")
        _check_vec(r, self.num_dofs, torch.float64, "r")
        arr = (C.c_void_p * self.fes.dim)(*[f.data_ptr() for f in dfaces])
        L.check(
            L.lib().gd_adv_grad_beta_f64(self.handle, adv_coef, u.data_ptr(), r.data_ptr(), dcell.data_ptr(), arr,
                                         _stream()),
            "gd_adv_grad_beta_f64",
        )


class _OperatorBase:
    mass_coef = 0.0
    adv_coef = 0.0

    def __init__(self, plan: Plan):
        self.plan = plan

    @property
    def num_dofs(self):
        return self.plan.num_dofs

    def __call__(self, x: torch.Tensor, y: torch.Tensor):
        """y = op(x), y overwritten."""
        return self.plan.apply(self.mass_coef, self.adv_coef, x, y)

    Apply = __call__

    def transpose(self, r: torch.Tensor, z: torch.Tensor):
        return self.plan.apply_transpose(self.mass_coef, self.adv_coef, r, z)


class MassOperator(_OperatorBase):
    mass_coef = 1.0


class AdvectionOperator(_OperatorBase):
    adv_coef = 1.0


class LinearCombinationOperator(_OperatorBase):
    """mass_coef * M + adv_coef * A on one plan, applied by one fused kernel."""

    def __init__(self, plan: Plan, mass_coef: float, adv_coef: float):
        super().__init__(plan)
        self.mass_coef = float(mass_coef)
        self.adv_coef = float(adv_coef)


def MakeMassFiniteElementOperator(fes, ir, sigma=None) -> MassOperator:
    plan = Plan(fes, ir)
    plan.set_sigma(sigma)
    return MassOperator(plan)


def MakeAdvectionOperator(fes, ir, adv, bc=None) -> AdvectionOperator:
    """Without `bc`, boundary faces are skipped (GenDiL's Empty BC); with `bc`
    the inflow value u+ = g(X) is used on boundary faces."""
    plan = Plan(fes, ir)
    plan.set_beta(adv)
    plan.set_bc(bc)
    return AdvectionOperator(plan)


def MakeMassAdvectionOperator(fes, ir, adv, sigma=None, bc=None, mass_coef=1.0, adv_coef=1.0):
    """mass_coef * M(sigma) + adv_coef * A(adv, bc) in one pass (GenDiL's fused
    MassAdvectionOperator, massadvectionoperator.hpp:447, has no boundary terms;
    here boundary terms follow `bc` as in MakeAdvectionOperator)."""
    plan = Plan(fes, ir)
    plan.set_sigma(sigma)
    plan.set_beta(adv)
    plan.set_bc(bc)
    return LinearCombinationOperator(plan, mass_coef, adv_coef)


# ---- face-centric operators (accumulating) ----------------------------------

class CartesianInteriorFaceConnectivity:
    """The interior faces of a Cartesian mesh, every axis
    (MakeCartesianInteriorFaceConnectivity, cartesianglobalfaceconnectivity.hpp:
    89-108): no periodic wrap, no boundary faces."""

    def __init__(self, sizes):
        self.sizes = tuple(int(s) for s in sizes)


def MakeCartesianInteriorFaceConnectivity(sizes) -> CartesianInteriorFaceConnectivity:
    return CartesianInteriorFaceConnectivity(sizes)


class CartesianIntermeshFaceConnectivity:
    """Face `face` (GenDiL's local numbering: 0..D-1 the -axis faces,
    D..2D-1 the +axis faces) of the minus mesh glued to the opposite face of
    the plus mesh (cartesianintermeshfaceconnectivity.hpp:36-180)."""

    def __init__(self, dim: int, face: int, sizes_minus, sizes_plus):
        self.dim, self.face = int(dim), int(face)
        if not 0 <= self.face < 2 * self.dim:
            raise ValueError("face out of range")
        self.sizes_minus = tuple(int(s) for s in sizes_minus)
        self.sizes_plus = tuple(int(s) for s in sizes_plus)
        axis = self.face % self.dim
        for d in range(self.dim):
            if d != axis and self.sizes_minus[d] != self.sizes_plus[d]:
                raise ValueError("tangential sizes of the two meshes must match")


class NonconformingCartesianIntermeshFaceConnectivity(CartesianIntermeshFaceConnectivity):
    """Face `face` of a coarse mesh (minus) glued to a mesh with an integer
    number of fine cells per coarse cell along each tangential axis (plus)
    (nonconformingcartesianintermeshfaceconnectivity.hpp:73-243)."""

    def __init__(self, dim: int, face: int, sizes_coarse, sizes_fine):
        self.dim, self.face = int(dim), int(face)
        if not 0 <= self.face < 2 * self.dim:
            raise ValueError("face out of range")
        self.sizes_minus = tuple(int(s) for s in sizes_coarse)
        self.sizes_plus = tuple(int(s) for s in sizes_fine)
        axis = self.face % self.dim
        for d in range(self.dim):
            if d != axis and self.sizes_plus[d] % self.sizes_minus[d] != 0:
                raise ValueError("fine cells per tangential axis must be a multiple of the coarse ones")


class AdvectionFaceOperator:
    """y += the upwind fluxes of every interior face, computed once per face
    (AdvectionFaceOperator, advectionoperator.hpp:1105-1180).  Like GenDiL's,
    op(x, y) ACCUMULATES into y (read-write view, :1173-1179)."""

    def __init__(self, plan: Plan):
        self.plan = plan

    @property
    def num_dofs(self):
        return self.plan.num_dofs

    def __call__(self, x: torch.Tensor, y: torch.Tensor):
        xs, ys = _span(x, self.plan, "x"), _span(y, self.plan, "y")
        _check_vec(xs, self.plan.num_dofs, torch.float64, "x")
        _check_vec(ys, self.plan.num_dofs, torch.float64, "y")
        if _overlap(xs, ys):
            raise ValueError("x and y must not overlap")
        L.check(L.lib().gd_adv_face_apply_f64(self.plan.handle, xs.data_ptr(), ys.data_ptr(), _stream()),
                "gd_adv_face_apply_f64")
        return y

    Apply = __call__


class AdvectionNonconformingFaceOperator:
    """Fluxes across the interface of two meshes whose spaces may differ in
    order (AdvectionNonconformingFaceOperator, advectionoperator.hpp:1210-1396):
    + on the minus space's part of y, - on the plus space's, accumulated.  The
    two spaces live in one vector through their L2Restriction shifts."""

    def __init__(self, minus: Plan, plus: Plan, face: int, refined: bool = False):
        self.minus, self.plus, self.face = minus, plus, int(face)
        self._entry = "gd_adv_hmesh_apply_f64" if refined else "gd_adv_intermesh_apply_f64"

    def __call__(self, x: torch.Tensor, y: torch.Tensor):
        nm, np_ = self.minus.num_dofs, self.plus.num_dofs
        xm, xp = _span(x, self.minus, "x"), _span(x, self.plus, "x")
        ym, yp = _span(y, self.minus, "y"), _span(y, self.plus, "y")
        for t, n, name in ((xm, nm, "x"), (xp, np_, "x"), (ym, nm, "y"), (yp, np_, "y")):
            _check_vec(t, n, torch.float64, name)
        # the kernels read x while they add into y (the h-refined interface in
        # one launch per subface position), so no part of x may share storage
        # with a part of y
        if any(_overlap(a, b) for a in (xm, xp) for b in (ym, yp)):
            raise ValueError("x and y must not overlap")
        L.check(
            getattr(L.lib(), self._entry)(self.minus.handle, self.plus.handle, self.face, xm.data_ptr(),
                                          xp.data_ptr(), ym.data_ptr(), yp.data_ptr(), _stream()),
            self._entry,
        )
        return y

    Apply = __call__


class DiffusionOperator:
    """SIPDG diffusion operator (GenDiL's DiffusionOperator,
    diffusionoperator.hpp:396-477): op(x, y) overwrites y."""

    def __init__(self, plan: Plan):
        self.plan = plan

    @property
    def num_dofs(self):
        return self.plan.num_dofs

    def __call__(self, x: torch.Tensor, y: torch.Tensor):
        out = y
        x, y = _span(x, self.plan, "x"), _span(y, self.plan, "y")
        if x.dtype not in (torch.float64, torch.float32):
            raise TypeError("x must be float64 or float32")
        _check_vec(x, self.num_dofs, x.dtype, "x")
        _check_vec(y, self.num_dofs, x.dtype, "y")
        if _overlap(x, y):
            raise ValueError("x and y must not overlap")
        fn = "gd_diff_apply_f64" if x.dtype == torch.float64 else "gd_diff_apply_f32"
        L.check(getattr(L.lib(), fn)(self.plan.handle, x.data_ptr(), y.data_ptr(), _stream()), fn)
        return out

    Apply = __call__


def MakeDiffusionOperator(fes, ir, velocity, sigma: float, kappa: float) -> DiffusionOperator:
    """MakeDiffusionOperator(fes, int_rule, velocity, sigma, kappa)
    (diffusionoperator.hpp:519-537): `velocity` is the diffusivity mu(X) (a
    SeparableField or a constant), sigma the symmetry parameter (-1 SIPG,
    0 IIPG, +1 NIPG), kappa the penalty factor (GenDiL's tests use (p+1)^2)."""
    plan = Plan(fes, ir)
    plan.set_diffusion(velocity, sigma, kappa)
    return DiffusionOperator(plan)


def MakeAdvectionFaceOperator(*args):
    """MakeAdvectionFaceOperator(fes, face_mesh, ir, adv) or
    MakeAdvectionFaceOperator(fes_minus, fes_plus, face_mesh, ir, adv)
    (advectionoperator.hpp:1546-1607); no boundary data, as in GenDiL."""
    if len(args) == 4:
        fes, conn, ir, adv = args
        if not isinstance(conn, CartesianInteriorFaceConnectivity):
            raise TypeError("single-space face operators take MakeCartesianInteriorFaceConnectivity(sizes)")
        if tuple(conn.sizes) != tuple(fes.mesh.sizes):
            raise ValueError("face connectivity sizes differ from the mesh")
        plan = Plan(fes, ir)
        plan.set_beta(adv)
        return AdvectionFaceOperator(plan)
    if len(args) == 5:
        fm, fp, conn, ir, adv = args
        if not isinstance(conn, CartesianIntermeshFaceConnectivity):
            raise TypeError("two-space face operators take a CartesianIntermeshFaceConnectivity")
        if tuple(conn.sizes_minus) != tuple(fm.mesh.sizes) or tuple(conn.sizes_plus) != tuple(fp.mesh.sizes):
            raise ValueError("face connectivity sizes differ from the meshes")
        pm, pp = Plan(fm, ir), Plan(fp, ir)
        pm.set_beta(adv)
        return AdvectionNonconformingFaceOperator(
            pm, pp, conn.face, refined=isinstance(conn, NonconformingCartesianIntermeshFaceConnectivity))
    raise TypeError("MakeAdvectionFaceOperator takes 4 or 5 arguments")
