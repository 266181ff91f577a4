"""CPU oracle for the MI355X DG operator apply.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package, and only as the checker: the product path never calls it.

  liboracle.so      project-owned CPU restatement of GenDiL's element-centric
                    apply (gd_oracle.cpp; each function cites the reference).
  _ref/gendil_ref   GenDiL's own SerialKernelConfiguration operators compiled
                    from the reference headers (built only where they exist).

Parity status: gdo_apply pinned by golden vectors from _ref/gendil_ref
(tests/golden); gdo_apply_transpose pinned by dense transposes and the
dot-product identity; gdo_grad_beta pinned by central finite differences;
the p-adaptive apply pinned by GenDiL's own p-adaptivity (uniform-order blocks
coupled by its intermesh face operator, tests/golden/padapt*), by its uniform
special case and by conservation;
bicgstab pinned by GenDiL's own BiCGSTAB_no_alloc run through _ref/gendil_ref
(tests/golden/bicgstab3d_c2_p1q2_bc0.npz) and by dense solves; the face-centric,
intermesh and h-refined face operators (advection_faces, hmesh_operator) pinned
by GenDiL's own face operators through _ref/gendil_ref (tests/golden/face*,
imesh*, hmesh*), conservation and the split-mesh identity; the SIPDG
diffusion apply pinned by GenDiL's own DiffusionOperator through
_ref/gendil_ref (tests/golden/diff*), symmetry for sigma = -1 and the constant
null space on periodic meshes.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_BIN = HERE / "_ref" / "gendil_ref"

MAX_DIM = 6
MAX_TERMS = 8


class GdoSeparable(C.Structure):
    _fields_ = [
        ("rank", C.c_int32),
        ("scale", C.c_double * MAX_TERMS),
        ("offset", (C.c_int64 * MAX_DIM) * MAX_TERMS),
        ("data", C.POINTER(C.c_double)),
    ]


class GdoProblem(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("sizes", C.c_int64 * MAX_DIM),
        ("h", C.c_double * MAX_DIM),
        ("origin", C.c_double * MAX_DIM),
        ("periodic", C.c_int32 * MAX_DIM),
        ("order", C.c_int32),
        ("nq", C.c_int32),
        ("basis", C.c_int32),
        ("mass_coef", C.c_double),
        ("adv_coef", C.c_double),
        ("sigma", GdoSeparable),
        ("beta_kind", C.c_int32),
        ("beta", GdoSeparable * MAX_DIM),
        ("beta_cell", C.POINTER(C.c_double)),
        ("beta_face", C.POINTER(C.c_double) * MAX_DIM),
        ("beta_affine_A", (C.c_double * MAX_DIM) * MAX_DIM),
        ("beta_affine_c", C.c_double * MAX_DIM),
        ("bc_mode", C.c_int32),
        ("bc", GdoSeparable),
        ("elem_order", C.POINTER(C.c_int8)),
        ("elem_nq", C.POINTER(C.c_int8)),
        ("dof_offset", C.POINTER(C.c_int64)),
        ("num_threads", C.c_int32),
        ("plant", C.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = C.CDLL(str(LIB_PATH))
        P = C.POINTER(GdoProblem)
        pd = C.POINTER(C.c_double)
        h.gdo_apply.argtypes = [P, pd, pd]
        h.gdo_apply_transpose.argtypes = [P, pd, pd]
        h.gdo_apply_elements.argtypes = [P, pd, pd, C.POINTER(C.c_int64), C.c_int64]
        h.gdo_grad_beta.argtypes = [P, pd, pd, pd, C.POINTER(pd)]
        h.gdo_points.argtypes = [C.c_int32, C.c_int32, pd, pd]
        h.gdo_axis_points.argtypes = [P, C.c_int32, pd]
        h.gdo_num_dofs.argtypes = [P]
        h.gdo_diffusion_apply.argtypes = [P, C.POINTER(GdoSeparable), C.c_double, C.c_double, pd, pd]
        h.gdo_num_dofs.restype = C.c_int64
        _lib = h
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def points(basis: int, n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().gdo_points(basis, n, _dp(x), _dp(w))
    return x, w


class Problem:
    """Oracle view of a discretization + coefficients (numpy, host)."""

    def __init__(self, fes, ir, mass_coef=0.0, adv_coef=0.0, sigma=None, adv=None, bc=None, num_threads=0,
                 plant=0):
        from gendil_amd import fields as F
        from gendil_amd._lib import BC_INFLOW, BC_NONE

        self.fes = fes
        self.q = ir.num_points if ir is not None else fes.fe.order + 1
        m = fes.mesh
        p = GdoProblem()
        p.dim = m.dim
        for d in range(m.dim):
            p.sizes[d] = m.sizes[d]
            p.h[d] = m.h[d]
            p.origin[d] = m.origin[d]
            p.periodic[d] = int(m.periodic[d])
        p.order = fes.fe.order
        p.nq = self.q
        p.basis = fes.fe.basis
        p.mass_coef = mass_coef
        p.adv_coef = adv_coef
        p.num_threads = num_threads
        p.plant = plant  # a planted mistake (gd_oracle.h), for tests of the fixtures only
        self._keep = []
        if sigma is None:
            sigma = F.SeparableField.constant(1.0)
        elif isinstance(sigma, (int, float)):
            sigma = F.SeparableField.constant(float(sigma))
        p.sigma = self._sep(sigma)
        p.beta_kind = 0
        if isinstance(adv, F.VectorField):
            for d, comp in enumerate(adv.components):
                p.beta[d] = self._sep(comp)
        elif isinstance(adv, F.AffineField):
            p.beta_kind = 2
            A = np.asarray(adv.A, dtype=np.float64).reshape(m.dim, m.dim)
            c = np.asarray(adv.c, dtype=np.float64).reshape(m.dim)
            for i in range(m.dim):
                p.beta_affine_c[i] = c[i]
                for j in range(m.dim):
                    p.beta_affine_A[i][j] = A[i, j]
        elif adv is not None:  # pointwise: numpy (cell, [faces])
            cell, faces = adv
            cell = np.ascontiguousarray(cell, dtype=np.float64)
            faces = [np.ascontiguousarray(f, dtype=np.float64) for f in faces]
            self._keep += [cell, *faces]
            p.beta_kind = 1
            p.beta_cell = _dp(cell)
            for d, f in enumerate(faces):
                p.beta_face[d] = _dp(f)
        if bc is None:
            p.bc_mode = BC_NONE
        else:
            p.bc_mode = BC_INFLOW
            if isinstance(bc, (int, float)):
                bc = F.SeparableField.constant(float(bc))
            p.bc = self._sep(bc)
        if fes.padaptive:
            eo = np.ascontiguousarray(fes.elem_order, dtype=np.int8)
            eq = np.ascontiguousarray(fes.elem_quad, dtype=np.int8)
            off = np.ascontiguousarray(fes.dof_offsets(), dtype=np.int64)
            self._keep += [eo, eq, off]
            p.elem_order = eo.ctypes.data_as(C.POINTER(C.c_int8))
            p.elem_nq = eq.ctypes.data_as(C.POINTER(C.c_int8))
            p.dof_offset = off.ctypes.data_as(C.POINTER(C.c_int64))
        self.p = p
        self.num_dofs = int(lib().gdo_num_dofs(C.byref(p)))

    def _sep(self, f):
        s, data = f.tabulate(self.fes.mesh, self.q)
        self._keep.append(data)
        o = GdoSeparable()
        o.rank = s.rank
        for r in range(MAX_TERMS):
            o.scale[r] = s.scale[r]
            for d in range(MAX_DIM):
                o.offset[r][d] = s.offset[r][d]
        o.data = _dp(data)
        return o

    def apply(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.num_dofs)
        rc = lib().gdo_apply(C.byref(self.p), _dp(x), _dp(y))
        if rc:
            raise RuntimeError(f"gdo_apply failed ({rc})")
        return y

    def apply_elements(self, x: np.ndarray, elems) -> np.ndarray:
        """The forward apply at the DoFs of the listed elements only; returns
        their outputs concatenated in list order."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        el = np.ascontiguousarray(np.asarray(elems, dtype=np.int64).reshape(-1))
        y = np.zeros(self.num_dofs)
        rc = lib().gdo_apply_elements(C.byref(self.p), _dp(x), _dp(y), el.ctypes.data_as(C.POINTER(C.c_int64)),
                                      el.size)
        if rc:
            raise RuntimeError(f"gdo_apply_elements failed ({rc})")
        return y[self.element_dofs(el)]

    def element_dofs(self, elems) -> np.ndarray:
        """Indices of the DoFs of the listed elements, concatenated in list order."""
        off = self.fes.dof_offsets()
        el = np.asarray(elems, dtype=np.int64)
        return np.concatenate([np.arange(off[e], off[e + 1]) for e in el]) if el.size else np.zeros(0, np.int64)

    def apply_transpose(self, r: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.num_dofs)
        rc = lib().gdo_apply_transpose(C.byref(self.p), _dp(r), _dp(z))
        if rc:
            raise RuntimeError(f"gdo_apply_transpose failed ({rc})")
        return z

    def diffusion(self, x: np.ndarray, mu, sigma: float, kappa: float) -> np.ndarray:
        """SIPDG diffusion apply (gdo_diffusion_apply) on this problem's mesh,
        basis and rule, diffusivity `mu` (a separable field or a constant)."""
        from gendil_amd import fields as F

        if isinstance(mu, (int, float)):
            mu = F.SeparableField.constant(float(mu))
        m = self._sep(mu)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.num_dofs)
        rc = lib().gdo_diffusion_apply(C.byref(self.p), C.byref(m), sigma, kappa, _dp(x), _dp(y))
        if rc:
            raise RuntimeError(f"gdo_diffusion_apply failed ({rc})")
        return y

    def grad_beta(self, u, r, face_counts):
        m = self.fes.mesh
        D = m.dim
        q = self.q
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        dcell = np.zeros(m.num_cells * D * q ** D)
        dfaces = [np.zeros(fc * q ** (D - 1)) for fc in face_counts]
        arr = (C.POINTER(C.c_double) * D)(*[_dp(f) for f in dfaces])
        rc = lib().gdo_grad_beta(C.byref(self.p), _dp(u), _dp(r), _dp(dcell), arr)
        if rc:
            raise RuntimeError(f"gdo_grad_beta failed ({rc})")
        return dcell, dfaces


def ref_available() -> bool:
    return REF_BIN.exists()


def run_ref(args, timeout=600):
    """Run GenDiL's own apply via the _ref harness (see ref_harness.cpp)."""
    res = subprocess.run([str(REF_BIN), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    if res.returncode != 0:
        raise RuntimeError(f"gendil_ref failed: {res.stderr}")
    return res.stdout


def ref_fixture(op, D, P, Q, basis, n, periodic, coef, bc, seed, path):
    run_ref(["fixture", op, D, P, Q, basis, n, int(periodic), coef, bc, seed, path])
    d = np.fromfile(path, dtype=np.float64)
    nd = int(d[:1].view(np.int64)[0])
    return d[1 : 1 + nd].copy(), d[1 + nd : 1 + 2 * nd].copy()


def ref_bench(op, D, P, Q, basis, n, periodic, coef, bc, seed, pairs, env=None, timeout=900):
    e = dict(os.environ)
    if env:
        e.update(env)
    res = subprocess.run(
        [str(REF_BIN), "bench", op, str(D), str(P), str(Q), basis, str(n), str(int(periodic)), coef, bc, str(seed),
         str(pairs)],
        capture_output=True, text=True, timeout=timeout, env=e,
    )
    if res.returncode != 0:
        raise RuntimeError(f"gendil_ref bench failed: {res.stderr}")
    return json.loads(res.stdout.strip().splitlines()[-1])


def ref_bicgstab(D, P, Q, basis, n, coef, bc, seed, path):
    """GenDiL's BiCGSTAB_no_alloc (bicgstab.hpp:44) on M + 0.05 A, b =
    FillRandom(seed), x0 = 0, tol 1e-10 (ref_common.hpp, solve_bicgstab).
    Returns (b, x, iterations, success, relative residual)."""
    return ref_solve("bicgstab", D, P, Q, basis, n, coef, bc, seed, path)


def ref_solve(mode, D, P, Q, basis, n, coef, bc, seed, path):
    """GenDiL's own solver `mode` through _ref/gendil_ref (ref_common.hpp):
    bicgstab (BiCGSTAB_no_alloc) or gmres (GMRES_no_alloc, restart 30) on
    M + 0.05 A(coef, inflow bc), cg (ConjugateGradient) on M(coef); b =
    FillRandom(seed), x0 = 0, tol 1e-10, at most 500 iterations.
    Returns (b, x, iterations, success, relative residual)."""
    run_ref([mode, "madv", D, P, Q, basis, n, 0, coef, bc, seed, path])
    d = np.fromfile(path, dtype=np.float64)
    nd = int(d[:1].view(np.int64)[0])
    b = d[1 : 1 + nd].copy()
    x = d[1 + nd : 1 + 2 * nd].copy()
    it, ok, rel = d[1 + 2 * nd : 4 + 2 * nd]
    return b, x, int(it), bool(ok), float(rel)


def bicgstab(apply, b, x, max_iter, tol):
    """Unpreconditioned BiCGSTAB restated from BiCGSTAB_no_alloc
    (bicgstab.hpp:44-157): same recurrences, early exit on ||s||, <t,t>
    fallback and 1e-30 breakdown guards.  x is updated in place; returns
    (success, iterations, relative residual)."""
    tiny = 1e-30                                   # :58
    t = apply(x)                                   # :61-64
    r = b - t
    r_hat = r.copy()                               # :66
    norm_rhs = np.sqrt(b @ b)                      # :68
    if norm_rhs == 0.0:
        x[:] = 0.0
        return True, 0, 0.0
    rho_prev = alpha = omega = 1.0                 # :75-77
    p = np.zeros_like(b)
    v = np.zeros_like(b)
    rel_res = np.sqrt(r @ r) / norm_rhs            # :82-83
    if rel_res < tol:
        return True, 0, rel_res
    it = 0
    while it < max_iter:                           # :90
        rho = r_hat @ r
        if abs(rho) < tiny:
            break
        if it == 0:
            p = r.copy()
        else:
            beta = (rho / rho_prev) * (alpha / omega)   # :101-104
            p = r + beta * (p - omega * v)
        v = apply(p)                               # :107
        rv = r_hat @ v
        if abs(rv) < tiny:
            break
        alpha = rho / rv                           # :111
        s = r - alpha * v                          # :113-114
        rel_s = np.sqrt(s @ s) / norm_rhs
        if rel_s < tol:                            # :117-124
            x += alpha * p
            return True, it + 1, rel_s
        t = apply(s)                               # :126
        ts, tt = t @ s, t @ t
        if abs(tt) < tiny:                         # :130-138
            x += alpha * p
            r = s
            rel_fb = np.sqrt(r @ r) / norm_rhs
            return rel_fb < tol, it + 1, rel_fb
        omega = ts / tt                            # :140
        if abs(omega) < tiny:
            break
        x += alpha * p + omega * s                 # :143-144
        r = s - omega * t                          # :146-147
        rel_res = np.sqrt(r @ r) / norm_rhs
        if rel_res < tol:
            return True, it + 1, rel_res
        rho_prev = rho
        it += 1
    return False, it, rel_res


# ---------------------------------------------------------------------------
# Face-centric upwind advection (numpy restatement, small meshes only)
# ---------------------------------------------------------------------------

def lagrange_matrix(nodes, x):
    """L[k, i] = l_i(x_k) for the Lagrange polynomials on `nodes`, product
    form (lagrangeshapefunctions.hpp:37-86)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    L = np.ones((x.size, nodes.size))
    for i in range(nodes.size):
        for j in range(nodes.size):
            if j != i:
                L[:, i] *= (x - nodes[j]) / (nodes[i] - nodes[j])
    return L


def interior_face_pairs(sizes, axis):
    """(minus, plus) element pairs of CartesianInteriorLocalFaceConnectivity
    <D, axis> (cartesianglobalfaceconnectivity.hpp:64-87): faces enumerated
    over the grid with one fewer cell along `axis`, first index fastest; the
    plus cell is the lower one, the minus cell the upper one (reference normal
    -e_axis of the minus cell)."""
    sizes = list(sizes)
    fs = sizes.copy()
    fs[axis] -= 1
    pairs = []
    for f in range(int(np.prod(fs))):
        idx = list(np.unravel_index(f, fs, order="F"))
        plus = int(np.ravel_multi_index(idx, sizes, order="F"))
        idx[axis] += 1
        minus = int(np.ravel_multi_index(idx, sizes, order="F"))
        pairs.append((minus, plus))
    return pairs


def intermesh_face_pairs(sizes_m, sizes_p, face):
    """(minus, plus) pairs of CartesianIntermeshFaceConnectivity<D, face>
    (cartesianintermeshfaceconnectivity.hpp:150-178): tangential positions
    first index fastest; a +axis face puts the minus cells on the last slice
    and the plus cells on the first, a -axis face the other way round."""
    D = len(sizes_m)
    axis, sign = face % D, (1 if face >= D else -1)
    fs = list(sizes_m)
    fs[axis] = 1
    pairs = []
    for f in range(int(np.prod(fs))):
        t = list(np.unravel_index(f, fs, order="F"))
        im, ip = t.copy(), t.copy()
        im[axis] = sizes_m[axis] - 1 if sign > 0 else 0
        ip[axis] = 0 if sign > 0 else sizes_p[axis] - 1
        pairs.append((int(np.ravel_multi_index(im, sizes_m, order="F")),
                      int(np.ravel_multi_index(ip, sizes_p, order="F"))))
    return pairs


def advection_faces(axis, sign, q, minus, plus, pairs, beta, u_m, u_p, y_m, y_p):
    """Accumulate the upwind face fluxes of GenDiL's AdvectionLocalFaceOperator
    / AdvectionNonconformingLocalFaceOperator (advectionoperator.hpp:478-548,
    596-676) for each (minus, plus) element pair into y_m (+) and y_p (-).

    minus / plus: dicts with `sizes`, `h`, `origin` (per axis), `n` (nodes
    per axis) and `basis` (0 GLL, 1 GL nodes).  The face carries the minus
    side's reference normal sign * e_axis.  Both sides are interpolated to the
    same q Gauss points per tangential axis (the shared rule, :1210-1229);
    the normal direction uses the side's basis at its face (ZeroPoint /
    OnePoint, faceintegrationrules.hpp:29-56).  Geometry, weight and beta come
    from the minus cell: X = origin + h (sub + xi), detJ = prod h, physical
    normal = sign e_axis / h_axis (computephysicalnormal.hpp:27-35), flux =
    w detJ (dot > 0 ? dot u- : dot u+) (:533-540).  beta maps an array
    [..., D] of points to [..., D]."""
    D = len(minus["sizes"])
    xq, wq = points(1, q)
    tdims = [d for d in range(D) if d != axis]
    sides = []
    for s, xi in ((minus, 1.0 if sign > 0 else 0.0), (plus, 0.0 if sign > 0 else 1.0)):
        nodes, _ = points(s["basis"], s["n"])
        sides.append((s["n"], lagrange_matrix(nodes, xq), lagrange_matrix(nodes, [xi])[0], xi))
    (nm, Bm, bm, xim), (npl, Bp, bp, _) = sides
    hm = np.asarray(minus["h"], dtype=np.float64)
    om = np.asarray(minus["origin"], dtype=np.float64)
    detJ = float(np.prod(hm))
    # face points (tangential axes in increasing order, first fastest -> numpy axis order)
    grids = np.meshgrid(*([xq] * (D - 1)), indexing="ij")
    wts = np.ones([q] * (D - 1))
    for k in range(D - 1):
        shape = [1] * (D - 1)
        shape[k] = q
        wts = wts * wq.reshape(shape)

    def trace(u, n, B, b):
        t = np.tensordot(u.reshape([n] * D, order="F"), b, axes=([axis], [0]))
        for k in range(D - 1):
            t = np.moveaxis(np.tensordot(B, t, axes=([1], [k])), 0, k)
        return t

    def spread(F, n, B, b):
        t = F
        for k in range(D - 1):
            t = np.moveaxis(np.tensordot(B.T, t, axes=([1], [k])), 0, k)
        full = np.multiply.outer(t, b)                 # normal axis last
        full = np.moveaxis(full, D - 1, axis)
        return full.reshape(-1, order="F")

    for em, ep in pairs:
        sub = np.unravel_index(em, minus["sizes"], order="F")
        X = np.zeros([q] * (D - 1) + [D])
        for k, d in enumerate(tdims):
            X[..., d] = om[d] + hm[d] * (sub[d] + grids[k])
        X[..., axis] = om[axis] + hm[axis] * (sub[axis] + xim)
        dot = beta(X)[..., axis] * (sign / hm[axis])
        um_q = trace(u_m[em * nm**D:(em + 1) * nm**D], nm, Bm, bm)
        up_q = trace(u_p[ep * npl**D:(ep + 1) * npl**D], npl, Bp, bp)
        F = wts * detJ * np.where(dot > 0.0, dot * um_q, dot * up_q)
        y_m[em * nm**D:(em + 1) * nm**D] += spread(F, nm, Bm, bm)
        y_p[ep * npl**D:(ep + 1) * npl**D] -= spread(F, npl, Bp, bp)


def face_operator(sizes, h, origin, n, basis, q, beta, u, y):
    """GenDiL's AdvectionFaceOperator over the Cartesian interior faces of one
    space (advectionoperator.hpp:1105-1180): every interior face of every
    axis once, accumulated into y."""
    side = dict(sizes=list(sizes), h=list(h), origin=list(origin), n=n, basis=basis)
    for axis in range(len(sizes)):
        advection_faces(axis, -1, q, side, side, interior_face_pairs(sizes, axis), beta, u, u, y, y)


def intermesh_operator(face, q, minus, plus, beta, u_m, u_p, y_m, y_p):
    """GenDiL's AdvectionNonconformingFaceOperator over
    CartesianIntermeshFaceConnectivity<D, face> (advectionoperator.hpp:
    1210-1396), accumulated into y_m and y_p."""
    D = len(minus["sizes"])
    pairs = intermesh_face_pairs(minus["sizes"], plus["sizes"], face)
    advection_faces(face % D, 1 if face >= D else -1, q, minus, plus, pairs, beta, u_m, u_p, y_m, y_p)


def hmesh_operator(face, q, coarse, fine, beta, u_c, u_f, y_c, y_f):
    """GenDiL's nonconforming face operator over
    NonconformingCartesianIntermeshFaceConnectivity<D, face>
    (nonconformingcartesianintermeshfaceconnectivity.hpp:115-243, operator
    advectionoperator.hpp:596-676): face `face` of the coarse (minus) mesh
    glued to a mesh refined by an integer ratio r_d = fine/coarse cells along
    each tangential axis.  One face per fine cell on the interface (fine
    tangential index first fastest); the coarse cell is fine index // r_d and
    the subface offset o_d = (fine index % r_d) / r_d, size s_d = 1 / r_d.
    The coarse side is interpolated at its mapped points o + s xi and its
    weights carry s (the nonconforming face DofToQuad, facetquaddata.hpp:
    86-155, used by GetWeight at advectionoperator.hpp:651), while the
    geometry and field are evaluated at the coarse face's own, unmapped rule
    points (ComputePhysicalCoordinatesAndJacobian with the mesh face data,
    :646); pinned by GenDiL's fixtures, which differ by 10 % for a
    non-constant field if X is mapped.  u_c / y_c, u_f / y_f: the two
    spaces' parts of the vectors."""
    D = len(coarse["sizes"])
    axis, sign = face % D, (1 if face >= D else -1)
    sc, sf = list(coarse["sizes"]), list(fine["sizes"])
    r = [1 if d == axis else sf[d] // sc[d] for d in range(D)]
    hc = np.asarray(coarse["h"], dtype=np.float64)
    oc = np.asarray(coarse["origin"], dtype=np.float64)
    xq, wq = points(1, q)
    nodes_c, _ = points(coarse["basis"], coarse["n"])
    nodes_f, _ = points(fine["basis"], fine["n"])
    nc, nf = coarse["n"], fine["n"]
    xi_m = 1.0 if sign > 0 else 0.0
    bm = lagrange_matrix(nodes_c, [xi_m])[0]
    bp = lagrange_matrix(nodes_f, [1.0 - xi_m])[0]
    Bf = lagrange_matrix(nodes_f, xq)
    tdims = [d for d in range(D) if d != axis]
    grids = np.meshgrid(*([xq] * (D - 1)), indexing="ij")

    def trace(u, n, b, Bs):
        t = np.tensordot(u.reshape([n] * D, order="F"), b, axes=([axis], [0]))
        for k in range(D - 1):
            t = np.moveaxis(np.tensordot(Bs[k], t, axes=([1], [k])), 0, k)
        return t

    def spread(F, b, Bs):
        t = F
        for k in range(D - 1):
            t = np.moveaxis(np.tensordot(Bs[k].T, t, axes=([1], [k])), 0, k)
        return np.moveaxis(np.multiply.outer(t, b), D - 1, axis).reshape(-1, order="F")

    fs = list(sf)
    fs[axis] = 1
    for f in range(int(np.prod(fs))):
        ft = list(np.unravel_index(f, fs, order="F"))
        im = [ft[d] // r[d] for d in range(D)]
        im[axis] = sc[axis] - 1 if sign > 0 else 0
        ip = list(ft)
        ip[axis] = 0 if sign > 0 else sf[axis] - 1
        em = int(np.ravel_multi_index(im, sc, order="F"))
        ep = int(np.ravel_multi_index(ip, sf, order="F"))
        Bm = [lagrange_matrix(nodes_c, (ft[d] % r[d] + xq) / r[d]) for d in tdims]
        X = np.zeros([q] * (D - 1) + [D])
        wts = np.ones([q] * (D - 1))
        for k, d in enumerate(tdims):
            X[..., d] = oc[d] + hc[d] * (im[d] + grids[k])
            shape = [1] * (D - 1)
            shape[k] = q
            wts = wts * wq.reshape(shape) / r[d]
        X[..., axis] = oc[axis] + hc[axis] * (im[axis] + xi_m)
        dot = beta(X)[..., axis] * (sign / hc[axis])
        um = trace(u_c[em * nc**D:(em + 1) * nc**D], nc, bm, Bm)
        up = trace(u_f[ep * nf**D:(ep + 1) * nf**D], nf, bp, [Bf] * (D - 1))
        F = wts * float(np.prod(hc)) * np.where(dot > 0.0, dot * um, dot * up)
        y_c[em * nc**D:(em + 1) * nc**D] += spread(F, bm, Bm)
        y_f[ep * nf**D:(ep + 1) * nf**D] -= spread(F, bp, [Bf] * (D - 1))
