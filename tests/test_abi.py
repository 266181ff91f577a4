"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point include/gendil_amd.h declares, and its host-only helpers behave."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import gendil_amd as G
from gendil_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "gendil_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gd_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_bound_and_exported():
    names = declared_symbols()
    assert len(names) >= 40
    lib = L.lib()
    for n in names:
        assert hasattr(lib, n), n
        assert n in L.SIGNATURES, f"{n} missing from the ctypes signature table"
    assert set(L.SIGNATURES) == set(names)


def test_version_and_error_string():
    lib = L.lib()
    assert lib.gd_version() == 1
    rc = lib.gd_rule_points(7, 3, None, None)
    assert rc == L.GD_ERR_INVALID
    assert b"gd_rule_points" in lib.gd_last_error()


def test_plan_create_rejects_bad_descriptors_without_touching_the_gpu():
    lib = L.lib()
    d = L.GdDesc()
    d.dim = 7
    h = C.c_void_p()
    assert lib.gd_plan_create(C.byref(d), C.byref(h)) == L.GD_ERR_INVALID
    d.dim = 3
    for i in range(3):
        d.sizes[i] = 0
        d.h[i] = 1.0
    assert lib.gd_plan_create(C.byref(d), C.byref(h)) == L.GD_ERR_INVALID
    assert not h.value


def test_axis_points_layout():
    mesh = G.CartesianMesh([3], [0.25], [1.0])
    xs = G.axis_points(mesh, 2, 0).reshape(3, 4)
    xq, _ = G.rule_points(1, 2)
    for e in range(3):
        assert np.allclose(xs[e, :2], 1.0 + 0.25 * e + 0.25 * xq)
        assert xs[e, 2] == 1.0 + 0.25 * e and xs[e, 3] == (1.0 + 0.25 * e) + 0.25


def test_separable_tabulation_offsets():
    mesh = G.CartesianMesh.unit(3, 4)
    f = G.SeparableField([(2.0, {1: lambda y: y}), (1.0, {})])
    s, data = f.tabulate(mesh, 3)
    assert s.rank == 2 and s.scale[0] == 2.0
    assert s.offset[0][1] == 0 and s.offset[0][0] == -1 and s.offset[1][1] == -1
    assert data.size == 4 * 5


def test_product_mesh_is_the_concatenated_cartesian_mesh():
    # MakeCartesianProductMesh(3D, 2D, 1D) numbers and sizes cells like the
    # Cartesian mesh over the concatenated axes (productmesh.hpp:91-94)
    import gendil_amd as G

    m = G.MakeCartesianProductMesh(G.Cartesian3DMesh(1.0, 2, 3, 4), G.Cartesian2DMesh(1.2345, 5, 6),
                                   G.Cartesian1DMesh(0.5, 7, origin=-1.0, periodic=True))
    assert m.sizes == (2, 3, 4, 5, 6, 7)
    assert m.h == (1.0, 1.0, 1.0, 1.2345, 1.2345, 0.5)
    assert m.origin == (0.0, 0.0, 0.0, 0.0, 0.0, -1.0)
    assert m.periodic == (False,) * 5 + (True,)
    assert m.num_cells == 2 * 3 * 4 * 5 * 6 * 7
    # cell (head, tail) = head + cells(head) * tail, head x-fastest
    head, tail = (1, 2, 3), (4, 5, 6)
    h_idx = head[0] + 2 * (head[1] + 3 * head[2])
    t_idx = tail[0] + 5 * (tail[1] + 6 * tail[2])
    e = h_idx + 24 * t_idx
    sub = []
    r = e
    for n in m.sizes:
        sub.append(r % n)
        r //= n
    assert tuple(sub) == head + tail
    with pytest.raises(ValueError):
        G.MakeCartesianProductMesh()


def test_pointwise_beta_sizes_are_checked_before_any_device_call():
    # a wrong cell or face array size is refused on the host: the kernels
    # would otherwise read (and grad_beta write) out of bounds
    import torch

    from gendil_amd.operators import Plan

    mesh = G.CartesianMesh.unit(3, 2, False)
    fes = G.MakeFiniteElementSpace(mesh, G.MakeLobattoFiniteElement(1))
    plan = Plan.__new__(Plan)  # no gd_plan behind it: nothing may reach the C-ABI
    plan.fes, plan.q = fes, 2
    E, q = mesh.num_cells, 2
    faces = [torch.zeros(n * q * q, dtype=torch.float64) for n in G.face_counts(mesh)]
    cell = torch.zeros(E * 3 * q ** 3, dtype=torch.float64)
    with pytest.raises(ValueError, match="cell array"):
        plan.set_beta_pointwise(torch.zeros(E * 3 * 27, dtype=torch.float64), faces)
    with pytest.raises(ValueError, match="face array 1"):
        plan.set_beta_pointwise(cell, [faces[0], torch.zeros(faces[1].numel() + 4, dtype=torch.float64), faces[2]])
    with pytest.raises(ValueError, match="one face array per axis"):
        plan.set_beta_pointwise(cell, faces[:2])
    # right sizes, but host tensors: refused before the C-ABI as well
    with pytest.raises(ValueError, match="device tensors"):
        plan.set_beta_pointwise(cell, faces)


def test_face_operator_wrappers_reject_overlapping_storage():
    from gendil_amd.operators import _overlap
    import torch

    v = torch.zeros(100, dtype=torch.float64)
    assert _overlap(v[:50], v[49:])
    assert not _overlap(v[:50], v[50:])
    assert _overlap(v, v[10:20])
