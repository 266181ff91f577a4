"""ctypes binding of the C-ABI declared in include/gendil_amd.h.

The shared library is built in-tree (gendil_amd/libgendil_amd.so).  There is
no fallback: if it is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

MAX_DIM = 6
MAX_TERMS = 8

GD_OK = 0
GD_ERR_INVALID = 1
GD_ERR_UNSUPPORTED = 2
GD_ERR_HIP = 3
GD_ERR_ALLOC = 4
GD_ERR_NOT_CONVERGED = 5

BASIS_GLL = 0
BASIS_GL = 1
BETA_SEPARABLE = 0
BETA_POINTWISE = 1
BETA_AFFINE = 2
BC_NONE = 0
BC_INFLOW = 1


class GdSeparable(C.Structure):
    _fields_ = [
        ("rank", C.c_int32),
        ("scale", C.c_double * MAX_TERMS),
        ("offset", (C.c_int64 * MAX_DIM) * MAX_TERMS),
        ("data", C.POINTER(C.c_double)),
        ("data_len", C.c_int64),
    ]


class GdDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("sizes", C.c_int64 * MAX_DIM),
        ("h", C.c_double * MAX_DIM),
        ("origin", C.c_double * MAX_DIM),
        ("periodic", C.c_int32 * MAX_DIM),
        ("order", C.c_int32),
        ("num_quad", C.c_int32),
        ("basis", C.c_int32),
        ("elem_order", C.POINTER(C.c_int8)),
        ("elem_quad", C.POINTER(C.c_int8)),
    ]


class GdSolveParams(C.Structure):
    _fields_ = [
        ("restart", C.c_int32),
        ("max_iters", C.c_int32),
        ("rel_tol", C.c_double),
        ("max_refinements", C.c_int32),
        ("inner_rel_tol", C.c_double),
    ]


class GdSolveResult(C.Structure):
    _fields_ = [
        ("success", C.c_int32),
        ("iterations", C.c_int32),
        ("refinements", C.c_int32),
        ("final_rel_residual", C.c_double),
        ("seconds", C.c_double),
    ]


_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_d = C.c_double
_f = C.c_float
_pd = C.POINTER(C.c_double)
_ppd = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every function declared in include/gendil_amd.h
SIGNATURES = {
    "gd_plan_create": (C.c_int, [C.POINTER(GdDesc), C.POINTER(C.c_void_p)]),
    "gd_plan_destroy": (C.c_int, [_p]),
    "gd_last_error": (C.c_char_p, []),
    "gd_version": (C.c_int, []),
    "gd_plan_num_dofs": (_i64, [_p]),
    "gd_plan_num_elements": (_i64, [_p]),
    "gd_plan_dof_offsets": (_p, [_p]),
    "gd_plan_axis_points": (C.c_int, [_p, _i32, _pd]),
    "gd_rule_points": (C.c_int, [_i32, _i32, _pd, _pd]),
    "gd_plan_set_sigma": (C.c_int, [_p, C.POINTER(GdSeparable)]),
    "gd_plan_set_beta_separable": (C.c_int, [_p, C.POINTER(GdSeparable)]),
    "gd_plan_set_beta_affine": (C.c_int, [_p, _pd, _pd]),
    "gd_plan_set_beta_pointwise_f64": (C.c_int, [_p, _p, _ppd]),
    "gd_plan_set_bc": (C.c_int, [_p, _i32, C.POINTER(GdSeparable)]),
    "gd_plan_set_diffusion": (C.c_int, [_p, C.POINTER(GdSeparable), C.c_double, C.c_double]),
    "gd_apply_f64": (C.c_int, [_p, _d, _d, _p, _p, _p]),
    "gd_apply_f32": (C.c_int, [_p, _f, _f, _p, _p, _p]),
    "gd_mass_apply_f64": (C.c_int, [_p, _p, _p, _p]),
    "gd_mass_apply_f32": (C.c_int, [_p, _p, _p, _p]),
    "gd_adv_apply_f64": (C.c_int, [_p, _p, _p, _p]),
    "gd_adv_apply_f32": (C.c_int, [_p, _p, _p, _p]),
    "gd_diff_apply_f64": (C.c_int, [_p, _p, _p, _p]),
    "gd_diff_apply_f32": (C.c_int, [_p, _p, _p, _p]),
    "gd_apply_T_f64": (C.c_int, [_p, _d, _d, _p, _p, _p]),
    "gd_adv_apply_T_f64": (C.c_int, [_p, _p, _p, _p]),
    "gd_adv_grad_beta_f64": (C.c_int, [_p, _d, _p, _p, _p, _ppd, _p]),
    "gd_adv_face_apply_f64": (C.c_int, [_p, _p, _p, _p]),
    "gd_adv_intermesh_apply_f64": (C.c_int, [_p, _p, C.c_int32, _p, _p, _p, _p, _p]),
    "gd_adv_hmesh_apply_f64": (C.c_int, [_p, _p, C.c_int32, _p, _p, _p, _p, _p]),
    "gd_plan_autotune": (C.c_int, [_p, _i32, _d, _d, _i32, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "gd_plan_set_variant": (C.c_int, [_p, _i32, _i32]),
    "gd_plan_query_variant": (C.c_int, [_p, _i32, _d, _d, C.POINTER(C.c_int32)]),
    "gd_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "gd_free": (C.c_int, [_p]),
    "gd_memcpy_h2d": (C.c_int, [_p, _p, C.c_size_t]),
    "gd_memcpy_d2h": (C.c_int, [_p, _p, C.c_size_t]),
    "gd_memset": (C.c_int, [_p, C.c_int, C.c_size_t, _p]),
    "gd_stream_sync": (C.c_int, [_p]),
    "gd_stream_query": (C.c_int, [_p, C.POINTER(C.c_int32)]),
    "gd_device_sync": (C.c_int, []),
    "gd_multidot_f64": (C.c_int, [_i64, _i32, _p, _i64, _p, _p, _p]),
    "gd_multidot_f64_noalloc": (C.c_int, [_i64, _i32, _p, _i64, _p, _p, _p]),
    "gd_multidot_f32": (C.c_int, [_i64, _i32, _p, _i64, _p, _p, _p]),
    "gd_multiaxpy_f64": (C.c_int, [_i64, _i32, _p, _p, _i64, _p, _p]),
    "gd_multiaxpy_f32": (C.c_int, [_i64, _i32, _p, _p, _i64, _p, _p]),
    "gd_axpby_f64": (C.c_int, [_i64, _d, _p, _d, _p, _p]),
    "gd_axpby_f32": (C.c_int, [_i64, _f, _p, _f, _p, _p]),
    "gd_convert_f64_to_f32": (C.c_int, [_i64, _d, _p, _p, _p]),
    "gd_convert_f32_to_f64": (C.c_int, [_i64, _d, _p, _p, _p]),
    "gd_add_f32_to_f64": (C.c_int, [_i64, _d, _p, _p, _p]),
    "gd_gmres_workspace_bytes": (C.c_size_t, []),
    "gd_gmres_begin": (C.c_int, [_p, _d, _d, _p]),
    "gd_arnoldi_step_f64": (C.c_int, [_i64, _i32, _i32, _p, _i64, _p, _p, _p]),
    "gd_arnoldi_step_f32": (C.c_int, [_i64, _i32, _i32, _p, _i64, _p, _p, _p]),
    "gd_gmres_update_f64": (C.c_int, [_i64, _i32, _i32, _p, _i64, _p, _p, _p]),
    "gd_gmres_update_f32": (C.c_int, [_i64, _i32, _i32, _p, _i64, _p, _p, _p]),
    "gd_host_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "gd_host_free": (C.c_int, [_p]),
    "gd_event_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "gd_event_destroy": (C.c_int, [_p]),
    "gd_event_record": (C.c_int, [_p, _p]),
    "gd_event_sync": (C.c_int, [_p]),
    "gd_gmres_f64": (C.c_int, [_p, _d, _d, _p, _p, C.POINTER(GdSolveParams), C.POINTER(GdSolveResult), _p]),
    "gd_gmres_mixed": (C.c_int, [_p, _d, _d, _p, _p, C.POINTER(GdSolveParams), C.POINTER(GdSolveResult), _p]),
    "gd_cg_f64": (C.c_int, [_p, _d, _d, _p, _p, C.POINTER(GdSolveParams), C.POINTER(GdSolveResult), _p]),
    "gd_cg_mixed": (C.c_int, [_p, _d, _d, _p, _p, C.POINTER(GdSolveParams), C.POINTER(GdSolveResult), _p]),
    "gd_bicgstab_f64": (C.c_int, [_p, _d, _d, _p, _p, C.POINTER(GdSolveParams), C.POINTER(GdSolveResult), _p]),
    "gd_ksolve_workspace_bytes": (C.c_size_t, []),
    "gd_cg_begin_f64": (C.c_int, [_i64, _p, _d, _d, _p, _p]),
    "gd_cg_step_f64": (C.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p]),
    "gd_cg_begin_f32": (C.c_int, [_i64, _p, _d, _d, _p, _p]),
    "gd_cg_step_f32": (C.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p]),
    "gd_bicgstab_begin_f64": (C.c_int, [_i64, _p, _p, _d, _d, _p, _p]),
    "gd_bicgstab_p_f64": (C.c_int, [_i64, _p, _p, _p, _p, _p]),
    "gd_bicgstab_s_f64": (C.c_int, [_i64, _p, _p, _p, _p, _p, _p]),
    "gd_bicgstab_xr_f64": (C.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "gd_ksolve_status": (C.c_int, [_p, _p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
}


class GdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[gd status {code}] {msg}")
        self.code = code


LIB_PATH = Path(__file__).resolve().parent / "libgendil_amd.so"
_lib = None


def lib():
    """Load libgendil_amd.so (built in-tree by gendil_amd.build / __graft_entry__.build)."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("GENDIL_AMD_LIB", LIB_PATH))
        if not path.exists():
            raise RuntimeError(
                f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            )
        handle = C.CDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != GD_OK:
        msg = lib().gd_last_error().decode(errors="replace")
        raise GdError(rc, f"{what}: {msg}" if what else msg)


def ptr(t) -> int:
    """Raw device pointer of a torch tensor (or None -> 0)."""
    if t is None:
        return None
    return t.data_ptr()
