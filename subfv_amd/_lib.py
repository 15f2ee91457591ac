"""ctypes binding of libsubfv_hip.so (declarations in include/subfv.h).

There is no CPU fallback: if the library is missing or cannot be loaded the
import of the device API fails with an error naming the file.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsubfv_hip.so")
MNBC = 10

OK = 0
E_INVALID = 1
E_DEVICE = 2
E_BAD_OMEGA = 3
E_NEG_STAR = 4
E_MESH = 5
E_BAD_BC = 6
E_NOT_CONVERGED = 7

SCHEMES = {
    "multi_point": 1,
    "multi_point_iso": 2,
    "three_wave": 3,
    "two_wave": 4,
    "modified_three_wave": 5,
    "multi_point_pressure": 6,
    "multi_point_pressure_ph": 8,
}
BCS = {
    "wall": 1,
    "": 1,
    "adherence_wall": 2,
    "freestream": 3,
    "outflowsupersonic": 4,
    "inflow_pond": 5,
    "inout_double_mach": 6,
    "double_mach_bottom": 7,
    "potential_flow_2d": 8,
    "potential_flow_3d": 9,
}


class Config(C.Structure):
    """subfv_config: the namelist entries the path reads."""

    _fields_ = [
        ("scheme_id", C.c_int32),
        ("bc_style", C.c_int32),
        ("boundary_2d", C.c_int32),
        ("exclude_bound_vert", C.c_int32),
        ("local_time_step", C.c_int32),
        ("second_order", C.c_int32),
        ("method", C.c_int32),
        ("n_bc", C.c_int32),
        ("bc_euler_id", C.c_int32 * MNBC),
        ("bc_val", (C.c_double * 5) * MNBC),
        ("cfl", C.c_double),
        ("tmax", C.c_double),
    ]


def make_config(scheme="multi_point", bc_style=2, boundary_2d=True, exclude_bound_vert=False,
                local_time_step=False, second_order=False, method=4, bcs=(), cfl=0.4,
                tmax=1e30) -> Config:
    c = Config()
    c.scheme_id = SCHEMES[scheme]
    c.bc_style = bc_style
    c.boundary_2d = int(boundary_2d)
    c.exclude_bound_vert = int(exclude_bound_vert)
    c.local_time_step = int(local_time_step)
    c.second_order = int(second_order)
    c.method = method
    c.n_bc = len(bcs)
    for i, (typ, val) in enumerate(bcs):
        c.bc_euler_id[i] = BCS[typ]
        for k in range(5):
            c.bc_val[i][k] = float(val[k])
    c.cfl = cfl
    c.tmax = tmax
    return c


# every function declared in include/subfv.h with its ctypes signature
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_i64 = C.c_int64
_i32 = C.c_int32
SIGNATURES = {
    "subfv_create": (C.c_int, [_PP, C.c_int]),
    "subfv_destroy": (None, [_P]),
    "subfv_error": (C.c_char_p, [_P]),
    "subfv_mesh_build": (C.c_int, [_P, _i32, _dp, _i32, _ip, _ip, _i32, _ip, _ip, _ip, _i32]),
    "subfv_mesh_upload": (C.c_int, [_P, _P, _i32]),
    "subfv_mesh_count": (_i64, [_P, C.c_char_p]),
    "subfv_mesh_get_i32": (_i64, [_P, C.c_char_p, _ip, _i64]),
    "subfv_mesh_get_f64": (_i64, [_P, C.c_char_p, _dp, _i64]),
    "subfv_host_mesh_build": (C.c_int, [_PP, _i32, _dp, _i32, _ip, _ip, _i32, _ip, _ip, _ip]),
    "subfv_host_mesh_build_on_device": (C.c_int, [_PP, _P, _i32, _dp, _i32, _ip, _ip, _i32, _ip, _ip, _ip]),
    "subfv_host_mesh_make_periodic": (C.c_int, [_P]),
    "subfv_host_mesh_get_i32": (_i64, [_P, C.c_char_p, _ip, _i64]),
    "subfv_host_mesh_free": (None, [_P]),
    "subfv_set_config": (C.c_int, [_P, C.POINTER(Config)]),
    "subfv_set_arith": (C.c_int, [_P, _i32]),
    "subfv_set_time": (C.c_int, [_P, C.c_double]),
    "subfv_state_put": (C.c_int, [_P, _dp]),
    "subfv_state_get": (C.c_int, [_P, _dp]),
    "subfv_get_f64": (_i64, [_P, C.c_char_p, _dp, _i64]),
    "subfv_put_f64": (_i64, [_P, C.c_char_p, _dp, _i64]),
    "subfv_compute_euler_flux": (C.c_int, [_P, _i32, _i32]),
    "subfv_compute_cell_grad": (C.c_int, [_P, _i32]),
    "subfv_flux_limiting": (C.c_int, [_P, C.c_double]),
    "subfv_add_timedisc_and_solve": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp]),
    "subfv_compute_rhs_ns": (C.c_int, [_P]),
    "subfv_set_diffusion": (C.c_int, [_P, _i32, C.c_double, C.c_double, C.c_double, _i32, _ip, _dp, _ip, _dp]),
    "subfv_set_sutherland": (C.c_int, [_P, _i32, C.c_double, C.c_double]),
    "subfv_compute_mu": (C.c_int, [_P, _dp, _dp, _dp]),
    "subfv_add_diffusion": (C.c_int, [_P, _dp, _dp, C.c_double, _i32, _dp, _dp]),
    "subfv_run": (C.c_int, [_P, _i32, _i32, _dp, _dp]),
    "subfv_profile_enable": (C.c_int, [_P, _i32]),
    "subfv_profile_read": (C.c_int, [_P, C.c_char_p, _dp, C.POINTER(C.c_int64)]),
    "subfv_batch_create": (C.c_int, [_PP, C.c_int]),
    "subfv_batch_destroy": (None, [_P]),
    "subfv_batch_error": (C.c_char_p, [_P]),
    "subfv_batch_run": (C.c_int, [_P, C.POINTER(C.c_void_p), _i32, _i32, _i32]),
    "subfv_batch_set_grid": (C.c_int, [_P, _i32]),
    "subfv_batch_wait": (C.c_int, [_P]),
    "subfv_init_1drp": (C.c_int, [_P, C.c_double, _dp, _dp]),
    "subfv_compute_rhs_lagrange": (C.c_int, [_P, _dp, _dp, _i32, _dp, _dp]),
    "subfv_acoustic_run": (C.c_int, [_P, _dp, _i32, C.c_double, _i32, _ip, _dp, _dp, _dp]),
    "subfv_bcsr_create": (C.c_int, [_P, _PP, _i32, _i32, _i32, _ip, _ip, _dp]),
    "subfv_bcsr_destroy": (None, [_P]),
    "subfv_bcsr_matmul": (C.c_int, [_P, _dp, _dp]),
    "subfv_bcsr_diag": (C.c_int, [_P, _dp, _dp]),
    "subfv_bcsr_set_values": (C.c_int, [_P, _dp]),
    "subfv_bcsr_gmres_diag": (C.c_int, [_P, _dp, _dp, C.c_double, _i32, _i32, _ip, _dp]),
    "subfv_bcsr_gmres_diag_noalloc": (C.c_int, [_P, _dp, _dp, C.c_double, _i32, _i32, _ip, _dp]),
    "subfv_bcsr_gmres": (C.c_int, [_P, _dp, _dp, C.c_double, _i32, _i32, _i32, _ip, _dp]),
    "subfv_bcsr_lusgs": (C.c_int, [_P, _dp, _dp]),
    "subfv_bcsr_set_colors": (C.c_int, [_P, _ip, _ip]),
    "subfv_bcsr_parallel_lusgs": (C.c_int, [_P, _dp, _dp]),
    "subfv_bcsr_pcgm": (C.c_int, [_P, _dp, _dp, C.c_double, _dp]),
    "subfv_bcsr_add_block": (C.c_int, [_P, _i32, _i32, _dp]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libsubfv_hip.so (built by subfv_amd.build); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)
