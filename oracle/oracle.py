"""TEST INFRASTRUCTURE ONLY: ctypes front end of the oracle library.

Loaded by tests/, by the smoke check and by the CPU-baseline leg of bench.py
as the checker; the product never imports this module.  Arrays use the
reference's Fortran layouts (sol(5,C) -> shape (C, 5), grad(3,5,C) ->
shape (C, 5, 3), ...).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libsubfv_oracle.so")
_lib = None

MNBC = 10
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
    _fields_ = [
        ("scheme_id", C.c_int),
        ("bc_style", C.c_int),
        ("boundary_2d", C.c_int),
        ("exclude_bound_vert", C.c_int),
        ("local_time_step", C.c_int),
        ("second_order", C.c_int),
        ("method", C.c_int),
        ("n_bc", C.c_int),
        ("bc_euler_id", C.c_int * (MNBC + 1)),
        ("bc_val", (C.c_double * 5) * (MNBC + 1)),
        ("cfl", C.c_double),
        ("tmax", C.c_double),
        ("t", C.c_double),
        ("activate_diffusion", C.c_int),
        ("mu_p", C.c_double),
        ("Cv_p", C.c_double),
        ("prandtl", C.c_double),
        ("bc_V_id", C.c_int * (MNBC + 1)),
        ("bc_T_id", C.c_int * (MNBC + 1)),
        ("bc_val_V", (C.c_double * 3) * (MNBC + 1)),
        ("bc_val_T", C.c_double * (MNBC + 1)),
        ("use_sutherland", C.c_int),
        ("mu0", C.c_double),
        ("T0", C.c_double),
    ]


def make_config(scheme="multi_point", bc_style=2, boundary_2d=True, exclude_bound_vert=False,
                local_time_step=False, second_order=False, method=4, bcs=(), cfl=0.4, tmax=1e30,
                t=0.0, diffusion=False, mu=1e-3, cv=720.19471, prandtl=0.71, bcs_v=(), bcs_t=(),
                sutherland=False, mu0=1.716e-5, T0=273.15) -> Config:
    """bcs: sequence of (bc_type, (rho,u,v,w,p)) in BC index order 1..n_bc;
    bcs_v / bcs_t: ("neumann"|"dirichlet", value) per BC index for the
    viscous terms (velocity (3,) / temperature), default neumann;
    sutherland: per-cell viscosity mu0 (T/T0)^1.5 (T0 + 110.4) / (T + 110.4)."""
    c = Config()
    c.scheme_id = SCHEMES[scheme]
    c.bc_style = bc_style
    c.boundary_2d = int(boundary_2d)
    c.exclude_bound_vert = int(exclude_bound_vert)
    c.local_time_step = int(local_time_step)
    c.second_order = int(second_order)
    c.method = method
    c.n_bc = len(bcs)
    for i, (typ, val) in enumerate(bcs, start=1):
        c.bc_euler_id[i] = BCS[typ]
        for k in range(5):
            c.bc_val[i][k] = float(val[k])
    c.cfl = cfl
    c.tmax = tmax
    c.t = t
    c.activate_diffusion = int(diffusion)
    c.mu_p = mu
    c.Cv_p = cv
    c.prandtl = prandtl
    c.use_sutherland = int(sutherland)
    c.mu0 = mu0
    c.T0 = T0
    for i in range(1, MNBC + 1):
        c.bc_V_id[i] = 1
        c.bc_T_id[i] = 1
    for i, (typ, val) in enumerate(bcs_v, start=1):
        c.bc_V_id[i] = 2 if typ == "dirichlet" else 1
        for k in range(3):
            c.bc_val_V[i][k] = float(val[k])
    for i, (typ, val) in enumerate(bcs_t, start=1):
        c.bc_T_id[i] = 2 if typ == "dirichlet" else 1
        c.bc_val_T[i] = float(val)
    return c


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.om_mesh_build.restype = P
        L.om_mesh_build.argtypes = [C.c_int, dp, C.c_int, ip, ip, C.c_int, ip, ip, ip, C.c_int,
                                    C.POINTER(C.c_int)]
        L.om_mesh_build_p.restype = P
        L.om_mesh_build_p.argtypes = [C.c_int, dp, C.c_int, ip, ip, C.c_int, ip, ip, ip, C.c_int, C.c_int,
                                      C.POINTER(C.c_int)]
        L.om_mesh_free.argtypes = [P]
        L.om_mesh_counts.argtypes = [P, ip]
        L.om_mesh_get_i.restype = C.c_long
        L.om_mesh_get_i.argtypes = [P, C.c_char_p, ip]
        L.om_mesh_get_d.restype = C.c_long
        L.om_mesh_get_d.argtypes = [P, C.c_char_p, dp]
        cp = C.POINTER(Config)
        L.om_compute_euler_flux.argtypes = [P, cp, dp, dp, P, dp, C.c_int, P]
        L.om_compute_cell_grad.argtypes = [P, dp, dp, dp, C.c_int]
        L.om_flux_limiting.argtypes = [P, dp, dp, dp, dp, C.c_double]
        L.om_add_timedisc_and_solve.argtypes = [P, cp, dp, dp, dp, dp, dp, C.POINTER(C.c_double), dp,
                                                C.c_double, C.c_double, C.c_double, C.c_double,
                                                C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.om_run.argtypes = [P, cp, dp, C.c_int, C.c_int, dp]
        L.om_max_omega_for_positivity.restype = C.c_double
        L.om_max_omega_for_positivity.argtypes = [dp, dp]
        L.om_bcsr_matmul.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp]
        L.om_bcsr_diag.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp]
        L.om_bcsr_gmres_diag.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_int, dp]
        L.om_bcsr_lusgs.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp]
        L.om_bcsr_parallel_lusgs.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, C.c_int, ip, ip, ip]
        L.om_bcsr_add_block.argtypes = [C.c_int, C.c_int, ip, ip, dp, C.c_int, C.c_int, dp]
        L.om_bcsr_gmres_kind.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_int,
                                         C.c_int, ip, ip, ip, dp]
        L.om_bcsr_pcgm.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, C.c_double, dp]
        L.om_add_diffusion.argtypes = [P, cp, dp, dp, dp]
        L.om_add_diffusion_mu.argtypes = [P, cp, dp, dp, C.c_double, C.c_int, dp, dp]
        L.om_compute_mu.restype = C.c_double
        L.om_compute_mu.argtypes = [C.c_int, cp, dp, dp]
        L.om_test_inv.argtypes = [C.c_int, dp, dp]
        L.om_inverse_3_by_3.argtypes = [dp, dp]
        L.om_potential_flow.argtypes = [C.c_int, dp, dp]
        L.om_compute_rhs_lagrange.argtypes = [P, dp, dp, C.c_int, dp, dp]
        L.om_acoustic_step.argtypes = [P, dp, C.c_int, ip, dp, C.c_double, C.POINTER(C.c_double)]
        _lib = L
    return _lib


_INT_NAMES = [
    "vert_nto", "elem_nto", "vert_is_bound", "vert_n_sub_faces", "vert_n_sub_elems", "vert_n_elems",
    "vert_sub_face", "vert_sub_elem", "vert_elem", "elem_n_vert", "elem_kind", "elem_vert",
    "elem_face", "elem_sub_elem", "elem_tag", "elem_n_neigh_by_vert", "elem_neigh_by_vert",
    "face_n_vert", "face_vert", "face_left", "face_right", "face_sub_face", "sub_face_vert",
    "sub_face_face", "sub_face_loc", "sub_face_left_elem", "sub_face_right_elem",
    "sub_face_left_sub_elem", "sub_face_right_sub_elem", "sub_elem_loc", "sub_elem_vert",
    "sub_elem_elem", "sub_elem_sub_face",
]
_DBL_NAMES = ["vert_coord", "vert_volume", "elem_coord", "elem_volume", "face_coord", "face_norm",
              "face_area", "sub_face_norm", "sub_face_area", "sub_elem_volume"]


class OracleMesh:
    """build_mesh + compute_geometry_mesh on the CPU."""

    def __init__(self, mesh, periodic: bool = False):
        """periodic: stitch the faces on x, y, z = -0.5 / 0.5 and extend the
        stencils of the vertices there (build_mesh with periodic_mesh)."""
        L = lib()
        st = C.c_int(0)
        self.h = L.om_mesh_build_p(mesh.n_vert, np.ascontiguousarray(mesh.coord, np.float64).ravel(),
                                   mesh.n_elems, np.ascontiguousarray(mesh.kind, np.int32),
                                   np.ascontiguousarray(mesh.elem_vert, np.int32), int(mesh.bf_nvert.size),
                                   np.ascontiguousarray(mesh.bf_nvert, np.int32),
                                   np.ascontiguousarray(mesh.bf_vert, np.int32),
                                   np.ascontiguousarray(mesh.bf_tag, np.int32), int(mesh.boundary_2d),
                                   int(periodic), C.byref(st))
        self.status = st.value
        if not self.h:
            raise ValueError(f"oracle mesh build failed (status {self.status})")
        cnt = np.zeros(6, np.int32)
        L.om_mesh_counts(self.h, cnt)
        self.n_vert, self.n_elems, self.n_faces, self.n_sub_faces, self.n_sub_elems, self.nmax_sfn = map(int, cnt)
        self.boundary_2d = bool(mesh.boundary_2d)
        self._cache = {}

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().om_mesh_free(self.h)
                self.h = None
        except Exception:
            pass

    def _cap(self):
        return 64 * (self.n_sub_faces + self.n_sub_elems + self.n_elems + self.n_vert + self.n_faces) + 64

    def geti(self, name: str) -> np.ndarray:
        if name not in self._cache:
            buf = np.zeros(self._cap(), np.int32)
            n = lib().om_mesh_get_i(self.h, name.encode(), buf)
            if n < 0:
                raise KeyError(name)
            self._cache[name] = buf[:n].copy()
        return self._cache[name]

    def getd(self, name: str) -> np.ndarray:
        if name not in self._cache:
            buf = np.zeros(self._cap(), np.float64)
            n = lib().om_mesh_get_d(self.h, name.encode(), buf)
            if n < 0:
                raise KeyError(name)
            self._cache[name] = buf[:n].copy()
        return self._cache[name]

    # convenience views
    @property
    def elem_coord(self):
        return self.getd("elem_coord").reshape(-1, 3)

    @property
    def elem_volume(self):
        return self.getd("elem_volume")

    # ------------------------------------------------------------ physics
    def compute_euler_flux(self, cfg: Config, sol: np.ndarray, grad=None, second_order=False, want_vp=False):
        ne, nv = self.n_elems, self.n_vert
        sol = np.ascontiguousarray(sol, np.float64).reshape(ne, 5)
        rhs = np.zeros((ne, 5))
        sl = np.zeros(ne)
        vp = np.zeros((nv, 3)) if want_vp else None
        g = None
        if grad is not None:
            g = np.ascontiguousarray(grad, np.float64).reshape(ne, 5, 3)
        st = lib().om_compute_euler_flux(self.h, C.byref(cfg), sol.ravel(), rhs.reshape(-1),
                                         None if g is None else g.ctypes.data_as(C.c_void_p),
                                         sl, int(second_order),
                                         None if vp is None else vp.ctypes.data_as(C.c_void_p))
        if st:
            raise RuntimeError(f"oracle compute_euler_flux status {st}")
        return (rhs, sl, vp) if want_vp else (rhs, sl)

    def compute_rhs_lagrange(self, sol: np.ndarray, mass: np.ndarray, b2d: bool = True):
        """compute_rhs_lagrange: (rhs (ne, 5), vp (nv, 3))."""
        ne, nv = self.n_elems, self.n_vert
        sol = np.ascontiguousarray(sol, np.float64).reshape(ne, 5)
        rhs = np.zeros((ne, 5))
        vp = np.zeros((nv, 3))
        lib().om_compute_rhs_lagrange(self.h, sol.ravel(), np.ascontiguousarray(mass, np.float64), int(b2d),
                                      rhs.reshape(-1), vp.reshape(-1))
        return rhs, vp

    def acoustic_step(self, sol: np.ndarray, cfl: float, bc_kind=(), bc_val=()):
        """One step of the acoustic loop on sol (ne, 4) = (p, u, v, w).
        bc_kind[t-1] for tag t: 0 wall, 1 freestream, 2 outflow; bc_val[t-1] the
        freestream state.  Returns (new sol, dt)."""
        ne = self.n_elems
        out = np.ascontiguousarray(sol, np.float64).reshape(ne, 4).copy()
        nb = len(bc_kind)
        kinds = np.zeros(nb + 1, np.int32)
        vals = np.zeros((nb + 1, 4))
        for t in range(nb):
            kinds[t + 1] = bc_kind[t]
            if len(bc_val) > t:
                vals[t + 1] = bc_val[t]
        dt = C.c_double(0.0)
        st = lib().om_acoustic_step(self.h, out.reshape(-1), nb, kinds, vals.reshape(-1), cfl, C.byref(dt))
        if st:
            raise RuntimeError("oracle acoustic_step: unsupported boundary tag")
        return out, dt.value

    def add_diffusion(self, cfg: Config, sol: np.ndarray, rhs=None, diag=None):
        """The viscous block of compute_rhs_ns (momentum, then heat) added to
        rhs (n_elems, 5) and diag (n_elems, 5, 5 column-major blocks, i.e.
        diag[c].T is the 5x5 matrix); zeros if not given."""
        ne = self.n_elems
        sol = np.ascontiguousarray(sol, np.float64).reshape(ne, 5)
        rhs = np.zeros((ne, 5)) if rhs is None else np.ascontiguousarray(rhs, np.float64).copy()
        diag = np.zeros((ne, 5, 5)) if diag is None else np.ascontiguousarray(diag, np.float64).copy()
        st = lib().om_add_diffusion(self.h, C.byref(cfg), sol.ravel(), rhs.reshape(-1), diag.reshape(-1))
        if st:
            raise RuntimeError(f"oracle add_diffusion status {st}")
        return rhs, diag

    def add_diffusion_mu(self, cfg: Config, sol: np.ndarray, mu_arr, mu0corr: float, parts: int = 3,
                         rhs=None, diag=None):
        """build_momentum_diff_mat_into_global_mat_diag (parts 1) and
        build_diff_mat_into_global_mat_diag (parts 2) for given cell
        viscosities mu_arr (n_elems) and mu0corr, added to rhs and diag in the
        layouts of add_diffusion."""
        ne = self.n_elems
        sol = np.ascontiguousarray(sol, np.float64).reshape(ne, 5)
        mu = np.ascontiguousarray(mu_arr, np.float64).reshape(ne)
        rhs = np.zeros((ne, 5)) if rhs is None else np.ascontiguousarray(rhs, np.float64).copy()
        diag = np.zeros((ne, 5, 5)) if diag is None else np.ascontiguousarray(diag, np.float64).copy()
        st = lib().om_add_diffusion_mu(self.h, C.byref(cfg), sol.ravel(), mu, float(mu0corr), int(parts),
                                       rhs.reshape(-1), diag.reshape(-1))
        if st:
            raise RuntimeError(f"oracle add_diffusion_mu status {st}")
        return rhs, diag

    def compute_cell_grad(self, sol: np.ndarray, method: int = 4):
        ne, nv = self.n_elems, self.n_vert
        grad = np.zeros((ne, 5, 3))
        ng = np.zeros((nv, 5, 3))
        st = lib().om_compute_cell_grad(self.h, np.ascontiguousarray(sol, np.float64).ravel(),
                                        grad.reshape(-1), ng.reshape(-1), method)
        if st:
            raise RuntimeError(f"oracle compute_cell_grad status {st}")
        return grad, ng

    def flux_limiting(self, sol, rhs, rhs_o2, sum_lambda, cfl):
        out = np.ascontiguousarray(rhs, np.float64).copy()
        lib().om_flux_limiting(self.h, np.ascontiguousarray(sol, np.float64).ravel(), out.reshape(-1),
                               np.ascontiguousarray(rhs_o2, np.float64).ravel(),
                               np.ascontiguousarray(sum_lambda, np.float64), cfl)
        return out

    def add_timedisc_and_solve(self, cfg, sol, rhs, sum_lambda, a1, a2, cfl, t, diag=None):
        ne = self.n_elems
        sol = np.ascontiguousarray(sol, np.float64).copy()
        rhs = np.ascontiguousarray(rhs, np.float64).copy()
        sl = np.ascontiguousarray(sum_lambda, np.float64).copy()
        dg = np.zeros((ne, 5, 5)) if diag is None else np.ascontiguousarray(diag, np.float64).copy()
        ds = np.zeros((ne, 5))
        res = C.c_double(0)
        rv = np.zeros(5)
        dt = C.c_double(0)
        om = C.c_double(0)
        st = lib().om_add_timedisc_and_solve(self.h, C.byref(cfg), sol.reshape(-1), rhs.reshape(-1),
                                             dg.reshape(-1), sl, ds.reshape(-1), C.byref(res), rv, a1, a2,
                                             cfl, t, C.byref(dt), C.byref(om))
        return dict(status=st, sol=sol, delta_sol=ds, res=res.value, res_vect=rv, dt=dt.value, omega=om.value)

    def run(self, cfg: Config, sol: np.ndarray, n_steps: int, time_scheme: int = 0):
        sol = np.ascontiguousarray(sol, np.float64).copy()
        out = np.zeros(3)
        st = lib().om_run(self.h, C.byref(cfg), sol.reshape(-1), n_steps, time_scheme, out)
        return st, sol, dict(t=out[0], dt=out[1], res=out[2])


def max_omega_for_positivity(U, dU) -> float:
    return lib().om_max_omega_for_positivity(np.ascontiguousarray(U, np.float64),
                                             np.ascontiguousarray(dU, np.float64))


# ------------------------------------------------------------------ BCSR
def bcsr_matmul(m, b, row, col, v, x):
    y = np.zeros(b * m)
    lib().om_bcsr_matmul(m, b, np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32),
                         np.ascontiguousarray(v, np.float64).ravel(), np.ascontiguousarray(x, np.float64), y)
    return y


def bcsr_diag(m, b, row, col, v, rhs):
    x = np.zeros(b * m)
    lib().om_bcsr_diag(m, b, np.ascontiguousarray(row, np.int32), np.ascontiguousarray(col, np.int32),
                       np.ascontiguousarray(v, np.float64).ravel(), x, np.ascontiguousarray(rhs, np.float64))
    return x


def bcsr_gmres_diag(m, b, row, col, v, rhs, tol, restart=30):
    x = np.zeros(b * m)
    info = np.zeros(2)
    conv = lib().om_bcsr_gmres_diag(m, b, np.ascontiguousarray(row, np.int32),
                                    np.ascontiguousarray(col, np.int32),
                                    np.ascontiguousarray(v, np.float64).ravel(), x,
                                    np.ascontiguousarray(rhs, np.float64), tol, restart, info)
    return x, bool(conv), dict(restarts=int(info[0]), rho=info[1])


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _f64(a):
    return np.ascontiguousarray(a, np.float64).ravel()


def color_sets(color):
    """(nsets, set_off, set_rows) from a 1-based colour per row: the rows of
    set cs in increasing order."""
    color = np.asarray(color, np.int32)
    nsets = int(color.max())
    rows = np.argsort(color, kind="stable").astype(np.int32) + 1
    off = np.concatenate([[0], np.cumsum(np.bincount(color, minlength=nsets + 1)[1:])]).astype(np.int32)
    return nsets, off, rows


def bcsr_lusgs(m, b, row, col, v, rhs):
    x = np.zeros(b * m)
    lib().om_bcsr_lusgs(m, b, _i32(row), _i32(col), _f64(v), x, _f64(rhs))
    return x


def bcsr_parallel_lusgs(m, b, row, col, v, rhs, color):
    x = np.zeros(b * m)
    nsets, off, rows = color_sets(color)
    lib().om_bcsr_parallel_lusgs(m, b, _i32(row), _i32(col), _f64(v), x, _f64(rhs), nsets, off, rows, _i32(color))
    return x


def bcsr_add_block(m, b, row, col, v, i, j, M):
    """add_block_to_bcsr on a copy of v; returns the new values."""
    out = _f64(v).copy()
    lib().om_bcsr_add_block(m, b, _i32(row), _i32(col), out, i, j, _f64(np.asarray(M).T))
    return out


def bcsr_gmres(m, b, row, col, v, rhs, tol, restart=30, precond="none", color=None):
    """bcsr_gmres (precond "none"), bcsr_gmres_lusgs ("lusgs") and
    bcsr_gmres_parallel_lusgs ("parallel_lusgs", with a 1-based colour per row)."""
    kind = {"none": 0, "lusgs": 2, "parallel_lusgs": 3}[precond]
    x = np.zeros(b * m)
    info = np.zeros(2)
    if kind == 3:
        nsets, off, rows = color_sets(color)
        col_arr = _i32(color)
    else:
        nsets, off, rows, col_arr = 0, np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1, np.int32)
    conv = lib().om_bcsr_gmres_kind(kind, m, b, _i32(row), _i32(col), _f64(v), x, _f64(rhs), tol, restart,
                                    nsets, off, rows, col_arr, info)
    return x, bool(conv), dict(restarts=int(info[0]), rho=info[1])


def bcsr_pcgm(m, b, row, col, v, rhs, tol, x0=None):
    x = np.zeros(b * m) if x0 is None else _f64(x0).copy()
    info = np.zeros(2)
    conv = lib().om_bcsr_pcgm(m, b, _i32(row), _i32(col), _f64(v), x, _f64(rhs), tol, info)
    return x, int(conv), dict(iterations=int(info[0]), rho=info[1])


def potential_flow(x, three_d: bool) -> np.ndarray:
    """sol_potential_flow_2d / _3d at the point x: primitive (rho, u, v, w, p)."""
    w = np.zeros(5)
    lib().om_potential_flow(int(three_d), np.ascontiguousarray(x, np.float64), w)
    return w


def inverse_3_by_3(A: np.ndarray) -> np.ndarray:
    out = np.zeros(9)
    lib().om_inverse_3_by_3(np.ascontiguousarray(np.asarray(A, np.float64).T).ravel(), out)
    return out.reshape(3, 3).T


def inv(A: np.ndarray) -> np.ndarray:
    """inv_lapack restatement (LU + unblocked dgetri) of a square matrix."""
    A = np.asarray(A, np.float64)
    n = A.shape[0]
    out = np.zeros(n * n)
    lib().om_test_inv(n, np.ascontiguousarray(A.T).ravel(), out)
    return out.reshape(n, n).T


def compute_mu(cfg: Config, sol: np.ndarray):
    """compute_mu restatement: (mu of every cell, mu0corr) for sol (C, 5)."""
    sol = np.ascontiguousarray(sol, np.float64).reshape(-1, 5)
    mu = np.zeros(sol.shape[0])
    mn = lib().om_compute_mu(sol.shape[0], C.byref(cfg), sol.ravel(), mu)
    return mu, mn
