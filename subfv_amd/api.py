"""Host-side mirror of subfv's explicit residual interface over the C-ABI.

Method names, argument meaning and error behaviour follow the reference's
module procedures (compute_euler_flux, compute_cell_grad_from_nodal_grad,
flux_limiting, add_timedisc_and_solve, compute_rhs_ns, the ns_main time
loop); the reference's fatal ``error stop`` conditions raise ``SubfvError``
with the reference message.  Arrays use the Fortran layouts of the
reference: sol(5,C) -> shape (C, 5), grad(3,5,C) -> shape (C, 5, 3).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Config, make_config  # noqa: F401  (re-exported)


class SubfvError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[subfv {code}] {msg}")
        self.code = code


class HostMesh:
    """Connectivity of build_mesh in a mesh handle: computed on the host, or
    on the device of `solver` (subfv_host_mesh_build_on_device)."""

    def __init__(self, mesh, periodic: bool = False, solver: "Solver | None" = None):
        """periodic: build_mesh with periodic_mesh (faces on x, y, z = -0.5 / 0.5
        stitched to their partners, subfv_host_mesh_make_periodic)."""
        L = _lib.lib()
        h = C.c_void_p()
        coord = _lib.f64(mesh.coord).ravel()
        kind = _lib.i32(mesh.kind)
        ev = _lib.i32(mesh.elem_vert)
        bn, bv, bt = _lib.i32(mesh.bf_nvert), _lib.i32(mesh.bf_vert), _lib.i32(mesh.bf_tag)
        if solver is not None:
            st = L.subfv_host_mesh_build_on_device(C.byref(h), solver.h, mesh.n_vert, _lib.dptr(coord),
                                                   mesh.n_elems, _lib.iptr(kind), _lib.iptr(ev), int(bn.size),
                                                   _lib.iptr(bn), _lib.iptr(bv), _lib.iptr(bt))
            if st:
                solver._check(st)
        else:
            st = L.subfv_host_mesh_build(C.byref(h), mesh.n_vert, _lib.dptr(coord), mesh.n_elems, _lib.iptr(kind),
                                         _lib.iptr(ev), int(bn.size), _lib.iptr(bn), _lib.iptr(bv), _lib.iptr(bt))
        if st:
            raise SubfvError(st, "host mesh build failed")
        self.h = h
        self._cap = 64 * (ev.size + mesh.n_vert + mesh.n_elems) + 1024
        if periodic:
            st = L.subfv_host_mesh_make_periodic(self.h)
            if st:
                raise SubfvError(st, "Error periodic boundaries !")

    def geti(self, name: str) -> np.ndarray:
        out = np.zeros(self._cap, np.int32)
        n = _lib.lib().subfv_host_mesh_get_i32(self.h, name.encode(), _lib.iptr(out), out.size)
        if n < 0:
            raise KeyError(name)
        return out[:n].copy()

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.lib().subfv_host_mesh_free(self.h)
                self.h = None
        except Exception:
            pass


class Solver:
    """One device context: mesh, configuration and resident state."""

    def __init__(self, device: int = 0):
        L = _lib.lib()
        h = C.c_void_p()
        st = L.subfv_create(C.byref(h), device)
        if st:
            raise SubfvError(st, "subfv_create failed (no usable HIP device?)")
        self.h = h
        self.n_vert = self.n_elems = 0

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().subfv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int):
        if st:
            msg = _lib.lib().subfv_error(self.h)
            raise SubfvError(st, msg.decode() if msg else "")

    # ------------------------------------------------------------- mesh
    def build_mesh(self, mesh, periodic: bool = False):
        """read_mesh_msh_2 + build_mesh + compute_geometry_mesh (periodic:
        with periodic_mesh, through a host connectivity)."""
        if periodic:
            self.upload_mesh(HostMesh(mesh, periodic=True), mesh.boundary_2d)
            return
        L = _lib.lib()
        coord = _lib.f64(mesh.coord).ravel()
        kind = _lib.i32(mesh.kind)
        ev = _lib.i32(mesh.elem_vert)
        bn, bv, bt = _lib.i32(mesh.bf_nvert), _lib.i32(mesh.bf_vert), _lib.i32(mesh.bf_tag)
        self._check(L.subfv_mesh_build(self.h, mesh.n_vert, _lib.dptr(coord), mesh.n_elems, _lib.iptr(kind),
                                       _lib.iptr(ev), int(bn.size), _lib.iptr(bn), _lib.iptr(bv), _lib.iptr(bt),
                                       int(mesh.boundary_2d)))
        self._counts()

    def upload_mesh(self, host_mesh: "HostMesh", boundary_2d: bool):
        """Device side of build_mesh for a connectivity built beforehand on the
        host (HostMesh): upload, geometry and BC data."""
        self._check(_lib.lib().subfv_mesh_upload(self.h, host_mesh.h, int(boundary_2d)))
        self._counts()

    def _counts(self):
        L = _lib.lib()
        self.n_vert = int(L.subfv_mesh_count(self.h, b"n_vert"))
        self.n_elems = int(L.subfv_mesh_count(self.h, b"n_elems"))
        self.n_faces = int(L.subfv_mesh_count(self.h, b"n_faces"))
        self.n_sub_faces = int(L.subfv_mesh_count(self.h, b"n_sub_faces"))
        self.n_sub_elems = int(L.subfv_mesh_count(self.h, b"n_sub_elems"))
        self._cap = 64 * (self.n_sub_faces + self.n_sub_elems + self.n_elems + self.n_vert + self.n_faces) + 64

    def count(self, what: str) -> int:
        return int(_lib.lib().subfv_mesh_count(self.h, what.encode()))

    def mesh_i(self, name: str) -> np.ndarray:
        out = np.zeros(self._cap, np.int32)
        n = _lib.lib().subfv_mesh_get_i32(self.h, name.encode(), _lib.iptr(out), out.size)
        if n < 0:
            raise KeyError(name)
        return out[:n].copy()

    def mesh_d(self, name: str) -> np.ndarray:
        out = np.zeros(self._cap, np.float64)
        n = _lib.lib().subfv_mesh_get_f64(self.h, name.encode(), _lib.dptr(out), out.size)
        if n < 0:
            raise KeyError(name)
        return out[:n].copy()

    @property
    def elem_coord(self) -> np.ndarray:
        return self.mesh_d("elem_coord").reshape(-1, 3)

    @property
    def elem_volume(self) -> np.ndarray:
        return self.mesh_d("elem_volume")

    # ------------------------------------------------------------- config/state
    def set_config(self, cfg: Config):
        self._check(_lib.lib().subfv_set_config(self.h, C.byref(cfg)))

    def set_arith(self, fast: bool):
        """The node sweep's arithmetic: exact (the reference's expressions as
        written, bit-identical) or fast (contracted multiply-adds and
        reciprocal products: within 1e-12 relative, subfv_set_arith)."""
        self._check(_lib.lib().subfv_set_arith(self.h, 1 if fast else 0))

    def set_time(self, t: float):
        self._check(_lib.lib().subfv_set_time(self.h, float(t)))

    def put_sol(self, sol: np.ndarray):
        s = _lib.f64(sol).reshape(self.n_elems, 5)
        self._check(_lib.lib().subfv_state_put(self.h, _lib.dptr(s)))

    def init_1drp(self, x1drp: float, wl, wr):
        """init_1drp on the device: left/right primitive states by centroid x."""
        wl = _lib.f64(wl).reshape(5)
        wr = _lib.f64(wr).reshape(5)
        self._check(_lib.lib().subfv_init_1drp(self.h, float(x1drp), _lib.dptr(wl), _lib.dptr(wr)))

    def get_sol(self) -> np.ndarray:
        s = np.zeros((self.n_elems, 5))
        self._check(_lib.lib().subfv_state_get(self.h, _lib.dptr(s)))
        return s

    _SHAPES = {"sol": (5,), "rhs": (5,), "rhs_o2": (5,), "sum_lambda": (), "sum_lambda_o2": (),
               "cell_grad": (5, 3), "nodal_grad": (5, 3), "vp": (3,), "diag": (5, 5), "rhs_diff": (4,)}
    _PER_VERT = {"nodal_grad", "vp"}

    def get(self, name: str) -> np.ndarray:
        n = self.n_vert if name in self._PER_VERT else self.n_elems
        shape = (n,) + self._SHAPES[name]
        out = np.zeros(shape)
        r = _lib.lib().subfv_get_f64(self.h, name.encode(), _lib.dptr(out), out.size)
        if r < 0:
            raise KeyError(name)
        return out

    def put(self, name: str, arr: np.ndarray):
        a = _lib.f64(arr)
        r = _lib.lib().subfv_put_f64(self.h, name.encode(), _lib.dptr(a), a.size)
        if r < 0:
            raise KeyError(name)

    # ------------------------------------------------------------- procedures
    def compute_euler_flux(self, sol=None, grad=None, second_order=False, want_vp=False):
        """compute_euler_flux: returns (rhs, sum_lambda[, vp])."""
        if sol is not None:
            self.put_sol(sol)
        if second_order:
            if grad is not None:
                self.put("cell_grad", grad)
            self._check(_lib.lib().subfv_compute_euler_flux(self.h, 1, int(want_vp)))
            out = (self.get("rhs_o2"), self.get("sum_lambda_o2"))
        else:
            self._check(_lib.lib().subfv_compute_euler_flux(self.h, 0, int(want_vp)))
            out = (self.get("rhs"), self.get("sum_lambda"))
        return out + (self.get("vp"),) if want_vp else out

    def compute_cell_grad_from_nodal_grad(self, sol=None, method: int = 4):
        if sol is not None:
            self.put_sol(sol)
        self._check(_lib.lib().subfv_compute_cell_grad(self.h, method))
        return self.get("cell_grad"), self.get("nodal_grad")

    def flux_limiting(self, cfl: float):
        self._check(_lib.lib().subfv_flux_limiting(self.h, cfl))
        return self.get("rhs")

    def add_timedisc_and_solve(self, a1: float, a2: float, cfl: float, t: float):
        dt = C.c_double(0.0)
        om = C.c_double(0.0)
        self._check(_lib.lib().subfv_add_timedisc_and_solve(self.h, a1, a2, cfl, t, C.byref(dt), C.byref(om)))
        return dt.value, om.value

    def set_diffusion(self, activate: bool = True, mu: float = 1e-3, cv: float = 720.19471,
                      prandtl: float = 0.71, bcs_v=(), bcs_t=()):
        """activate_diffusion with constant viscosity mu (mu_p), Cv_p and the
        Prandtl number; bcs_v / bcs_t: ("neumann"|"dirichlet", value) per BC
        index 1..n (velocity (3,) / temperature)."""
        n = max(len(bcs_v), len(bcs_t))
        vid = np.ones(max(n, 1), np.int32)
        tid = np.ones(max(n, 1), np.int32)
        vv = np.zeros(3 * max(n, 1))
        tv = np.zeros(max(n, 1))
        for i, (typ, val) in enumerate(bcs_v):
            vid[i] = 2 if typ == "dirichlet" else 1
            vv[3 * i:3 * i + 3] = val
        for i, (typ, val) in enumerate(bcs_t):
            tid[i] = 2 if typ == "dirichlet" else 1
            tv[i] = val
        self._check(_lib.lib().subfv_set_diffusion(self.h, int(activate), mu, cv, prandtl, n, _lib.iptr(vid),
                                                   _lib.dptr(vv), _lib.iptr(tid), _lib.dptr(tv)))

    def set_sutherland(self, activate: bool = True, mu0: float = 1.716e-5, T0: float = 273.15):
        """use_sutherland of compute_mu: per-cell viscosity from the cell
        temperature (ns_vectorial_diffusion_module.F90:322-359)."""
        self._check(_lib.lib().subfv_set_sutherland(self.h, int(activate), mu0, T0))

    def compute_mu(self, sol):
        """compute_mu (ns_vectorial_diffusion_module.F90:322-347): (mu_arr
        (n_elems), mu0corr) of the state sol (n_elems, 5)."""
        sol = _lib.f64(sol).reshape(self.n_elems, 5)
        mu = np.zeros(self.n_elems)
        m0 = np.zeros(1)
        self._check(_lib.lib().subfv_compute_mu(self.h, _lib.dptr(sol), _lib.dptr(mu), _lib.dptr(m0)))
        return mu, float(m0[0])

    def add_diffusion(self, sol, mu_arr, mu0corr: float, parts: int = 3, rhs=None, diag=None):
        """build_momentum_diff_mat_into_global_mat_diag (parts 1,
        ns_vectorial_diffusion_module.F90:200-320) and
        build_diff_mat_into_global_mat_diag (parts 2,
        ns_diffusion_module.F90:112-184) for the state sol (n_elems, 5) and the
        cell viscosities mu_arr: (rhs, diag) with their contributions added to
        the given rhs (n_elems, 5) and diag (n_elems, 5, 5 column-major blocks),
        zeros if not given."""
        C_ = self.n_elems
        sol = _lib.f64(sol).reshape(C_, 5)
        mu = _lib.f64(mu_arr).reshape(C_)
        rhs = np.zeros((C_, 5)) if rhs is None else np.array(rhs, np.float64).reshape(C_, 5)
        diag = np.zeros((C_, 5, 5)) if diag is None else np.array(diag, np.float64).reshape(C_, 5, 5)
        self._check(_lib.lib().subfv_add_diffusion(self.h, _lib.dptr(sol), _lib.dptr(mu), float(mu0corr),
                                                   int(parts), _lib.dptr(rhs), _lib.dptr(diag)))
        return rhs, diag

    def compute_rhs_ns(self):
        self._check(_lib.lib().subfv_compute_rhs_ns(self.h))
        return self.get("rhs"), self.get("sum_lambda")

    def compute_rhs_lagrange(self, sol, mass, b2d: bool = True):
        """compute_rhs_lagrange on the device: (rhs (C, 5), vp (V, 3))."""
        C_, V_ = self.n_elems, self.n_vert
        sol = _lib.f64(sol).reshape(C_, 5)
        mass = _lib.f64(mass).reshape(C_)
        rhs = np.zeros((C_, 5))
        vp = np.zeros((V_, 3))
        self._check(_lib.lib().subfv_compute_rhs_lagrange(self.h, _lib.dptr(sol), _lib.dptr(mass), int(b2d),
                                                          _lib.dptr(rhs), _lib.dptr(vp)))
        return rhs, vp

    def acoustic_run(self, sol, n_steps: int, cfl: float, bc_kind=(), bc_val=(), t: float = 0.0):
        """n_steps of the acoustic loop on sol (C, 4) = (p, u, v, w);
        bc_kind per tag: 0 wall, 1 freestream, 2 outflow.  Returns (sol, t, dt)."""
        out = _lib.f64(sol).reshape(self.n_elems, 4).copy()
        nb = len(bc_kind)
        kinds = _lib.i32(np.asarray(bc_kind, np.int32).reshape(nb)) if nb else np.zeros(1, np.int32)
        vals = np.zeros((max(nb, 1), 4))
        for i in range(min(nb, len(bc_val))):
            vals[i] = bc_val[i]
        tt, dt = C.c_double(t), C.c_double(0.0)
        self._check(_lib.lib().subfv_acoustic_run(self.h, _lib.dptr(out), int(n_steps), float(cfl), nb,
                                                  _lib.iptr(kinds), _lib.dptr(vals), C.byref(tt), C.byref(dt)))
        return out, tt.value, dt.value

    def profile(self, on: bool = True):
        self._check(_lib.lib().subfv_profile_enable(self.h, int(on)))

    def profile_read(self, phase: str):
        """(total milliseconds, launches) of one phase since profile(True)."""
        ms = C.c_double(0.0)
        n = C.c_int64(0)
        self._check(_lib.lib().subfv_profile_read(self.h, phase.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def run(self, n_steps: int, time_scheme: int = 0):
        """Time loop on device-resident state; returns (t, dt)."""
        t = C.c_double(0.0)
        dt = C.c_double(0.0)
        self._check(_lib.lib().subfv_run(self.h, n_steps, time_scheme, C.byref(t), C.byref(dt)))
        return t.value, dt.value


class Bcsr:
    """bcsr_matrix_type resident on the device (sparse_bcsr_linear_module.F90:6-11):
    m block rows of b x b blocks, 1-based row(m+1) and col(nnzb), v(b,b,nnzb)
    column-major blocks."""

    def __init__(self, solver: "Solver", m: int, b: int, row, col, v):
        self.m, self.b = int(m), int(b)
        self._row = _lib.i32(row)
        self._col = _lib.i32(col)
        self._v = _lib.f64(v).ravel()
        self.nnzb = int(self._col.size)
        if self._row.size != self.m + 1 or self._v.size != self.nnzb * self.b * self.b:
            raise ValueError("inconsistent BCSR arrays")
        h = C.c_void_p()
        st = _lib.lib().subfv_bcsr_create(solver.h, C.byref(h), self.m, self.b, self.nnzb, _lib.iptr(self._row),
                                          _lib.iptr(self._col), _lib.dptr(self._v))
        if st:
            raise SubfvError(st, _lib.lib().subfv_error(solver.h).decode())
        self.h = h
        self.solver = solver

    @property
    def n(self) -> int:
        return self.m * self.b

    def _vec(self, x):
        x = _lib.f64(x).ravel()
        if x.size != self.n:
            raise ValueError(f"vector of length {x.size}, expected {self.n}")
        return x

    def matmul(self, x):
        """y = A x (bcsr_matmul)."""
        x = self._vec(x)
        y = np.zeros(self.n)
        self.solver._check(_lib.lib().subfv_bcsr_matmul(self.h, _lib.dptr(x), _lib.dptr(y)))
        return y

    def diag(self, rhs):
        """x = D^-1 rhs, block-Jacobi with lu_solve_noswp (bcsr_diag)."""
        rhs = self._vec(rhs)
        x = np.zeros(self.n)
        self.solver._check(_lib.lib().subfv_bcsr_diag(self.h, _lib.dptr(x), _lib.dptr(rhs)))
        return x

    def gmres_diag(self, rhs, tol: float, restart: int = 30, precision: int = 0):
        """bcsr_gmres_diag from x = 0 with absolute tolerance tol; precision 1
        runs the SpMV and block-Jacobi of each cycle in FP32.  Returns
        (x, conv, {"restarts", "rho", "spmv"})."""
        rhs = self._vec(rhs)
        x = np.zeros(self.n)
        conv = C.c_int32(0)
        stats = np.zeros(3)
        st = _lib.lib().subfv_bcsr_gmres_diag(self.h, _lib.dptr(x), _lib.dptr(rhs), tol, restart, precision,
                                              C.byref(conv), _lib.dptr(stats))
        if st not in (0, _lib.E_NOT_CONVERGED):
            self.solver._check(st)
        return x, bool(conv.value), {"restarts": int(stats[0]), "rho": float(stats[1]), "spmv": int(stats[2])}

    def gmres_diag_noalloc(self, rhs, tol: float, restart: int = 30, precision: int = 0):
        """gmres_diag with its Krylov basis in the library's fixed device
        scratch array (subfv_bcsr_gmres_diag_noalloc: no device allocation;
        n * (restart + 1) + (restart + 1) * restart <= 4 Mi doubles)."""
        rhs = self._vec(rhs)
        x = np.zeros(self.n)
        conv = C.c_int32(0)
        stats = np.zeros(3)
        st = _lib.lib().subfv_bcsr_gmres_diag_noalloc(self.h, _lib.dptr(x), _lib.dptr(rhs), tol, restart, precision,
                                                      C.byref(conv), _lib.dptr(stats))
        if st not in (0, _lib.E_NOT_CONVERGED):
            self.solver._check(st)
        return x, bool(conv.value), {"restarts": int(stats[0]), "rho": float(stats[1]), "spmv": int(stats[2])}

    PRECONDITIONERS = {"none": 0, "diag": 1, "lusgs": 2, "parallel_lusgs": 3}

    def gmres(self, rhs, tol: float, restart: int = 30, precond: str = "none", precision: int = 0):
        """bcsr_gmres ("none"), bcsr_gmres_diag ("diag"), bcsr_gmres_lusgs
        ("lusgs") or bcsr_gmres_parallel_lusgs ("parallel_lusgs", colour sets
        from set_colors) from x = 0.  Returns (x, conv, stats) as gmres_diag."""
        rhs = self._vec(rhs)
        x = np.zeros(self.n)
        conv = C.c_int32(0)
        stats = np.zeros(3)
        st = _lib.lib().subfv_bcsr_gmres(self.h, _lib.dptr(x), _lib.dptr(rhs), tol, restart,
                                         self.PRECONDITIONERS[precond], precision, C.byref(conv), _lib.dptr(stats))
        if st not in (0, _lib.E_NOT_CONVERGED):
            self.solver._check(st)
        return x, bool(conv.value), {"restarts": int(stats[0]), "rho": float(stats[1]), "spmv": int(stats[2])}

    def lusgs(self, rhs):
        """bcsr_lusgs: (L + I) y = rhs, then (L + D) x = y."""
        rhs = self._vec(rhs)
        x = np.zeros(self.n)
        self.solver._check(_lib.lib().subfv_bcsr_lusgs(self.h, _lib.dptr(x), _lib.dptr(rhs)))
        return x

    def set_colors(self, color=None) -> int:
        """Colour sets of bcsr_parallel_lusgs: a 1-based colour per block row,
        or None for the library's greedy colouring.  Returns the number of sets."""
        n = C.c_int32(0)
        if color is None:
            st = _lib.lib().subfv_bcsr_set_colors(self.h, None, C.byref(n))
        else:
            self._color = _lib.i32(color)
            if self._color.size != self.m:
                raise ValueError("one colour per block row")
            st = _lib.lib().subfv_bcsr_set_colors(self.h, _lib.iptr(self._color), C.byref(n))
        self.solver._check(st)
        return int(n.value)

    def parallel_lusgs(self, rhs):
        """bcsr_parallel_lusgs with the colour sets of set_colors."""
        rhs = self._vec(rhs)
        x = np.zeros(self.n)
        self.solver._check(_lib.lib().subfv_bcsr_parallel_lusgs(self.h, _lib.dptr(x), _lib.dptr(rhs)))
        return x

    def pcgm(self, rhs, tol: float, x0=None):
        """bcsr_pcgm from x0 (zero by default).  Returns (x, conv, stats);
        a singular diagonal block raises SubfvError."""
        rhs = self._vec(rhs)
        x = np.zeros(self.n) if x0 is None else self._vec(x0).copy()
        stats = np.zeros(2)
        st = _lib.lib().subfv_bcsr_pcgm(self.h, _lib.dptr(x), _lib.dptr(rhs), tol, _lib.dptr(stats))
        if st not in (0, _lib.E_NOT_CONVERGED):
            self.solver._check(st)
        return x, st == 0, {"iterations": int(stats[0]), "rho": float(stats[1])}

    def add_block(self, i: int, j: int, M):
        """add_block_to_bcsr: block (i, j) (1-based) += M (b x b)."""
        M = np.ascontiguousarray(np.asarray(M, np.float64).T).ravel()
        if M.size != self.b * self.b:
            raise ValueError("M must be b x b")
        self.solver._check(_lib.lib().subfv_bcsr_add_block(self.h, int(i), int(j), _lib.dptr(M)))

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().subfv_bcsr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
