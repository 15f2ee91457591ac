"""Implicit low-Mach system: a backward-Euler Jacobian in BCSR form.

The linear systems the BCSR / block-Jacobi / GMRES routines
(src/linear_algebra/sparse_bcsr_linear_module.F90) serve are built here: one
backward-Euler step of a first-order finite-volume discretisation of the 2-D
Euler equations (conservative variables rho, rho u, rho v, E: 4x4 blocks) on
an n x n jittered quad grid, periodic in x and y, around a seeded Mach-0.01
vortex.  Interface fluxes are Rusanov (local Lax-Friedrichs):

    F(UL, UR, n) = 1/2 (Fn(UL) + Fn(UR)) - 1/2 lam (UR - UL),
    lam = max(|u_L.n| + c_L, |u_R.n| + c_R)  (frozen in the linearisation),

so block row i of  J = vol_i/dt I + dR/dU  holds
    diagonal:  vol_i/dt I + sum_faces area/2 (A_n(U_i) + lam I),
    neighbour: area/2 (A_n(U_j) - lam I)
with A_n = dFn/dU the analytic Euler flux Jacobian.  Each row has 5 blocks
(itself and its four face neighbours).  The right-hand side is -R(U).

Node jitter uses the same counter-based generator as the explicit meshes
(meshgen.uniform), so the system is identical wherever it is built.
"""
from __future__ import annotations

import numpy as np

from .meshgen import uniform

GAMMA = 1.4


def periodic_grid(n: int, jitter: float = 0.2, seed: int = 20261202):
    """Node coordinates (n, n, 2) of a jittered periodic grid of [-0.5, 0.5]^2
    (node (i, j) is shared with (i + n, j) and (i, j + n))."""
    h = 1.0 / n
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ids = (i * n + j).ravel().astype(np.uint64)
    dx = (2.0 * uniform(seed, 11, ids) - 1.0) * jitter * h
    dy = (2.0 * uniform(seed, 12, ids) - 1.0) * jitter * h
    x = -0.5 + i * h + dx.reshape(n, n)
    y = -0.5 + j * h + dy.reshape(n, n)
    return np.stack([x, y], -1)


def _node(X, i, j):
    """Node (i, j) with periodic images shifted by whole periods."""
    n = X.shape[0]
    p = X[i % n, j % n].copy()
    p[..., 0] += np.floor_divide(i, n)
    p[..., 1] += np.floor_divide(j, n)
    return p


def geometry(X):
    """Cell areas and, for the +x (east) and +y (north) faces of each cell,
    the outward unit normal and length.  Cell (i, j) has corners
    (i, j), (i+1, j), (i+1, j+1), (i, j+1)."""
    n = X.shape[0]
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = _node(X, i, j)
    b = _node(X, i + 1, j)
    c = _node(X, i + 1, j + 1)
    d = _node(X, i, j + 1)
    cross = lambda p, q: p[..., 0] * q[..., 1] - p[..., 1] * q[..., 0]
    area = 0.5 * (cross(a, b) + cross(b, c) + cross(c, d) + cross(d, a))
    # east face b->c, outward normal (dy, -dx); north face c->d
    e = c - b
    ne = np.stack([e[..., 1], -e[..., 0]], -1)
    le = np.linalg.norm(ne, axis=-1)
    f = d - c
    nn = np.stack([f[..., 1], -f[..., 0]], -1)
    ln = np.linalg.norm(nn, axis=-1)
    centre = 0.25 * (a + b + c + d)
    return area, ne / le[..., None], le, nn / ln[..., None], ln, centre


def vortex_state(centre, mach: float = 0.01, seed: int = 20261202):
    """rho = 1, a Gresho-type vortex (steady for the Euler equations) centred at
    a seeded point, scaled so that max |u| / c = mach."""
    off = uniform(seed, 13, np.arange(2, dtype=np.uint64)) * 0.2 - 0.1
    x = centre[..., 0] - off[0]
    y = centre[..., 1] - off[1]
    x = x - np.round(x)
    y = y - np.round(y)
    r = np.sqrt(x * x + y * y)
    ut = np.where(r < 0.2, 5 * r, np.where(r < 0.4, 2 - 5 * r, 0.0))
    dp = np.where(r < 0.2, 12.5 * r * r, np.where(r < 0.4, 4 - 4 * np.log(0.2) + 12.5 * r * r - 20 * r + 4 * np.log(np.maximum(r, 1e-300)), 4 - 4 * np.log(0.2) + 12.5 * 0.16 - 8 + 4 * np.log(0.4)))
    p0 = 1.0 / (GAMMA * mach * mach)  # c = 1 / mach for the peak speed 1
    rho = np.ones_like(r)
    p = p0 + dp
    with np.errstate(invalid="ignore", divide="ignore"):
        u = np.where(r > 0, -ut * y / r, 0.0)
        v = np.where(r > 0, ut * x / r, 0.0)
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], -1)


def flux_n(U, n):
    rho, mx, my, E = U[..., 0], U[..., 1], U[..., 2], U[..., 3]
    u, v = mx / rho, my / rho
    p = (GAMMA - 1) * (E - 0.5 * rho * (u * u + v * v))
    vn = u * n[..., 0] + v * n[..., 1]
    return np.stack([rho * vn, mx * vn + p * n[..., 0], my * vn + p * n[..., 1], (E + p) * vn], -1)


def wave_speed(U, n):
    rho, mx, my, E = U[..., 0], U[..., 1], U[..., 2], U[..., 3]
    u, v = mx / rho, my / rho
    p = (GAMMA - 1) * (E - 0.5 * rho * (u * u + v * v))
    return np.abs(u * n[..., 0] + v * n[..., 1]) + np.sqrt(GAMMA * p / rho)


def flux_jacobian(U, n):
    """A_n = d(F.n)/dU, shape (..., 4, 4), row = flux component."""
    g1 = GAMMA - 1
    rho, mx, my, E = U[..., 0], U[..., 1], U[..., 2], U[..., 3]
    u, v = mx / rho, my / rho
    nx, ny = n[..., 0], n[..., 1]
    vn = u * nx + v * ny
    q2 = u * u + v * v
    p = g1 * (E - 0.5 * rho * q2)
    H = (E + p) / rho
    phi = 0.5 * g1 * q2
    A = np.zeros(U.shape[:-1] + (4, 4))
    A[..., 0, 1] = nx
    A[..., 0, 2] = ny
    A[..., 1, 0] = phi * nx - u * vn
    A[..., 1, 1] = vn + u * nx - g1 * u * nx
    A[..., 1, 2] = u * ny - g1 * v * nx
    A[..., 1, 3] = g1 * nx
    A[..., 2, 0] = phi * ny - v * vn
    A[..., 2, 1] = v * nx - g1 * u * ny
    A[..., 2, 2] = vn + v * ny - g1 * v * ny
    A[..., 2, 3] = g1 * ny
    A[..., 3, 0] = vn * (phi - H)
    A[..., 3, 1] = H * nx - g1 * u * vn
    A[..., 3, 2] = H * ny - g1 * v * vn
    A[..., 3, 3] = GAMMA * vn
    return A


def residual(U, geo, lam=None):
    """R(U) = sum over the cell's faces of area * Rusanov flux (outward).
    lam = (lam_east, lam_north) freezes the dissipation speeds (the
    linearisation used for the Jacobian)."""
    area, ne, le, nn, ln, _ = geo
    Ue = np.roll(U, -1, axis=0)  # east neighbour (i+1, j)
    Un = np.roll(U, -1, axis=1)  # north neighbour (i, j+1)

    def rus(UL, UR, nrm, lm):
        if lm is None:
            lm = np.maximum(wave_speed(UL, nrm), wave_speed(UR, nrm))
        return 0.5 * (flux_n(UL, nrm) + flux_n(UR, nrm)) - 0.5 * lm[..., None] * (UR - UL)

    fe = rus(U, Ue, ne, None if lam is None else lam[0]) * le[..., None]
    fn = rus(U, Un, nn, None if lam is None else lam[1]) * ln[..., None]
    R = fe + fn
    R = R - np.roll(fe, 1, axis=0) - np.roll(fn, 1, axis=1)
    return R


def build_system(n: int = 512, mach: float = 0.01, cfl: float = 5.0, seed: int = 20261202):
    """Backward-Euler system J dU = -R(U) in BCSR form.

    Returns dict(m, b=4, row, col (1-based), v (column-major blocks, flat),
    rhs, dt, U).  dt = cfl * h / max wave speed (acoustic CFL)."""
    X = periodic_grid(n, seed=seed)
    geo = geometry(X)
    area, ne, le, nn, ln, centre = geo
    U = vortex_state(centre, mach, seed)
    h = 1.0 / n
    smax = max(wave_speed(U, ne).max(), wave_speed(U, nn).max())
    dt = cfl * h / smax
    Ue = np.roll(U, -1, axis=0)
    Un = np.roll(U, -1, axis=1)
    lam_e = np.maximum(wave_speed(U, ne), wave_speed(Ue, ne))
    lam_n = np.maximum(wave_speed(U, nn), wave_speed(Un, nn))
    I4 = np.eye(4)
    # contributions of the east face of (i, j): to row (i,j) [+] and row (i+1,j) [-]
    Ae_own = flux_jacobian(U, ne)      # dF_e/dU_(i,j) uses A_n(U_L) with n = ne
    Ae_nbr = flux_jacobian(Ue, ne)     # dF_e/dU_(i+1,j)
    An_own = flux_jacobian(U, nn)
    An_nbr = flux_jacobian(Un, nn)
    he = 0.5 * le[..., None, None]
    hn = 0.5 * ln[..., None, None]
    # dF_e/dU_L = 1/2 (A(U_L) + lam I), dF_e/dU_R = 1/2 (A(U_R) - lam I)
    dFe_L = he * (Ae_own + lam_e[..., None, None] * I4)
    dFe_R = he * (Ae_nbr - lam_e[..., None, None] * I4)
    dFn_L = hn * (An_own + lam_n[..., None, None] * I4)
    dFn_R = hn * (An_nbr - lam_n[..., None, None] * I4)
    # row (i,j): + dF_e(i,j) + dF_n(i,j) - dF_e(i-1,j) - dF_n(i,j-1)
    diag = (area / dt)[..., None, None] * I4 + dFe_L + dFn_L - np.roll(dFe_R, 1, axis=0) - np.roll(dFn_R, 1, axis=1)
    east = dFe_R                           # column (i+1, j)
    north = dFn_R                          # column (i, j+1)
    west = -np.roll(dFe_L, 1, axis=0)      # column (i-1, j)
    south = -np.roll(dFn_L, 1, axis=1)     # column (i, j-1)
    m = n * n
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    rid = (ii * n + jj).ravel()
    cols = np.stack([rid, (((ii + 1) % n) * n + jj).ravel(), ((ii * n + (jj + 1) % n)).ravel(),
                     (((ii - 1) % n) * n + jj).ravel(), ((ii * n + (jj - 1) % n)).ravel()], 1)
    blocks = np.stack([diag.reshape(m, 4, 4), east.reshape(m, 4, 4), north.reshape(m, 4, 4),
                       west.reshape(m, 4, 4), south.reshape(m, 4, 4)], 1)
    order = np.argsort(cols, axis=1, kind="stable")
    cols = np.take_along_axis(cols, order, 1)
    blocks = np.take_along_axis(blocks, order[..., None, None], 1)
    row = (np.arange(m + 1) * 5 + 1).astype(np.int32)
    col = (cols.ravel() + 1).astype(np.int32)
    v = np.ascontiguousarray(np.swapaxes(blocks.reshape(-1, 4, 4), 1, 2)).ravel()  # column-major
    rhs = -residual(U, geo).reshape(-1)
    return {"m": m, "b": 4, "row": row, "col": col, "v": v, "rhs": rhs, "dt": dt, "U": U, "geo": geo,
            "n": n, "lam": (lam_e, lam_n)}
