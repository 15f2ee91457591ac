"""The implicit low-Mach test system (subfv_amd.lowmach), checked on the CPU:
flux Jacobian against finite differences of the flux, the assembled BCSR
operator against a finite-difference linearisation of the residual with
frozen dissipation speeds, and the oracle's BCSR SpMV / block-Jacobi / GMRES
on it."""
import numpy as np

import oracle as O
from subfv_amd import lowmach as lm


def test_flux_jacobian_matches_finite_differences():
    rng = np.random.default_rng(4)
    U = np.stack([1 + 0.1 * rng.random(8), 0.1 * rng.standard_normal(8), 0.1 * rng.standard_normal(8),
                  3 + rng.random(8)], -1)
    th = rng.random(8) * 2 * np.pi
    n = np.stack([np.cos(th), np.sin(th)], -1)
    A = lm.flux_jacobian(U, n)
    eps = 1e-6
    for k in range(4):
        d = np.zeros(4)
        d[k] = eps
        fd = (lm.flux_n(U + d, n) - lm.flux_n(U - d, n)) / (2 * eps)
        assert np.allclose(A[..., :, k], fd, rtol=1e-7, atol=1e-8)


def _dense_apply(sys, x):
    return O.bcsr_matmul(sys["m"], 4, sys["row"], sys["col"], sys["v"], x)


def test_bcsr_operator_is_the_frozen_speed_linearisation():
    s = lm.build_system(12, mach=0.1, cfl=2.0)
    m, n = s["m"], s["n"]
    assert s["row"][-1] - 1 == 5 * m  # five blocks per row
    U, geo = s["U"], s["geo"]
    x = np.random.default_rng(1).standard_normal(U.shape)
    eps = 1e-6 * np.abs(U).max()
    fd = (lm.residual(U + eps * x, geo, s["lam"]) - lm.residual(U - eps * x, geo, s["lam"])) / (2 * eps)
    area = geo[0]
    expect = (area / s["dt"])[..., None] * x + fd
    got = _dense_apply(s, x.reshape(-1)).reshape(n, n, 4)
    assert np.allclose(got, expect, rtol=1e-6, atol=1e-6 * np.abs(expect).max())


def test_oracle_gmres_solves_the_low_mach_system():
    s = lm.build_system(24, mach=0.01, cfl=5.0)
    rhs = s["rhs"]
    tol = 1e-10 * np.linalg.norm(rhs)
    x, conv, info = O.bcsr_gmres_diag(s["m"], 4, s["row"], s["col"], s["v"], rhs, tol, 30)
    assert conv, info
    assert np.linalg.norm(rhs - _dense_apply(s, x)) < tol
