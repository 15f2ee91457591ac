"""The reference-signature Fortran procedures (subfv_amd/fortran/
subfv_reference_api.f90) driven as the reference's own driver calls them:
compute_euler_flux (first and second order), compute_cell_grad_from_nodal_grad,
flux_limiting and add_timedisc_and_solve on one residual and stage update
(ns_main.F90:130-179, ns_euler_module.F90:521-609), then bcsr_matmul,
bcsr_diag and bcsr_gmres_diag on a block matrix whose values are refilled in
place between calls.  Results equal the oracle's bit for bit (GMRES: to
its tolerance)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from subfv_amd import build as B
from subfv_amd import meshgen as mg

EXE = os.path.join(B.FBUILD, "reference_api_check")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built (no flang)")]


def _bcsr(m, b, seed):
    rng = np.random.default_rng(seed)
    row, cols, vals = [1], [], []
    for i in range(m):
        nb = sorted(set([i] + list(rng.integers(0, m, 4))))
        for j in nb:
            blk = rng.standard_normal((b, b)) * 0.3 + (2.0 * b * np.eye(b) if j == i else 0.0)
            cols.append(j + 1)
            vals.append(blk.T.reshape(-1))
        row.append(row[-1] + len(nb))
    return np.array(row, np.int32), np.array(cols, np.int32), np.concatenate(vals)


def test_reference_signature_procedures(tmp_path):
    print("[gpu] reference-signature Fortran procedures", flush=True)
    mesh = mg.quad_mesh(12)
    mg.write_msh(mesh, str(tmp_path / "mesh.msh"))
    m, b = 200, 4
    row, col, v = _bcsr(m, b, 3)
    x = np.random.default_rng(4).standard_normal(m * b)
    tol = 1e-10 * float(np.linalg.norm(x))
    with open(tmp_path / "bcsr.bin", "wb") as f:
        np.array([m, b, len(col)], np.int32).tofile(f)
        row.tofile(f)
        col.tofile(f)
        v.astype(np.float64).tofile(f)
        x.tofile(f)
        np.array([tol]).tofile(f)
    out = subprocess.run([EXE, "mesh.msh", "bcsr.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr

    om = O.OracleMesh(mesh)
    ne, nv = om.n_elems, om.n_vert
    raw = np.fromfile(tmp_path / "residual.bin", dtype=np.uint8)
    assert tuple(raw[:8].view(np.int32)) == (ne, nv)
    d = raw[8:].view(np.float64)
    sizes = [5 * ne, 5 * ne, 15 * ne, 15 * nv, 5 * ne, 1]
    parts = np.split(d, np.cumsum(sizes)[:-1])
    rhs, rhs_o2, grad, ngrad, sol1, dt = (p.copy() for p in parts)
    # the oracle along the same calls; the initial state as the driver forms it
    xc = om.elem_coord
    sol = np.where((xc[:, 0] < 0.5)[:, None], np.array([1.0, 0, 0, 0, 1.0 / 0.4]),
                   np.array([0.125, 0, 0, 0, 0.1 / 0.4]))
    cfg = O.make_config("multi_point", bc_style=2, boundary_2d=True, second_order=True, method=4, cfl=0.4)
    r1, sl = om.compute_euler_flux(cfg, sol)
    g, ng = om.compute_cell_grad(sol, 4)
    r2, sl2 = om.compute_euler_flux(cfg, sol, grad=g, second_order=True)
    rl = om.flux_limiting(sol, r1, r2, sl, 0.4)
    assert np.array_equal(ng.reshape(-1), ngrad), np.abs(ng.reshape(-1) - ngrad).max()
    assert np.array_equal(g.reshape(-1), grad)
    assert np.array_equal(r2.reshape(-1), rhs_o2)
    assert np.array_equal(rl.reshape(-1), rhs)
    ref = om.add_timedisc_and_solve(cfg, sol, rl, sl, 1.0, 1.0, 0.4, 0.0)
    assert ref["status"] == 0
    assert dt[0] == ref["dt"]
    assert np.array_equal(sol1, ref["sol"].reshape(-1)), np.abs(sol1 - ref["sol"].reshape(-1)).max()

    o = np.fromfile(tmp_path / "bcsr_out.bin", dtype=np.uint8)
    n = m * b
    y, xd, xg = (o[:8 * 3 * n].view(np.float64)[k * n:(k + 1) * n] for k in range(3))
    conv = int(o[8 * 3 * n:].view(np.int32)[0])
    assert np.array_equal(y, O.bcsr_matmul(m, b, row, col, v, x))
    assert np.array_equal(xd, O.bcsr_diag(m, b, row, col, v, x))
    # bcsr_gmres_diag ran on the doubled values: A2 x = b gives x / 2
    v2 = 2.0 * v
    assert conv == 1
    xo, conv_o, _ = O.bcsr_gmres_diag(m, b, row, col, v2, x, tol)
    assert conv_o
    r = x - O.bcsr_matmul(m, b, row, col, v2, xg)
    assert np.linalg.norm(r) < tol
    assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo)
