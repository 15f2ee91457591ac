"""Viscous terms on the device (kernels_diff.hpp) against the oracle
(om_diffusion.c).  The device forms the nodal products once per mesh and
adds each cell's terms as one sum per vertex, where the reference adds them
one by one; results agree to round-off, not bit for bit: the residual within
1e-12 of the viscous terms' scale (measured: 2e-14 on tri 8^2, 4e-17 for the
point-implicit diagonal) and the state after 100 viscous steps within 1e-12
relative (measured: 1e-15 on tri, quad, 3-D hex and a 10-sub-element fan).
Tetrahedral meshes (36 sub-faces and 24 sub-elements at an interior vertex)
take the large-class kernels: setup matrices in a global workspace, one
vertex per apply workgroup."""
import numpy as np
import pytest

import oracle as O
from subfv_amd import meshgen as mg
from subfv_amd.api import Solver, SubfvError, make_config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    s = Solver(0)
    yield s
    s.close()


def smooth_state(xc, seed=5):
    rng = np.random.default_rng(seed)
    x, y, z = xc[:, 0], xc[:, 1], xc[:, 2]
    w = np.stack([1.0 + 0.2 * np.sin(3 * x + 1) * np.cos(2 * y) + 0.02 * rng.random(x.size),
                  0.3 * np.sin(2 * y + 0.5), -0.2 * np.cos(3 * x), 0.05 * np.sin(4 * z + x),
                  1.0 + 0.3 * np.cos(2 * x + y) + 0.02 * rng.random(x.size)], 1)
    return mg.primit_to_conserv(w)


CASES = {
    "quad9": (lambda: mg.quad_mesh(9), dict(), dict()),
    "tri7": (lambda: mg.tri_mesh(7), dict(), dict()),
    "quad8lr_dirichlet": (lambda: mg.quad_mesh(8, tags="lr"),
                          dict(bcs=[("freestream", (1.0, 0.2, 0.0, 0.0, 1.0)), ("outflowsupersonic", (0, 0, 0, 0, 0))]),
                          dict(bcs_v=[("dirichlet", (0.1, -0.05, 0.0)), ("neumann", (0, 0, 0))],
                               bcs_t=[("dirichlet", 1.5e-3), ("neumann", 0.0)])),
    "tet3": (lambda: mg.tet_mesh(3), dict(boundary_2d=False), dict()),  # large size class
}


@pytest.mark.parametrize("second", [False, True])
@pytest.mark.parametrize("name", list(CASES))
def test_rhs_and_diag_match_oracle(solver, name, second):
    print(f"[gpu] diffusion rhs {name} second={second}", flush=True)
    mk, cfgkw, dkw = CASES[name]
    mesh = mk()
    solver.build_mesh(mesh)
    om = O.OracleMesh(mesh)
    sol = smooth_state(om.elem_coord)
    mu = 2e-3
    solver.set_config(make_config("multi_point", second_order=second, method=4, **cfgkw))
    solver.set_diffusion(True, mu=mu, **dkw)
    solver.put_sol(sol)
    solver.compute_rhs_ns()
    rhs = solver.get("rhs")
    diag = solver.get("diag")
    ocfg = O.make_config("multi_point", second_order=second, method=4, diffusion=True, mu=mu, **cfgkw, **dkw)
    r0, sl = om.compute_euler_flux(ocfg, sol)
    if second:
        g = om.compute_cell_grad(sol, 4)[0]
        r2, _ = om.compute_euler_flux(ocfg, sol, grad=g, second_order=True)
        r0 = om.flux_limiting(sol, r0, r2, sl, ocfg.cfl)
    rd, dd = om.add_diffusion(ocfg, sol)
    ref_rhs = r0 + rd
    scale_r = np.abs(rd).max(axis=0) + 1e-300
    scale_d = np.abs(dd).max()
    # Euler part bit-exact; the viscous part to 1e-12 of its scale
    assert np.all(np.abs(rhs - ref_rhs) <= 1e-12 * scale_r[None, :] + 1e-14 * np.abs(ref_rhs)), \
        np.abs(rhs - ref_rhs).max(axis=0) / scale_r
    assert np.abs(diag - dd).max() <= 1e-12 * scale_d
    solver.set_diffusion(False)


@pytest.mark.parametrize("scheme,second,ts", [("multi_point", False, 0), ("multi_point", True, 0),
                                              ("two_wave", True, 1)])
def test_viscous_time_loop_matches_oracle(solver, scheme, second, ts):
    print(f"[gpu] diffusion run {scheme} second={second} ts={ts}", flush=True)
    mesh = mg.tri_mesh(8)
    solver.build_mesh(mesh)
    om = O.OracleMesh(mesh)
    sol = smooth_state(om.elem_coord, 9)
    mu = 5e-3
    solver.set_config(make_config(scheme, second_order=second, method=4, cfl=0.3))
    solver.set_diffusion(True, mu=mu)
    solver.set_time(0.0)
    solver.put_sol(sol)
    solver.run(10, ts)
    got = solver.get_sol()
    st, ref, info = om.run(O.make_config(scheme, second_order=second, method=4, cfl=0.3, diffusion=True, mu=mu),
                           sol, 10, ts)
    assert st == 0
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    assert rel <= 1e-12, rel
    solver.set_diffusion(False)


def fan_prism_mesh(k=5, layers=2):
    """k prisms around a vertical axis, extruded in `layers` layers: the axis
    vertices between layers have 2k sub-elements and 3k sub-faces (10 and
    15 for k = 5), more than the viscous apply stages in LDS."""
    ang = 2 * np.pi * np.arange(k) / k + 0.1
    pts = np.concatenate([[[0.0, 0.0]], np.stack([np.cos(ang), np.sin(ang)], 1)])
    coord = np.concatenate([np.column_stack([pts, np.full(len(pts), 0.4 * z)]) for z in range(layers + 1)])
    ev = []
    for lay in range(layers):
        o, o2 = lay * len(pts), (lay + 1) * len(pts)
        for q in range(k):
            a, b, c = 0, 1 + q, 1 + (q + 1) % k
            ev.append([o + a, o + b, o + c, o2 + a, o2 + b, o2 + c])
    ev = np.array(ev) + 1
    return mg.Mesh(coord, np.full(len(ev), 6, np.int32), ev.astype(np.int32).ravel(), boundary_2d=False, name="fan")


LONG_CASES = {
    "tri8": lambda: mg.tri_mesh(8),
    "hex3": lambda: mg.hex_mesh(3),  # 3-D interior vertices: 12 sub-faces, 8 sub-elements
    "fan5": fan_prism_mesh,          # 10 sub-elements: slot terms read from HBM
    "tet3": lambda: mg.tet_mesh(3),  # 36 sub-faces, 24 sub-elements: the large size class
}


@pytest.mark.parametrize("ts", [0, 1])
@pytest.mark.parametrize("name", list(LONG_CASES))
def test_viscous_100_steps_within_1e12(solver, name, ts):
    """100 second-order viscous steps (the reference's two-stage loop and
    SSP-RK3): the conserved states stay within 1e-12 relative of the
    oracle's, on a prism mesh, a 3-D hex mesh and a mesh whose vertices
    exceed the LDS staging of the viscous apply."""
    print(f"[gpu] viscous 100 steps {name} ts={ts}", flush=True)
    mesh = LONG_CASES[name]()
    solver.build_mesh(mesh)
    om = O.OracleMesh(mesh)
    sol = smooth_state(om.elem_coord, 9)
    kw = dict(boundary_2d=mesh.boundary_2d, second_order=True, method=4, cfl=0.3)
    solver.set_config(make_config("multi_point", **kw))
    solver.set_diffusion(True, mu=5e-3)
    solver.set_time(0.0)
    solver.put_sol(sol)
    solver.run(100, ts)
    got = solver.get_sol()
    solver.set_diffusion(False)
    st, ref, info = om.run(O.make_config("multi_point", diffusion=True, mu=5e-3, **kw), sol, 100, ts)
    assert st == 0
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    assert rel <= 1e-12, rel


def test_vertices_beyond_the_limit_are_rejected(solver):
    print("[gpu] diffusion limit", flush=True)
    solver.build_mesh(fan_prism_mesh(k=40))  # the axis vertex between the layers has 80 sub-elements
    solver.set_config(make_config("multi_point", boundary_2d=False))
    solver.set_diffusion(True, mu=1e-3)
    solver.put_sol(mg.primit_to_conserv(np.tile([1.0, 0, 0, 0, 1.0], (solver.n_elems, 1))))
    with pytest.raises(SubfvError):
        solver.compute_rhs_ns()
    solver.set_diffusion(False)


def warm_state(xc, seed=3):
    """air-like temperatures (250-350 K) so Sutherland's mu varies by ~30 %"""
    rng = np.random.default_rng(seed)
    x, y = xc[:, 0], xc[:, 1]
    rho = 1.0 + 0.2 * np.sin(3 * x + 1) * np.cos(2 * y) + 0.02 * rng.random(x.size)
    T = 300.0 + 50.0 * np.sin(2 * np.pi * x) * np.cos(np.pi * y) + rng.random(x.size)
    w = np.stack([rho, 30 * np.sin(2 * y + 0.5), -20 * np.cos(3 * x), 5 * np.sin(4 * xc[:, 2] + x),
                  rho * 0.4 * 720.19471 * T], 1)
    return mg.primit_to_conserv(w)


@pytest.mark.parametrize("second", [False, True])
@pytest.mark.parametrize("name", list(CASES))
def test_sutherland_rhs_and_diag_match_oracle(solver, name, second):
    """compute_mu with use_sutherland: mu per cell, mu0corr their minimum;
    T^1.5 goes through pow on both sides, so the per-cell viscosities may
    differ in the last bit: the same 1e-12 bound as the constant case."""
    print(f"[gpu] sutherland rhs {name} second={second}", flush=True)
    mk, cfgkw, dkw = CASES[name]
    mesh = mk()
    solver.build_mesh(mesh)
    om = O.OracleMesh(mesh)
    sol = warm_state(om.elem_coord)
    solver.set_config(make_config("multi_point", second_order=second, method=4, **cfgkw))
    solver.set_diffusion(True, mu=1.0, **dkw)  # mu_p unused with Sutherland
    solver.set_sutherland(True)
    try:
        solver.put_sol(sol)
        solver.compute_rhs_ns()
        rhs = solver.get("rhs")
        diag = solver.get("diag")
    finally:
        solver.set_sutherland(False)
        solver.set_diffusion(False)
    ocfg = O.make_config("multi_point", second_order=second, method=4, diffusion=True, mu=1.0, sutherland=True,
                         **cfgkw, **dkw)
    mu, mn = O.compute_mu(ocfg, sol)
    assert mu.max() > 1.2 * mn
    r0, sl = om.compute_euler_flux(ocfg, sol)
    if second:
        g = om.compute_cell_grad(sol, 4)[0]
        r2, _ = om.compute_euler_flux(ocfg, sol, grad=g, second_order=True)
        r0 = om.flux_limiting(sol, r0, r2, sl, ocfg.cfl)
    rd, dd = om.add_diffusion(ocfg, sol)
    ref_rhs = r0 + rd
    scale_r = np.abs(rd).max(axis=0) + 1e-300
    assert np.all(np.abs(rhs - ref_rhs) <= 1e-12 * scale_r[None, :] + 1e-14 * np.abs(ref_rhs)), \
        np.abs(rhs - ref_rhs).max(axis=0) / scale_r
    assert np.abs(diag - dd).max() <= 1e-12 * np.abs(dd).max()


def test_sutherland_time_loop_matches_oracle(solver):
    print("[gpu] sutherland run", flush=True)
    mesh = mg.tri_mesh(8)
    solver.build_mesh(mesh)
    om = O.OracleMesh(mesh)
    sol = warm_state(om.elem_coord, 4)
    solver.set_config(make_config("multi_point", second_order=True, method=4, cfl=0.3))
    solver.set_diffusion(True, mu=1.0)
    solver.set_sutherland(True, mu0=1.716e-5, T0=273.15)
    solver.set_time(0.0)
    try:
        solver.put_sol(sol)
        solver.run(10, 0)
        got = solver.get_sol()
    finally:
        solver.set_sutherland(False)
        solver.set_diffusion(False)
    st, ref, info = om.run(O.make_config("multi_point", second_order=True, method=4, cfl=0.3, diffusion=True,
                                         mu=1.0, sutherland=True), sol, 10, 0)
    assert st == 0
    assert np.abs(got - ref).max() / np.abs(ref).max() <= 1e-12
    # the viscous terms moved the state: the same run without them differs
    st, inv, _ = om.run(O.make_config("multi_point", second_order=True, method=4, cfl=0.3), sol, 10, 0)
    assert np.abs(got - inv).max() / np.abs(ref).max() > 1e-9


def test_sutherland_off_restores_constant_viscosity(solver):
    """the blocks formed per residual under Sutherland are replaced by the
    constant-viscosity ones when it is switched off: same bits as before"""
    print("[gpu] sutherland toggle", flush=True)
    mesh = mg.quad_mesh(9)
    solver.build_mesh(mesh)
    sol = warm_state(solver.elem_coord, 6)
    solver.set_config(make_config("multi_point"))
    solver.set_diffusion(True, mu=2e-5)
    solver.put_sol(sol)
    solver.compute_rhs_ns()
    a = solver.get("rhs"), solver.get("diag")
    solver.set_sutherland(True)
    solver.compute_rhs_ns()
    b = solver.get("rhs"), solver.get("diag")
    solver.set_sutherland(False)
    solver.set_diffusion(True, mu=2e-5)
    solver.compute_rhs_ns()
    c = solver.get("rhs"), solver.get("diag")
    solver.set_diffusion(False)
    assert not np.array_equal(a[1], b[1])
    assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_viscous_parameter_changes_reuse_the_buffers(solver):
    """Changing mu, toggling Sutherland's law and switching the viscous terms
    off and on re-form the blocks in the buffers of the first viscous
    residual: the context's device memory stays flat."""
    print("[gpu] diffusion parameter sweep memory", flush=True)
    mesh = mg.tri_mesh(12)
    solver.build_mesh(mesh)
    sol = warm_state(solver.elem_coord, 2)
    solver.set_config(make_config("multi_point"))
    solver.put_sol(sol)
    solver.set_diffusion(True, mu=2e-5)
    solver.compute_rhs_ns()
    first = solver.get("rhs")
    solver.set_sutherland(True)  # the per-cell viscosities join the buffers once
    solver.compute_rhs_ns()
    solver.set_sutherland(False)
    used = solver.count("device_bytes_used")
    for k in range(6):
        solver.set_diffusion(True, mu=2e-5 * (1 + k))
        solver.set_sutherland(k % 2 == 1)
        solver.compute_rhs_ns()
        solver.set_diffusion(False)
        solver.compute_rhs_ns()
    solver.set_sutherland(False)
    solver.set_diffusion(True, mu=2e-5)
    solver.compute_rhs_ns()
    assert solver.count("device_bytes_used") == used
    assert np.array_equal(solver.get("rhs"), first)
    solver.set_diffusion(False)


@pytest.mark.parametrize("name", list(CASES))
def test_reference_builders_for_given_viscosities(solver, name):
    """build_momentum_diff_mat_into_global_mat_diag (parts 1) and
    build_diff_mat_into_global_mat_diag (parts 2) for cell viscosities the
    caller gives (varying by a factor 3), each added to a nonzero rhs and
    diag as the reference's builders add them, against the oracle's builders;
    parts 3 is both.  Within 1e-12 of the viscous terms' scale.  A residual
    afterwards uses the constant viscosity again."""
    print(f"[gpu] reference builders {name}", flush=True)
    mk, cfgkw, dkw = CASES[name]
    mesh = mk()
    solver.build_mesh(mesh)
    om = O.OracleMesh(mesh)
    ne = om.n_elems
    sol = smooth_state(om.elem_coord)
    rng = np.random.default_rng(11)
    mu = 2e-3 * (0.5 + rng.random(ne))
    mu0 = float(mu.min())
    rhs0 = rng.standard_normal((ne, 5))
    diag0 = rng.standard_normal((ne, 5, 5))
    solver.set_config(make_config("multi_point", **cfgkw))
    solver.set_diffusion(True, mu=2e-3, **dkw)
    ocfg = O.make_config("multi_point", diffusion=True, mu=2e-3, **cfgkw, **dkw)
    try:
        for parts in (1, 2, 3):
            r, d = solver.add_diffusion(sol, mu, mu0, parts, rhs0, diag0)
            ro, do = om.add_diffusion_mu(ocfg, sol, mu, mu0, parts, rhs0, diag0)
            zr, zd = om.add_diffusion_mu(ocfg, sol, mu, mu0, parts)
            scale_r = np.abs(zr).max() + 1e-300
            scale_d = np.abs(zd).max() + 1e-300
            # the terms land in rhs0 and diag0 in another order: within 1e-12
            # of their scale plus the rounding of those sums
            assert np.all(np.abs(r - ro) <= 1e-12 * scale_r + 1e-14 * np.abs(ro)), (parts, np.abs(r - ro).max())
            assert np.all(np.abs(d - do) <= 1e-12 * scale_d + 1e-14 * np.abs(do)), (parts, np.abs(d - do).max())
            # the rows and blocks a builder does not touch stay as given
            assert np.array_equal(r[:, 0], rhs0[:, 0])
            # (blocks column-major: d[c, j, i] is diag(i + 1, j + 1, c + 1))
            if parts == 1:
                assert np.array_equal(d[:, :, 4], diag0[:, :, 4])  # the heat row diag(5, :)
            if parts == 2:
                assert np.array_equal(r[:, 1:4], rhs0[:, 1:4])
                assert np.array_equal(d[:, 1:4, 1:4], diag0[:, 1:4, 1:4])
        # compute_mu without Sutherland: mu_p everywhere
        m, m0 = solver.compute_mu(sol)
        assert np.all(m == 2e-3) and m0 == 2e-3
        # the context's own residual uses mu_p again
        solver.put_sol(sol)
        solver.compute_rhs_ns()
        rd, dd = om.add_diffusion(ocfg, sol)
        r0, _ = om.compute_euler_flux(ocfg, sol)
        scale_r = np.abs(rd).max(axis=0) + 1e-300
        assert np.all(np.abs(solver.get("rhs") - (r0 + rd)) <= 1e-12 * scale_r[None, :] + 1e-14 * np.abs(r0 + rd))
    finally:
        solver.set_diffusion(False)


def test_compute_mu_sutherland_matches_oracle(solver):
    """compute_mu with use_sutherland: per-cell viscosities and mu0corr (their
    minimum); T^1.5 goes through pow on both sides: within 4 ulp."""
    print("[gpu] compute_mu sutherland", flush=True)
    mesh = mg.tri_mesh(9)
    solver.build_mesh(mesh)
    om = O.OracleMesh(mesh)
    sol = warm_state(om.elem_coord, 3)
    solver.set_config(make_config("multi_point"))
    solver.set_diffusion(True, mu=1.0)
    solver.set_sutherland(True)
    try:
        m, m0 = solver.compute_mu(sol)
        # the builders on these viscosities: the context's Sutherland residual
        r, d = solver.add_diffusion(sol, m, m0, 3)
    finally:
        solver.set_sutherland(False)
        solver.set_diffusion(False)
    ocfg = O.make_config("multi_point", diffusion=True, mu=1.0, sutherland=True)
    mo, m0o = O.compute_mu(ocfg, sol)
    assert mo.max() > 1.2 * m0o
    assert np.abs(m - mo).max() <= 4 * np.spacing(mo).max()
    assert abs(m0 - m0o) <= 4 * np.spacing(m0o)
    rd, dd = om.add_diffusion(ocfg, sol)
    assert np.abs(r - rd).max() <= 1e-12 * np.abs(rd).max()
    assert np.abs(d - dd).max() <= 1e-12 * np.abs(dd).max()
