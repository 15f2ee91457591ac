"""Parity of the HIP path (through the C-ABI) with the oracle.

Floating-point results are compared bit for bit: both sides evaluate the
reference expressions in the same order without contraction.  Runs of many
steps are additionally held to the 1e-12 relative bound of the target.
"""
import numpy as np
import pytest

import oracle as O
from subfv_amd import meshgen as mg
from subfv_amd.api import Solver, SubfvError, make_config

pytestmark = pytest.mark.gpu

MESHES = {
    "quad9": lambda: mg.quad_mesh(9),
    "quad8lr": lambda: mg.quad_mesh(8, tags="lr"),
    "tri7": lambda: mg.tri_mesh(7),
    "tet3": lambda: mg.tet_mesh(3),
}
GEOM = ["vert_volume", "elem_coord", "elem_volume", "face_coord", "face_norm", "face_area",
        "sub_face_norm", "sub_face_area", "sub_elem_volume"]
SCHEMES = ["multi_point", "two_wave", "three_wave", "modified_three_wave", "multi_point_iso",
           "multi_point_pressure", "multi_point_pressure_ph"]


def smooth_state(xc, seed=7):
    rng = np.random.default_rng(seed)
    x, y, z = xc[:, 0], xc[:, 1], xc[:, 2]
    w = np.stack([1.0 + 0.2 * np.sin(3 * x + 1) * np.cos(2 * y) + 0.05 * rng.random(x.size),
                  0.3 * np.sin(2 * y + 0.5), -0.2 * np.cos(3 * x), 0.1 * np.sin(4 * z + x),
                  1.0 + 0.3 * np.cos(2 * x + y) + 0.05 * rng.random(x.size)], 1)
    return mg.primit_to_conserv(w)


@pytest.fixture(scope="module")
def solver():
    s = Solver(0)
    yield s
    s.close()


def _pair(solver, name):
    mesh = MESHES[name]()
    solver.build_mesh(mesh)
    return mesh, O.OracleMesh(mesh)


@pytest.mark.parametrize("name", list(MESHES))
def test_mesh_connectivity_and_geometry_bit_exact(solver, name):
    print(f"[gpu] geometry {name}", flush=True)
    mesh, om = _pair(solver, name)
    for n in ["vert_sub_face", "sub_face_left_elem", "sub_face_right_elem", "vert_sub_elem", "elem_vert",
              "face_vert", "sub_elem_sub_face", "vert_is_bound"]:
        assert np.array_equal(solver.mesh_i(n), om.geti(n)), n
    for g in GEOM:
        a, b = solver.mesh_d(g), om.getd(g)
        assert a.shape == b.shape, g
        assert np.array_equal(a, b), (g, np.abs(a - b).max())


def _bcs_for(name):
    if name == "quad8lr":
        return [("freestream", (1.0, 0.2, 0.0, 0.0, 1.0)), ("outflowsupersonic", (0, 0, 0, 0, 0))]
    return []


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("name", list(MESHES))
def test_euler_flux_first_order_bit_exact(solver, name, scheme):
    print(f"[gpu] flux o1 {name} {scheme}", flush=True)
    mesh, om = _pair(solver, name)
    b2d = mesh.boundary_2d
    bcs = _bcs_for(name)
    solver.set_config(make_config(scheme, boundary_2d=b2d, bcs=bcs))
    cfg = O.make_config(scheme, boundary_2d=b2d, bcs=bcs)
    sol = smooth_state(om.elem_coord)
    rhs, sl, vp = solver.compute_euler_flux(sol, want_vp=True)
    r_rhs, r_sl, r_vp = om.compute_euler_flux(cfg, sol, want_vp=True)
    assert np.array_equal(rhs, r_rhs), np.abs(rhs - r_rhs).max()
    assert np.array_equal(sl, r_sl), np.abs(sl - r_sl).max()
    if scheme in ("multi_point", "multi_point_iso"):
        assert np.array_equal(vp, r_vp), np.abs(vp - r_vp).max()


@pytest.mark.parametrize("bc_style", [0, 1, 2])
@pytest.mark.parametrize("excl", [False, True])
def test_multi_point_bc_styles_bit_exact(solver, bc_style, excl):
    print(f"[gpu] multi_point bc_style={bc_style} excl={excl}", flush=True)
    mesh, om = _pair(solver, "quad8lr")
    bcs = _bcs_for("quad8lr")
    solver.set_config(make_config("multi_point", bc_style=bc_style, exclude_bound_vert=excl, bcs=bcs))
    cfg = O.make_config("multi_point", bc_style=bc_style, exclude_bound_vert=excl, bcs=bcs)
    sol = smooth_state(om.elem_coord, seed=3)
    rhs, sl = solver.compute_euler_flux(sol)
    r_rhs, r_sl = om.compute_euler_flux(cfg, sol)
    assert np.array_equal(rhs, r_rhs) and np.array_equal(sl, r_sl)


@pytest.mark.parametrize("method", [0, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("name", ["quad9", "tri7", "tet3"])
def test_cell_grad_bit_exact(solver, name, method):
    print(f"[gpu] cell grad {name} method={method}", flush=True)
    mesh, om = _pair(solver, name)
    solver.set_config(make_config("multi_point", boundary_2d=mesh.boundary_2d))
    sol = smooth_state(om.elem_coord, seed=11)
    g, ng = solver.compute_cell_grad_from_nodal_grad(sol, method)
    rg, rng_ = om.compute_cell_grad(sol, method)
    assert np.array_equal(ng, rng_), np.abs(ng - rng_).max()
    assert np.array_equal(g, rg), np.abs(g - rg).max()


@pytest.mark.parametrize("scheme", ["multi_point", "two_wave", "multi_point_pressure"])
@pytest.mark.parametrize("name", ["quad8lr", "tri7", "tet3"])
def test_euler_flux_second_order_bit_exact(solver, name, scheme):
    print(f"[gpu] flux o2 {name} {scheme}", flush=True)
    mesh, om = _pair(solver, name)
    bcs = _bcs_for(name)
    solver.set_config(make_config(scheme, boundary_2d=mesh.boundary_2d, bcs=bcs))
    cfg = O.make_config(scheme, boundary_2d=mesh.boundary_2d, bcs=bcs)
    sol = smooth_state(om.elem_coord, seed=5)
    rg, _ = om.compute_cell_grad(sol, 4)
    rhs, sl = solver.compute_euler_flux(sol, grad=rg, second_order=True)
    r_rhs, r_sl = om.compute_euler_flux(cfg, sol, grad=rg, second_order=True)
    assert np.array_equal(rhs, r_rhs), np.abs(rhs - r_rhs).max()
    assert np.array_equal(sl, r_sl)


def test_add_timedisc_and_solve_bit_exact(solver):
    print("[gpu] add_timedisc_and_solve", flush=True)
    mesh, om = _pair(solver, "tri7")
    solver.set_config(make_config("multi_point", cfl=0.4))
    cfg = O.make_config("multi_point", cfl=0.4)
    sol = smooth_state(om.elem_coord, seed=2)
    rhs, sl = solver.compute_euler_flux(sol)
    dt, omega = solver.add_timedisc_and_solve(1.0, 1.0, 0.4, 0.0)
    r_rhs, r_sl = om.compute_euler_flux(cfg, sol)
    ref = om.add_timedisc_and_solve(cfg, sol, r_rhs, r_sl, 1.0, 1.0, 0.4, 0.0)
    assert ref["status"] == 0
    assert dt == ref["dt"] and omega == ref["omega"]
    assert np.array_equal(solver.get_sol(), ref["sol"])


@pytest.mark.parametrize("case", [
    ("quad9", "multi_point", 0, False, 40),
    ("quad8lr", "two_wave", 0, False, 40),
    ("tri7", "multi_point", 0, True, 15),
    ("tet3", "multi_point", 0, True, 10),
    ("tri7", "multi_point", 1, True, 10),
    ("quad9", "three_wave", 1, False, 20),
])
def test_time_loop_bit_exact(solver, case):
    name, scheme, ts, o2, nsteps = case
    print(f"[gpu] time loop {case}", flush=True)
    mesh, om = _pair(solver, name)
    bcs = _bcs_for(name)
    solver.set_config(make_config(scheme, boundary_2d=mesh.boundary_2d, bcs=bcs, second_order=o2, cfl=0.4))
    solver.set_time(0.0)
    cfg = O.make_config(scheme, boundary_2d=mesh.boundary_2d, bcs=bcs, second_order=o2, cfl=0.4)
    xc = om.elem_coord
    sol = mg.init_1drp(xc, 0.5, [1, 0.1, 0, 0, 1], [0.125, 0, 0, 0, 0.1])
    solver.put_sol(sol)
    t, dt = solver.run(nsteps, ts)
    got = solver.get_sol()
    st, ref, info = om.run(cfg, sol, nsteps, ts)
    assert st == 0
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    assert rel <= 1e-12, rel
    assert np.array_equal(got, ref), rel
    assert t == info["t"]


def test_bad_omega_is_reported(solver):
    """A cell with negative pressure cannot be made positive by any omega >
    1e-6: the reference stops with BAD OMEGA (ns_euler_module.F90:595-598)."""
    print("[gpu] bad omega", flush=True)
    mesh, om = _pair(solver, "quad9")
    solver.set_config(make_config("two_wave", cfl=0.4))
    solver.set_time(0.0)
    sol = mg.init_1drp(om.elem_coord, 0.5, [1, 0, 0, 0, 1], [0.125, 0, 0, 0, 0.1])
    sol[3, 4] = -1.0
    solver.put_sol(sol)
    with pytest.raises(SubfvError) as e:
        solver.run(2)
    assert e.value.code == 3


@pytest.mark.parametrize("case", [("quad8lr", 0, False, 25), ("tri7", 0, True, 10), ("tet3", 0, True, 6)])
def test_local_time_step_loop_bit_exact(solver, case):
    """local_time_step = .true. (ns_euler_module.F90:557-560): dt stays the
    global CFL minimum but is not clamped to tmax, the point-implicit
    diagonal keeps a1*sum_lambda, and the positivity back-off does not
    shorten dt; the pseudo-time mode of the potential-flow tests."""
    name, ts, o2, nsteps = case
    print(f"[gpu] local time step {case}", flush=True)
    mesh, om = _pair(solver, name)
    bcs = _bcs_for(name)
    kw = dict(boundary_2d=mesh.boundary_2d, bcs=bcs, second_order=o2, cfl=0.8, local_time_step=True, tmax=0.05)
    solver.set_config(make_config("multi_point", **kw))
    solver.set_time(0.0)
    cfg = O.make_config("multi_point", **kw)
    sol = mg.init_1drp(om.elem_coord, 0.5, [1, 0.1, 0, 0, 1], [0.125, 0, 0, 0, 0.1])
    solver.put_sol(sol)
    t, dt = solver.run(nsteps, ts)
    got = solver.get_sol()
    st, ref, info = om.run(cfg, sol, nsteps, ts)
    assert st == 0
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    assert t == info["t"] and dt == info["dt"]
