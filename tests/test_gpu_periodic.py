"""Periodic meshes on the device, against the oracle (build_mesh with
periodic_mesh, src/mesh/mesh_connectivity_module.F90:44-54, 976-1328).

The node sweep, the nodal gradients, the wall normal and the vertex volumes
read the extended vertex lists of build_periodic_stencil; results are
compared bit for bit, as for non-periodic meshes.  The second-order case
keeps the reference's reconstruction at the seam (vertex and cell-centre
coordinates are not shifted by the period, ns_euler_recon_module.F90:441-446).
"""
import numpy as np
import pytest

import oracle as O
from subfv_amd import meshgen as mg
from subfv_amd.api import Solver, SubfvError, make_config

pytestmark = pytest.mark.gpu


def _tets():
    m = mg.tet_mesh(3)
    m.coord = m.coord - 0.5
    return m


MESHES = {
    "quad10": lambda: mg.quad_mesh(10, x0=-0.5, y0=-0.5),
    "tri9": lambda: mg.tri_mesh(9, x0=-0.5, y0=-0.5),
    "tet3": _tets,
}


def smooth_state(xc, seed=7):
    rng = np.random.default_rng(seed)
    x, y, z = xc[:, 0], xc[:, 1], xc[:, 2]
    w = np.stack([1.0 + 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.05 * rng.random(x.size),
                  0.3 * np.sin(2 * np.pi * y), -0.2 * np.cos(2 * np.pi * x), 0.1 * np.sin(2 * np.pi * z),
                  1.0 + 0.3 * np.cos(2 * np.pi * (x + y)) + 0.05 * rng.random(x.size)], 1)
    return mg.primit_to_conserv(w)


@pytest.fixture(scope="module")
def solver():
    s = Solver(0)
    yield s
    s.close()


def _pair(solver, name):
    mesh = MESHES[name]()
    solver.build_mesh(mesh, periodic=True)
    return mesh, O.OracleMesh(mesh, periodic=True)


@pytest.mark.parametrize("name", list(MESHES))
def test_periodic_connectivity_and_geometry_bit_exact(solver, name):
    print(f"[gpu] periodic geometry {name}", flush=True)
    mesh, om = _pair(solver, name)
    for n in ["vert_sub_face", "vert_sub_elem", "sub_face_right_elem", "sub_face_right_sub_elem", "vert_is_bound"]:
        assert np.array_equal(solver.mesh_i(n), om.geti(n)), n
    for g in ["vert_volume", "sub_face_norm", "sub_face_area", "face_norm", "elem_volume"]:
        a, b = solver.mesh_d(g), om.getd(g)
        assert a.shape == b.shape and np.array_equal(a, b), g


@pytest.mark.parametrize("scheme", ["multi_point", "multi_point_iso", "two_wave", "multi_point_pressure"])
@pytest.mark.parametrize("name", list(MESHES))
def test_periodic_flux_first_order_bit_exact(solver, name, scheme):
    print(f"[gpu] periodic flux o1 {name} {scheme}", flush=True)
    mesh, om = _pair(solver, name)
    solver.set_config(make_config(scheme, boundary_2d=mesh.boundary_2d))
    cfg = O.make_config(scheme, boundary_2d=mesh.boundary_2d)
    sol = smooth_state(om.elem_coord)
    rhs, sl, vp = solver.compute_euler_flux(sol, want_vp=True)
    r_rhs, r_sl, r_vp = om.compute_euler_flux(cfg, sol, want_vp=True)
    assert np.array_equal(rhs, r_rhs), np.abs(rhs - r_rhs).max()
    assert np.array_equal(sl, r_sl)
    if scheme.startswith("multi_point") and "pressure" not in scheme:
        assert np.array_equal(vp, r_vp)


@pytest.mark.parametrize("name", list(MESHES))
def test_periodic_gradients_and_second_order_bit_exact(solver, name):
    print(f"[gpu] periodic o2 {name}", flush=True)
    mesh, om = _pair(solver, name)
    solver.set_config(make_config("multi_point", boundary_2d=mesh.boundary_2d))
    cfg = O.make_config("multi_point", boundary_2d=mesh.boundary_2d)
    sol = smooth_state(om.elem_coord, seed=5)
    g, ng = solver.compute_cell_grad_from_nodal_grad(sol, 4)
    rg, rng_ = om.compute_cell_grad(sol, 4)
    assert np.array_equal(ng, rng_) and np.array_equal(g, rg)
    rhs, sl = solver.compute_euler_flux(sol, grad=rg, second_order=True)
    r_rhs, r_sl = om.compute_euler_flux(cfg, sol, grad=rg, second_order=True)
    assert np.array_equal(rhs, r_rhs), np.abs(rhs - r_rhs).max()
    assert np.array_equal(sl, r_sl)


@pytest.mark.parametrize("case", [("quad10", 0, False, 30), ("tri9", 1, True, 10), ("tet3", 1, True, 6)])
def test_periodic_time_loop_bit_exact(solver, case):
    name, ts, o2, nsteps = case
    print(f"[gpu] periodic time loop {case}", flush=True)
    mesh, om = _pair(solver, name)
    solver.set_config(make_config("multi_point", boundary_2d=mesh.boundary_2d, second_order=o2, cfl=0.4))
    solver.set_time(0.0)
    cfg = O.make_config("multi_point", boundary_2d=mesh.boundary_2d, second_order=o2, cfl=0.4)
    sol = smooth_state(om.elem_coord, seed=2)
    solver.put_sol(sol)
    t, _ = solver.run(nsteps, ts)
    got = solver.get_sol()
    st, ref, info = om.run(cfg, sol, nsteps, ts)
    assert st == 0
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    assert t == info["t"]


def test_viscous_terms_are_refused_on_periodic_meshes(solver):
    print("[gpu] periodic viscous refused", flush=True)
    _pair(solver, "quad10")
    solver.set_config(make_config("multi_point"))
    solver.set_diffusion(True)
    solver.put_sol(smooth_state(solver.elem_coord))
    with pytest.raises(SubfvError):
        solver.compute_rhs_ns()
    solver.set_diffusion(False)
