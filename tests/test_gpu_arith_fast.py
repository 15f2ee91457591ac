"""The node sweep's fast arithmetic (subfv_set_arith: fused multiply-adds and
reciprocal products, flux_fast.hip) against the oracle: not bit for bit, but
within 1e-12 relative after 200 steps on the first-order Sod workload (quad
128 x 128) and on the tet blast with first-order and limited space, after
100 second-order viscous SSP-RK3 steps, and mass still conserved to
round-off.  Switching back gives the exact kernel's bits again."""
import numpy as np
import pytest

import oracle as O
from subfv_amd import meshgen as mg
from subfv_amd.api import Solver, make_config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    s = Solver(0)
    yield s
    s.set_arith(False)
    s.close()


def _rel(got, ref):
    return float(np.abs(got - ref).max() / np.abs(ref).max())


def _mass(om, sol):
    return float((om.getd("elem_volume") * sol[:, 0]).sum())


def _run(solver, mesh, sol0, nsteps, ts, fast, diffusion=None, **kw):
    solver.build_mesh(mesh)
    solver.set_arith(fast)
    solver.set_config(make_config("multi_point", boundary_2d=mesh.boundary_2d, **kw))
    if diffusion:
        solver.set_diffusion(True, **diffusion)
    solver.set_time(0.0)
    solver.put_sol(sol0)
    t, _ = solver.run(nsteps, ts)
    got = solver.get_sol()
    if diffusion:
        solver.set_diffusion(False)
    return got, t


def test_sod_quad128_200_steps_within_1e12(solver):
    print("[gpu] fast arith sod quad128 x 200", flush=True)
    mesh = mg.quad_mesh(128)
    om = O.OracleMesh(mesh)
    sol0 = mg.init_1drp(om.elem_coord, 0.5, [1, 0, 0, 0, 1], [0.125, 0, 0, 0, 0.1])
    got, t = _run(solver, mesh, sol0, 200, 0, True, cfl=0.4)
    st, ref, info = om.run(O.make_config("multi_point", cfl=0.4), sol0, 200, 0)
    assert st == 0
    rel = _rel(got, ref)
    print(f"  max rel diff {rel:.3e}", flush=True)
    assert rel <= 1e-12, rel
    assert abs(t - info["t"]) <= 1e-12 * info["t"]
    m0, m1 = _mass(om, sol0), _mass(om, got)
    assert abs(m1 - m0) <= 1e-13 * m0


@pytest.mark.parametrize("method", [0, 4])
def test_blast_tet_200_steps_within_1e12(solver, method):
    print(f"[gpu] fast arith blast tet6 method={method} x 200", flush=True)
    mesh = mg.tet_mesh(6)
    om = O.OracleMesh(mesh)
    sol0 = mg.blast_state(om.elem_coord, om.getd("elem_volume"), 1.0 / 6)
    kw = dict(second_order=True, method=method, cfl=0.4)
    got, t = _run(solver, mesh, sol0, 200, 0, True, **kw)
    st, ref, info = om.run(O.make_config("multi_point", boundary_2d=False, **kw), sol0, 200, 0)
    assert st == 0
    rel = _rel(got, ref)
    print(f"  max rel diff {rel:.3e}", flush=True)
    assert rel <= 1e-12, rel
    m0, m1 = _mass(om, sol0), _mass(om, got)
    assert abs(m1 - m0) <= 1e-13 * m0


def test_viscous_rk3_100_steps_within_1e12(solver):
    print("[gpu] fast arith viscous rk3 tri32 x 100", flush=True)
    mesh = mg.tri_mesh(32)
    om = O.OracleMesh(mesh)
    sol0 = mg.vortex_state(om.elem_coord)
    kw = dict(second_order=True, method=4, cfl=0.4)
    got, t = _run(solver, mesh, sol0, 100, 1, True, diffusion=dict(mu=1e-3, prandtl=0.72), **kw)
    st, ref, info = om.run(O.make_config("multi_point", diffusion=True, mu=1e-3, prandtl=0.72, **kw), sol0, 100, 1)
    assert st == 0
    rel = _rel(got, ref)
    print(f"  max rel diff {rel:.3e}", flush=True)
    assert rel <= 1e-12, rel


def test_exact_again_after_fast(solver):
    print("[gpu] fast arith off again", flush=True)
    mesh = mg.tri_mesh(12)
    om = O.OracleMesh(mesh)
    sol0 = mg.init_1drp(om.elem_coord, 0.5, [1, 0, 0, 0, 1], [0.125, 0, 0, 0, 0.1])
    _run(solver, mesh, sol0, 5, 0, True, cfl=0.4)
    got, _ = _run(solver, mesh, sol0, 20, 0, False, cfl=0.4)
    st, ref, _ = om.run(O.make_config("multi_point", cfl=0.4), sol0, 20, 0)
    assert st == 0
    assert np.array_equal(got, ref)
