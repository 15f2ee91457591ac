"""The node sweeps of src/lagrange (compute_rhs_lagrange) and src/acoustic
(the acoustic time loop) on the device, bit for bit against the oracle: both
keep the reference's expression order and the order in which every cell
receives its terms (vertex, then sub-face)."""
import numpy as np
import pytest

import oracle as O
from subfv_amd import meshgen as mg
from subfv_amd.api import Solver, SubfvError

pytestmark = pytest.mark.gpu

MESHES = {
    "quad12": lambda: mg.quad_mesh(12),
    "tri10": lambda: mg.tri_mesh(10),
    "tet3": lambda: mg.tet_mesh(3),
}


@pytest.fixture(scope="module")
def solver():
    s = Solver(0)
    yield s
    s.close()


def lagrange_state(ne, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros((ne, 5))
    s[:, 0] = 1.0 + 0.3 * rng.random(ne)
    s[:, 1:4] = 0.2 * rng.standard_normal((ne, 3))
    s[:, 4] = 2.0 + rng.random(ne)
    return s


@pytest.mark.parametrize("name", list(MESHES))
def test_lagrange_rhs_bit_exact(solver, name):
    print(f"[gpu] lagrange rhs {name}", flush=True)
    mesh = MESHES[name]()
    b2d = name != "tet3"
    solver.build_mesh(mesh)
    om = O.OracleMesh(mesh)
    sol = lagrange_state(om.n_elems, 3)
    if not b2d:
        sol[:, 3] = 0.2 * np.random.default_rng(4).standard_normal(om.n_elems)
    mass = om.elem_volume * sol[:, 0]
    rhs, vp = solver.compute_rhs_lagrange(sol, mass, b2d)
    rhs_o, vp_o = om.compute_rhs_lagrange(sol, mass, b2d)
    assert np.array_equal(vp, vp_o)
    assert np.array_equal(rhs, rhs_o)


@pytest.mark.parametrize("name", ["quad12", "tri10"])
def test_acoustic_steps_bit_exact_with_walls(solver, name):
    print(f"[gpu] acoustic walls {name}", flush=True)
    mesh = MESHES[name]()
    solver.build_mesh(mesh)
    om = O.OracleMesh(mesh)
    rng = np.random.default_rng(5)
    ne = om.n_elems
    sol = np.column_stack([1.0 + 0.2 * rng.random(ne), 0.1 * rng.standard_normal((ne, 2)), np.zeros(ne)])
    got, t, dt = solver.acoustic_run(sol, 8, 0.4)
    ref, tr = sol.copy(), 0.0
    for _ in range(8):
        ref, dtr = om.acoustic_step(ref, 0.4)
        tr = tr + dtr
    assert np.array_equal(got, ref)
    assert t == tr and dt == dtr


def test_acoustic_freestream_and_outflow_bit_exact(solver):
    print("[gpu] acoustic freestream/outflow", flush=True)
    mesh = mg.quad_mesh(12, tags="lr")
    solver.build_mesh(mesh)
    om = O.OracleMesh(mesh)
    ne = om.n_elems
    rng = np.random.default_rng(6)
    sol = np.column_stack([1.0 + 0.1 * rng.random(ne), 0.2 + 0.05 * rng.standard_normal(ne),
                           0.05 * rng.standard_normal(ne), np.zeros(ne)])
    kinds, vals = [1, 2], [[1.2, 0.3, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0]]
    got, t, _ = solver.acoustic_run(sol, 6, 0.5, kinds, vals)
    ref, tr = sol.copy(), 0.0
    for _ in range(6):
        ref, dtr = om.acoustic_step(ref, 0.5, kinds, vals)
        tr = tr + dtr
    assert np.array_equal(got, ref) and t == tr
    with pytest.raises(SubfvError):
        solver.acoustic_run(sol, 1, 0.5, [1], [[1.0, 0.0, 0.0, 0.0]])  # tag 2 without a kind
