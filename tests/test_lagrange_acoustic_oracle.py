"""The oracle's node sweeps of src/lagrange (compute_rhs_lagrange,
lagrange_module.F90:14-147) and src/acoustic (one step of
acoustic_main.F90:171-290) pinned by what the mathematics fixes:

- Lagrange: for a uniform state the interior nodal velocity is the flow
  velocity and the interior cells' residual vanishes (closed cells); the
  wall projection leaves no normal velocity (Bp . vp = 0); the mass-like
  first component is conserved to round-off;
- acoustic: a state at rest under walls and a uniform flow under outflow
  boundaries are steady; dt is the minimum of vol / face area (a = 1);
- inverse_3_by_3 against numpy (the bit-exact pin against the reference's
  own routine is in test_oracle_vs_reference.py)."""
import numpy as np
import pytest

import oracle as O
from subfv_amd import meshgen as mg


@pytest.fixture(scope="module", params=["quad", "tri"])
def mesh_om(request):
    mesh = mg.quad_mesh(10) if request.param == "quad" else mg.tri_mesh(8)
    return mesh, O.OracleMesh(mesh)


def test_inverse_3_by_3_matches_numpy():
    rng = np.random.default_rng(0)
    for _ in range(50):
        A = rng.standard_normal((3, 3)) + 2 * np.eye(3)
        assert np.allclose(O.inverse_3_by_3(A), np.linalg.inv(A), rtol=1e-12, atol=1e-13)


def lagrange_state(ne, u=(0.3, -0.2, 0.0), rho=1.3, E=2.5):
    s = np.zeros((ne, 5))
    s[:, 0] = rho
    s[:, 1:4] = u
    s[:, 4] = E
    return s


def test_lagrange_uniform_state(mesh_om):
    mesh, om = mesh_om
    ne = om.n_elems
    sol = lagrange_state(ne)
    mass = om.elem_volume * sol[:, 0]
    rhs, vp = om.compute_rhs_lagrange(sol, mass, b2d=True)
    inner_v = ~om.geti("vert_is_bound").astype(bool)
    assert np.allclose(vp[inner_v], [0.3, -0.2, 0.0], rtol=0, atol=1e-13)
    # cells whose vertices are all interior see only interior vertices
    evoff = np.concatenate([[0], np.cumsum(om.geti("elem_n_vert"))])
    ev = om.geti("elem_vert") - 1
    cell_inner = np.array([inner_v[ev[evoff[c]:evoff[c + 1]]].all() for c in range(ne)])
    assert cell_inner.sum() > 0
    assert np.abs(rhs[cell_inner]).max() < 1e-12


def test_lagrange_wall_projection_and_conservation(mesh_om):
    mesh, om = mesh_om
    ne = om.n_elems
    rng = np.random.default_rng(1)
    sol = lagrange_state(ne)
    sol[:, 0] *= 1 + 0.1 * rng.random(ne)
    sol[:, 1:3] += 0.05 * rng.standard_normal((ne, 2))
    sol[:, 4] *= 1 + 0.1 * rng.random(ne)
    mass = om.elem_volume * sol[:, 0]
    rhs, vp = om.compute_rhs_lagrange(sol, mass, b2d=True)
    # sum over cells of mass * rhs(1): every interior sub-face term cancels, walls carry none
    total = (mass * rhs[:, 0]).sum()
    scale = (mass * np.abs(rhs[:, 0])).sum()
    assert abs(total) < 1e-12 * scale
    # wall vertices: the projection leaves Bp . vp = 0, Bp the sum of area * n
    # over the vertex's lateral wall sub-faces
    nsf = om.geti("vert_n_sub_faces")
    vsf = om.geti("vert_sub_face") - 1
    off = np.concatenate([[0], np.cumsum(nsf)])
    sff = om.geti("sub_face_face") - 1
    fr = om.geti("face_right")
    fn = om.getd("face_norm").reshape(-1, 3)
    sfa = om.getd("sub_face_area")
    checked = 0
    for v in range(om.n_vert):
        Bp = np.zeros(3)
        for s_ in vsf[off[v]:off[v + 1]]:
            f = sff[s_]
            if fr[f] <= 0 and abs(fn[f, 2]) <= 1e-8:
                Bp += sfa[s_] * fn[f]
        if np.abs(Bp).max() > 1e-8:
            checked += 1
            assert abs(Bp @ vp[v]) < 1e-12 * np.linalg.norm(Bp) * (np.abs(vp[v]).max() + 1.0)
    assert checked > 0


def test_acoustic_rest_and_uniform_flow_are_steady(mesh_om):
    mesh, om = mesh_om
    ne = om.n_elems
    rest = np.tile([1.7, 0.0, 0.0, 0.0], (ne, 1))
    new, dt = om.acoustic_step(rest, 0.5)
    assert np.allclose(new, rest, rtol=0, atol=1e-13)
    # uniform flow: the cells away from the walls keep it
    flow = np.tile([1.0, 0.2, -0.1, 0.0], (ne, 1))
    new, _ = om.acoustic_step(flow, 0.5)
    # interior cells (no wall face) keep the state
    evoff = np.concatenate([[0], np.cumsum(om.geti("elem_n_vert"))])
    ev = om.geti("elem_vert") - 1
    inner_v = ~om.geti("vert_is_bound").astype(bool)
    cell_inner = np.array([inner_v[ev[evoff[c]:evoff[c + 1]]].all() for c in range(ne)])
    assert np.allclose(new[cell_inner], flow[cell_inner], rtol=0, atol=1e-13)


def test_acoustic_dt_is_the_volume_over_face_area_minimum(mesh_om):
    mesh, om = mesh_om
    ne = om.n_elems
    rest = np.tile([1.0, 0.0, 0.0, 0.0], (ne, 1))
    _, dt = om.acoustic_step(rest, 0.5)
    fl, fr = om.geti("face_left"), om.geti("face_right")
    fa = om.getd("face_area")
    vol = om.elem_volume
    # faces touching a vertex: every face; both sides where interior
    cand = vol[fl - 1] / fa
    cand = np.concatenate([cand, vol[fr[fr > 0] - 1] / fa[fr > 0]])
    assert dt == cand.min()


def test_acoustic_freestream_and_outflow_tags():
    mesh = mg.quad_mesh(8, tags="lr")
    om = O.OracleMesh(mesh)
    ne = om.n_elems
    flow = np.tile([1.0, 0.25, 0.0, 0.0], (ne, 1))
    # inflow from the left at the flow state, outflow to the right: steady
    new, _ = om.acoustic_step(flow, 0.5, bc_kind=[1, 2], bc_val=[[1.0, 0.25, 0.0, 0.0]])
    evoff = np.concatenate([[0], np.cumsum(om.geti("elem_n_vert"))])
    ev = om.geti("elem_vert") - 1
    y = om.getd("vert_coord").reshape(-1, 3)[:, 1]
    # cells away from the (wall) top and bottom sides
    far = np.array([(y[ev[evoff[c]:evoff[c + 1]]] > 1e-9).all() and (y[ev[evoff[c]:evoff[c + 1]]] < 1 - 1e-9).all()
                    for c in range(ne)])
    assert np.allclose(new[far], flow[far], rtol=0, atol=1e-13)
    with pytest.raises(RuntimeError):
        om.acoustic_step(flow, 0.5, bc_kind=[1])  # tag 2 has no kind
