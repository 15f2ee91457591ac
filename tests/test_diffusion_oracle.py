"""Pins of the oracle's viscous terms (om_diffusion.c: the nodal momentum
and heat systems of ns_vectorial_diffusion_module.F90 and
ns_diffusion_module.F90 with constant viscosity).

The point-implicit form splits the viscous operator into a right-hand side
and a diagonal block D, so the operator applied to the state is
R = rhs - D U.  Pinned here:
- inv_lapack's restatement against numpy;
- R = 0 for uniform velocity and temperature (consistency), for a linear
  shear flow and for a linear temperature profile away from the walls
  (the nodal scheme is exact for linear fields);
- the momentum rows and the heat part of the energy row sum to zero over
  the domain between adiabatic, zero-traction walls; mass is conserved by
  viscous steps;
- a shear layer decays: viscosity removes kinetic energy.
"""
import numpy as np
import pytest

import oracle as O
from subfv_amd import meshgen as mg


def test_inverse_matches_numpy():
    for n in (1, 3, 12, 36, 50):
        A = np.random.default_rng(n).standard_normal((n, n)) + 4 * np.eye(n)
        X = O.inv(A)
        assert np.allclose(A @ X, np.eye(n), atol=1e-13 * np.linalg.cond(A))
        assert np.allclose(X, np.linalg.inv(A), rtol=1e-12, atol=1e-13 * np.linalg.cond(A))


def effective(om, cfg, sol):
    """R = rhs - D U per cell and the operator's scale max|D| max|U| + max|rhs|"""
    r, d = om.add_diffusion(cfg, sol)
    DU = np.einsum("cji,cj->ci", d, sol)  # diag[c].T is the 5x5 block
    scale = np.abs(d).max() * np.abs(sol).max() + np.abs(r).max()
    return r - DU, np.full(5, scale)


def interior(om, layers=2):
    """cells not within `layers` vertex-rings of the boundary"""
    bound = om.geti("vert_is_bound").astype(bool)
    ev = om.geti("elem_vert")
    nv = om.geti("elem_n_vert")
    off = np.concatenate([[0], np.cumsum(nv)])
    bad_v = bound.copy()
    for _ in range(layers):
        cell_bad = np.array([bad_v[ev[off[c]:off[c + 1]] - 1].any() for c in range(om.n_elems)])
        for c in np.nonzero(cell_bad)[0]:
            bad_v[ev[off[c]:off[c + 1]] - 1] = True
    return ~np.array([bad_v[ev[off[c]:off[c + 1]] - 1].any() for c in range(om.n_elems)])


@pytest.mark.parametrize("mesh", [mg.quad_mesh(8), mg.tri_mesh(7), mg.tet_mesh(3)])
def test_uniform_state_has_zero_viscous_operator(mesh):
    om = O.OracleMesh(mesh)
    sol = mg.primit_to_conserv(np.tile([1.2, 0.3, -0.2, 0.1, 0.9], (om.n_elems, 1)))
    cfg = O.make_config("multi_point", boundary_2d=mesh.boundary_2d, diffusion=True, mu=1e-2)
    R, scale = effective(om, cfg, sol)
    for k in range(1, 5):
        assert np.abs(R[:, k]).max() <= 1e-10 * scale[k], k


def test_linear_fields_are_exact_in_the_interior():
    """On a regular grid the nodal systems reproduce linear fields: zero
    momentum residual for shear and dilatation, zero heat-flux divergence
    for a linear temperature, and the viscous work of the shear u = y is
    tau : grad u = mu per unit volume.  (On jittered cells the scheme is
    first-order consistent only, as its sub-element gradients are.)"""
    mesh = mg.quad_mesh(10, jitter=0.0)
    om = O.OracleMesh(mesh)
    xc = om.elem_coord
    vol = om.elem_volume
    inner = interior(om)
    assert inner.sum() > 10
    mu = 1e-2
    cfg = O.make_config("multi_point", diffusion=True, mu=mu)
    one, zero = np.ones(om.n_elems), np.zeros(om.n_elems)
    for u, v in ((xc[:, 1], zero), (zero, xc[:, 0]), (xc[:, 0], zero), (0.3 * xc[:, 0], -0.2 * xc[:, 1])):
        sol = mg.primit_to_conserv(np.stack([one, u, v, zero, one], 1))
        R, scale = effective(om, cfg, sol)
        assert np.abs(R[inner, 1:4]).max() <= 1e-12 * scale[1]
    sol = mg.primit_to_conserv(np.stack([one, xc[:, 1], zero, zero, one], 1))
    R, _ = effective(om, cfg, sol)
    assert np.allclose(R[inner, 4] / vol[inner], mu, rtol=1e-9)
    cv = 720.19471
    theta = 1.0 + 0.5 * xc[:, 0]
    sol = np.stack([one, zero, zero, zero, cv * theta], 1)
    R, scale = effective(om, cfg, sol)
    assert np.abs(R[inner, 4]).max() <= 1e-12 * scale[4]


@pytest.mark.parametrize("mesh", [mg.quad_mesh(10), mg.tri_mesh(8), mg.tet_mesh(3)])
def test_momentum_and_heat_operators_are_conservative(mesh):
    """Between zero-traction, adiabatic walls the momentum rows and the heat
    part of the energy row sum to zero over the domain for any state (the
    nodal systems make the sub-face fluxes single-valued).  The viscous-work
    part is not summed: the reference builds it from the sub-element stress,
    which is continuous across sub-faces only to truncation error."""
    om = O.OracleMesh(mesh)
    n = om.n_elems
    rng = np.random.default_rng(3)
    z = np.zeros(n)
    cfg = O.make_config("two_wave", boundary_2d=mesh.boundary_2d, diffusion=True, mu=1e-2)
    w = np.stack([1 + 0.1 * rng.random(n), 0.2 * rng.standard_normal(n), 0.2 * rng.standard_normal(n),
                  0.1 * rng.standard_normal(n), 1 + 0.1 * rng.random(n)], 1)
    R, scale = effective(om, cfg, mg.primit_to_conserv(w))
    assert np.abs(R[:, 1:4].sum(axis=0)).max() <= 1e-13 * scale[0]
    w = np.stack([1 + 0.1 * rng.random(n), z, z, z, 1 + 0.1 * rng.random(n)], 1)
    R, scale = effective(om, cfg, mg.primit_to_conserv(w))
    assert abs(R[:, 4].sum()) <= 1e-13 * scale[0]


@pytest.mark.parametrize("second", [False, True])
def test_viscous_runs_conserve_mass(second):
    mesh = mg.quad_mesh(10)
    om = O.OracleMesh(mesh)
    xc = om.elem_coord
    w = np.stack([1 + 0.1 * np.sin(6 * xc[:, 0]), 0.2 * np.sin(2 * np.pi * xc[:, 1]), 0 * xc[:, 0],
                  0 * xc[:, 0], 1 + 0.1 * np.cos(5 * xc[:, 1])], 1)
    sol = mg.primit_to_conserv(w)
    vol = om.elem_volume
    cfg = O.make_config("two_wave", diffusion=True, mu=5e-3, second_order=second, cfl=0.3)
    st, out, info = om.run(cfg, sol, 10, 0)
    assert st == 0
    a, b = (vol * sol[:, 0]).sum(), (vol * out[:, 0]).sum()
    assert abs(a - b) <= 1e-13 * abs(a)


def test_shear_layer_loses_kinetic_energy_to_viscosity():
    mesh = mg.quad_mesh(12)
    om = O.OracleMesh(mesh)
    xc = om.elem_coord
    w = np.stack([np.ones(om.n_elems), 0.1 * np.sin(2 * np.pi * xc[:, 1]), 0 * xc[:, 0], 0 * xc[:, 0],
                  np.ones(om.n_elems)], 1)
    sol = mg.primit_to_conserv(w)
    vol = om.elem_volume

    def ke(s):
        return (vol * 0.5 * (s[:, 1] ** 2 + s[:, 2] ** 2) / s[:, 0]).sum()

    inv_cfg = O.make_config("two_wave", cfl=0.3)
    vis_cfg = O.make_config("two_wave", diffusion=True, mu=2e-2, cfl=0.3)
    _, a, _ = om.run(inv_cfg, sol, 10, 0)
    _, b, _ = om.run(vis_cfg, sol, 10, 0)
    assert ke(b) < ke(a)


def test_builders_split_compose_to_the_viscous_block():
    """The two builders for compute_mu's viscosities, momentum then heat on
    the same arrays, give the viscous block of compute_rhs_ns bit for bit,
    with and without Sutherland's law; with a zero viscosity of one builder's
    system only that builder's rows move."""
    mesh = mg.quad_mesh(7, tags="lr")
    om = O.OracleMesh(mesh)
    xc = om.elem_coord
    rng = np.random.default_rng(2)
    T = 300.0 + 40.0 * np.sin(2 * np.pi * xc[:, 0]) + rng.random(om.n_elems)
    w = np.stack([1.0 + 0.1 * np.cos(3 * xc[:, 1]), 20 * np.sin(2 * xc[:, 1]), -10 * np.cos(3 * xc[:, 0]),
                  0.0 * T, (1.0 + 0.1 * np.cos(3 * xc[:, 1])) * 0.4 * 720.19471 * T], 1)
    sol = mg.primit_to_conserv(w)
    for suth in (False, True):
        cfg = O.make_config("multi_point", diffusion=True, mu=2e-3, sutherland=suth,
                            bcs_v=[("dirichlet", (0.1, 0.0, 0.0)), ("neumann", (0, 0, 0))],
                            bcs_t=[("dirichlet", 300.0), ("neumann", 0.0)])
        mu, m0 = O.compute_mu(cfg, sol)
        r1, d1 = om.add_diffusion_mu(cfg, sol, mu, m0, 1)
        r2, d2 = om.add_diffusion_mu(cfg, sol, mu, m0, 2, r1, d1)
        rf, df = om.add_diffusion(cfg, sol)
        assert np.array_equal(r2, rf) and np.array_equal(d2, df)
        r3, d3 = om.add_diffusion_mu(cfg, sol, mu, m0, 3)
        assert np.array_equal(r3, rf) and np.array_equal(d3, df)
        # momentum only: the heat row of diag untouched, the energy rhs is the work term
        assert np.all(d1[:, :, 4] == 0.0) and np.any(d1[:, 1:4, 1:4] != 0.0)
        rh, dh = om.add_diffusion_mu(cfg, sol, mu, m0, 2)
        assert np.all(rh[:, :4] == 0.0) and np.all(dh[:, 1:4, 1:4] == 0.0) and np.any(dh[:, :, 4] != 0.0)
        assert np.allclose(r1[:, 4] + rh[:, 4], rf[:, 4], rtol=1e-13, atol=1e-13 * np.abs(rf[:, 4]).max())
