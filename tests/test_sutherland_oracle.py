"""Pins of the oracle's Sutherland viscosity (compute_mu with
use_sutherland, ns_vectorial_diffusion_module.F90:322-359): mu of every
cell from its temperature, mu0corr their minimum, kappa of each cell's mu in
the heat system (ns_diffusion_module.F90:50,78), each side's cell mu in the
momentum blocks (:79,108,137,170) and the sub-element's cell mu in the
viscous work (:301).

Pinned here:
- the law itself against the textbook value for air at 300 K
  (1.846e-5 Pa s with mu0 = 1.716e-5 at 273.15 K, C = 110.4 K) and at T0;
- at uniform temperature the viscous operator equals the constant-viscosity
  one with mu_p = mu(T) (every place mu enters, and mu0corr);
- locality: a hot patch changes the operator only in cells that share a
  vertex with a cell that shares a vertex with the patch; every other cell
  matches the constant-viscosity operator of the cold temperature;
- the momentum rows and the heat part of the energy row still sum to zero
  over the domain when mu varies from cell to cell (with the wall cells at
  the minimum, see the test).
"""
import numpy as np
import pytest

import oracle as O
from subfv_amd import meshgen as mg
from test_diffusion_oracle import interior

CV = 720.19471


def state(om, T, rng, u_amp=0.2):
    """primitive rho, u, v, w, p with temperature T (scalar or per cell)"""
    n = om.n_elems
    rho = 1.0 + 0.1 * rng.random(n)
    T = np.broadcast_to(np.asarray(T, np.float64), (n,))
    w = np.stack([rho, u_amp * rng.standard_normal(n), u_amp * rng.standard_normal(n),
                  0.5 * u_amp * rng.standard_normal(n), rho * 0.4 * CV * T], 1)
    return mg.primit_to_conserv(w)


def test_law_against_textbook_values():
    om = O.OracleMesh(mg.quad_mesh(3))
    cfg = O.make_config("multi_point", diffusion=True, sutherland=True, cv=CV)
    rng = np.random.default_rng(0)
    mu, mn = O.compute_mu(cfg, state(om, 300.0, rng, u_amp=0.0))
    assert np.allclose(mu, 1.846e-5, rtol=5e-4) and mn == mu.min()
    mu, _ = O.compute_mu(cfg, state(om, 273.15, rng, u_amp=0.0))
    assert np.allclose(mu, 1.716e-5, rtol=1e-14)
    mu, mn = O.compute_mu(O.make_config("multi_point", diffusion=True, mu=3e-3), state(om, 300.0, rng))
    assert np.all(mu == 3e-3) and mn == 3e-3


@pytest.mark.parametrize("mesh", [mg.quad_mesh(8), mg.tri_mesh(7), mg.tet_mesh(3)])
def test_uniform_temperature_matches_constant_viscosity(mesh):
    om = O.OracleMesh(mesh)
    rng = np.random.default_rng(4)
    sol = state(om, 310.0, rng)
    cs = O.make_config("multi_point", boundary_2d=mesh.boundary_2d, diffusion=True, sutherland=True, cv=CV)
    mu, _ = O.compute_mu(cs, sol)
    assert mu.max() - mu.min() <= 1e-14 * mu.max()
    cc = O.make_config("multi_point", boundary_2d=mesh.boundary_2d, diffusion=True, mu=float(mu.mean()), cv=CV)
    r1, d1 = om.add_diffusion(cs, sol)
    r0, d0 = om.add_diffusion(cc, sol)
    sr = np.abs(d0).max() * np.abs(sol).max() + np.abs(r0).max()  # the operator's scale
    assert np.abs(r1 - r0).max() <= 1e-12 * sr
    assert np.abs(d1 - d0).max() <= 1e-12 * np.abs(d0).max()


def test_hot_patch_changes_only_its_two_vertex_rings():
    mesh = mg.tri_mesh(12)
    om = O.OracleMesh(mesh)
    n = om.n_elems
    xc = om.elem_coord
    hot = (np.abs(xc[:, 0] - 0.5) < 0.12) & (np.abs(xc[:, 1] - 0.5) < 0.12)
    assert 4 <= hot.sum() < n // 10
    ev = om.geti("elem_vert")
    off = np.concatenate([[0], np.cumsum(om.geti("elem_n_vert"))])
    verts = [ev[off[c]:off[c + 1]] - 1 for c in range(n)]
    vh = np.zeros(om.geti("vert_is_bound").size, bool)
    for c in np.nonzero(hot)[0]:
        vh[verts[c]] = True
    ring1 = np.array([vh[verts[c]].any() for c in range(n)])
    v2 = vh.copy()
    for c in np.nonzero(ring1)[0]:
        v2[verts[c]] = True
    near = np.array([v2[verts[c]].any() for c in range(n)])
    rng = np.random.default_rng(9)
    T = np.where(hot, 360.0, 280.0)
    sol = state(om, T, rng)
    cs = O.make_config("multi_point", diffusion=True, sutherland=True, cv=CV)
    mu, mn = O.compute_mu(cs, sol)
    mu_cold = float(mu[~hot].mean())
    assert mn == mu.min() and abs(mn - mu_cold) <= 1e-14 * mu_cold and mu[hot].min() > 1.1 * mu_cold
    r1, d1 = om.add_diffusion(cs, sol)
    r0, d0 = om.add_diffusion(O.make_config("multi_point", diffusion=True, mu=mu_cold, cv=CV), sol)
    sd = np.abs(d0).max()
    sr = sd * np.abs(sol).max() + np.abs(r0).max()
    assert (~near).sum() > n // 2
    assert np.abs(r1[~near] - r0[~near]).max() <= 1e-12 * sr
    assert np.abs(d1[~near] - d0[~near]).max() <= 1e-12 * sd
    # and inside the patch mu is ~25 % larger: the blocks and the momentum rows move by that order
    assert np.abs(d1[hot] - d0[hot]).max() > 0.05 * np.abs(d0[hot]).max()
    assert np.abs(r1[hot, 1:4] - r0[hot, 1:4]).max() > 0.05 * np.abs(r0[hot, 1:4]).max()


@pytest.mark.parametrize("mesh", [mg.quad_mesh(10), mg.tri_mesh(8), mg.tet_mesh(6)])
def test_variable_viscosity_operators_are_conservative(mesh):
    """mu varies freely in the cells away from the walls; the cells touching
    a wall vertex hold the minimum, so mu0corr is their viscosity.  (Where a
    wall cell's mu exceeds mu0corr, the reference's wall rows -- face velocity
    given, its (mu - mu0corr) terms kept in S~^T only -- let momentum through
    the wall: one interior cell colder than the walls moves the domain sum by
    a few per cent.  Interior vertices conserve for any mu.)"""
    om = O.OracleMesh(mesh)
    n = om.n_elems
    inner = interior(om, layers=1)
    assert inner.sum() >= 30
    rng = np.random.default_rng(5)
    T = np.where(inner, 300.0 + 100.0 * rng.random(n), 300.0)
    cfg = O.make_config("two_wave", boundary_2d=mesh.boundary_2d, diffusion=True, sutherland=True, cv=CV)
    mu, mn = O.compute_mu(cfg, state(om, T, rng))
    assert mu.max() > 1.1 * mn
    sol = state(om, T, rng)
    r, d = om.add_diffusion(cfg, sol)
    R = r - np.einsum("cji,cj->ci", d, sol)
    assert np.abs(R[:, 1:4].sum(axis=0)).max() <= 1e-11 * np.abs(R[:, 1:4]).max()
    # heat only: fluid at rest, temperatures varying
    sol = state(om, T, rng, u_amp=0.0)
    r, d = om.add_diffusion(cfg, sol)
    R = r - np.einsum("cji,cj->ci", d, sol)
    assert abs(R[:, 4].sum()) <= 1e-12 * np.abs(R[:, 4]).max()
