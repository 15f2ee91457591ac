"""The potential-flow boundary states (compute_right_state with
BC_EULER_POTENTIAL_FLOW_2D/3D, ns_euler_recon_module.F90:398-407, states of
ns_euler_exact_sol_module.F90:105-170) on the device against the oracle.

The 3-D state is plain arithmetic: bit for bit.  The 2-D state goes through
atan2, cos and sin, which the device library and the host's libm round
differently in the last place, so runs through it are held to 1e-12
relative."""
import numpy as np
import pytest

import oracle as O
from subfv_amd import meshgen as mg
from subfv_amd.api import Solver, make_config

pytestmark = pytest.mark.gpu


def run_both(kinds, steps=6):
    mesh = mg.quad_mesh(10, tags="lr", x0=0.6, y0=-0.5)
    s = Solver(0)
    s.build_mesh(mesh)
    w0 = np.array([1.4, 1e-4, 0.0, 0.0, 1.0])
    bcs = [(k, w0) for k in kinds]
    s.set_config(make_config("multi_point", bcs=bcs, cfl=0.4))
    om = O.OracleMesh(mesh)
    u0 = mg.primit_to_conserv(np.tile(w0, (om.n_elems, 1)) * (1 + 0.01 * np.random.default_rng(2).random((om.n_elems, 1))))
    s.put_sol(u0)
    s.run(steps)
    got = s.get_sol()
    s.close()
    st, ref, _ = om.run(O.make_config("multi_point", bcs=bcs, cfl=0.4), u0, steps)
    assert st == 0
    return got, ref


def test_potential_flow_3d_bit_exact():
    print("[gpu] potential flow 3d", flush=True)
    got, ref = run_both(["potential_flow_3d", "potential_flow_3d"])
    assert np.array_equal(got, ref)


def test_potential_flow_2d_within_rounding_of_the_trig_functions():
    print("[gpu] potential flow 2d", flush=True)
    got, ref = run_both(["potential_flow_2d", "outflowsupersonic"])
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
