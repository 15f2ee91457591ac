"""The VTU output of a device solve (subfv_amd/output.py): every point and
cell field of the reference's write_sol_vtu (ns_io_module.F90:582-870), from
a snapshot of the device state and the nodal and cell gradients computed on
the device, equals the same fields formed from the oracle's state and
gradients after the same steps."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle as O
from subfv_amd import meshgen as mg
from subfv_amd import output as OUT
from subfv_amd.api import Solver, make_config
from test_output import _oracle_snapshot

pytestmark = pytest.mark.gpu


def test_vtu_fields_of_a_device_solve_equal_the_oracle_fields(tmp_path):
    print("[gpu] vtu fields", flush=True)
    mesh = mg.quad_mesh(10, tags="lr")
    bcs = [("freestream", (1.0, 0.3, 0.0, 0.0, 1.0)), ("wall", (0, 0, 0, 0, 0))]
    s = Solver(0)
    try:
        s.build_mesh(mesh)
        kw = dict(bcs=bcs, second_order=True, method=4, cfl=0.4)
        s.set_config(make_config("multi_point", **kw))
        s.set_time(0.0)
        om = O.OracleMesh(mesh)
        sol0 = mg.init_1drp(om.elem_coord, 0.5, [1, 0.2, 0, 0, 1], [0.3, 0, 0, 0, 0.4])
        s.put_sol(sol0)
        s.run(8, 0)
        snap = OUT.solver_snapshot(s, gradients=True)
    finally:
        s.close()
    st, ref, _ = om.run(O.make_config("multi_point", **kw), sol0, 8, 0)
    assert st == 0 and np.array_equal(snap["sol"], ref)
    fkw = dict(second_order=True, boundary_2d=True, bc_kinds=(3, 1))
    path = tmp_path / "out.vtu"
    wr = OUT.AsyncWriter()
    fut = wr.vtu(str(path), snap, **fkw)
    wr.close()
    assert fut.exception() is None
    got = {a.get("Name"): np.array(a.text.split(), float) for a in ET.parse(path).getroot().iter("DataArray")}
    want = OUT.reference_fields(_oracle_snapshot(om, ref), **fkw)
    names = [nm for nm, _, _ in want["point"]] + [nm for nm, _, _ in want["cell"]]
    for required in ("Nodal_Velocity", "Nodal_Pressure", "Nodal_Grad_Density", "Nodal_Grad_Pressure", "Hessian Mach",
                     "Ducros", "Internal_energy", "H", "Mach", "Temperature", "mu", "r", "rxy", "Theta3",
                     "Density_Grad", "Pressure_Grad", "Residual"):
        assert required in names and required in got, required
    for nm, data, nc in want["point"] + want["cell"]:
        assert np.array_equal(got[nm], np.asarray(data, float).reshape(-1)), nm
