"""Solution output from host snapshots of the device state.

`write_sol_dat` writes the table of write_sol_dat (src/ns/ns_io_module.F90:
50-96): a blank first line, then one record per cell strictly inside the
box -- centre, conservative state, primitive state.  `write_sol_vtu` writes
an ASCII VTK unstructured grid with the reference's cell-data names for the
state (write_sol_vtu, :403-884: Energy per unit mass, Temperature from
temp_u; its other fields -- Hessians, nodal gradients, Ducros sensor -- are
not written), and `restart_from_vtu` reads the state back as the
reference's restart does.  `AsyncWriter` runs writes on
host threads: the caller hands over a snapshot (arrays it no longer
changes) and goes on, so a served loop never waits on a file.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET
from concurrent.futures import Future, ThreadPoolExecutor

import numpy as np

GAMMA = 1.4
# gmsh element kind -> VTK cell type (vertex orders agree for these kinds)
VTK_TYPE = {4: 10, 5: 12, 6: 13, 7: 14}
N_VERT = {4: 4, 5: 8, 6: 6, 7: 5}


def primitives(sol: np.ndarray) -> np.ndarray:
    """conserv_to_primit (ns_euler_primitives_module.F90:19-30) row by row."""
    sol = np.asarray(sol, np.float64)
    w = np.empty_like(sol)
    w[:, 0] = sol[:, 0]
    w[:, 1:4] = sol[:, 1:4] / sol[:, 0:1]
    w[:, 4] = (GAMMA - 1) * (sol[:, 4] - 0.5 * w[:, 0] * (w[:, 1] ** 2 + w[:, 2] ** 2 + w[:, 3] ** 2))
    return w


def write_sol_dat(path: str, elem_coord: np.ndarray, sol: np.ndarray, box=None) -> int:
    """Cells whose centre lies strictly inside box = (xmin, xmax, ymin, ymax,
    zmin, zmax) (all cells when None); returns the number of records."""
    xc = np.asarray(elem_coord, np.float64).reshape(-1, 3)
    sol = np.asarray(sol, np.float64).reshape(-1, 5)
    inside = np.ones(xc.shape[0], bool)
    if box is not None:
        for d in range(3):
            inside &= (xc[:, d] < box[2 * d + 1]) & (xc[:, d] > box[2 * d])
    rows = np.concatenate([xc, sol, primitives(sol)], axis=1)[inside]
    with open(path, "w") as fh:
        fh.write("\n")
        np.savetxt(fh, rows, fmt="%.17g")
    return int(inside.sum())


def _array(fh, name, data, ncomp, typ="Float64", fmt="%.17g"):
    fh.write(f'<DataArray type="{typ}" Name="{name}" NumberOfComponents="{ncomp}" format="ascii">\n')
    np.savetxt(fh, np.asarray(data).reshape(-1, ncomp), fmt=fmt)
    fh.write("</DataArray>\n")


def write_sol_vtu(path: str, vert_coord: np.ndarray, elem_kind: np.ndarray, elem_vert: np.ndarray,
                  sol: np.ndarray, elem_coord: np.ndarray | None = None, elem_tag: np.ndarray | None = None,
                  cv: float = 720.19471) -> None:
    """vert_coord (V, 3); elem_kind gmsh kinds; elem_vert the elements'
    1-based vertex ids, concatenated; sol (C, 5) conservative."""
    X = np.asarray(vert_coord, np.float64).reshape(-1, 3)
    kind = np.asarray(elem_kind, np.int64)
    ev = np.asarray(elem_vert, np.int64) - 1
    sol = np.asarray(sol, np.float64).reshape(-1, 5)
    nv = np.array([N_VERT[int(k)] for k in kind], np.int64)
    if nv.sum() != ev.size or sol.shape[0] != kind.size:
        raise ValueError("elements, their vertices and the state do not match")
    w = primitives(sol)
    a = np.sqrt(GAMMA * w[:, 4] / w[:, 0])
    speed = np.sqrt(w[:, 1] ** 2 + w[:, 2] ** 2 + w[:, 3] ** 2)
    tau = 1.0 / sol[:, 0]  # temp_u (ns_euler_primitives_module.F90:52-63)
    temp = (sol[:, 4] * tau - 0.5 * (sol[:, 1] ** 2 + sol[:, 2] ** 2 + sol[:, 3] ** 2) * tau ** 2) / cv
    with open(path, "w") as fh:
        fh.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian">\n')
        fh.write(f'<UnstructuredGrid>\n<Piece NumberOfPoints="{X.shape[0]}" NumberOfCells="{kind.size}">\n')
        fh.write("<Points>\n")
        _array(fh, "Points", X, 3)
        fh.write("</Points>\n<Cells>\n")
        _array(fh, "connectivity", ev, 1, "Int64", "%d")
        _array(fh, "offsets", np.cumsum(nv), 1, "Int64", "%d")
        _array(fh, "types", np.array([VTK_TYPE[int(k)] for k in kind]), 1, "UInt8", "%d")
        fh.write("</Cells>\n<CellData>\n")
        if elem_coord is not None:
            _array(fh, "Centroid", elem_coord, 3)
        if elem_tag is not None:
            _array(fh, "Tag", elem_tag, 1, "Int32", "%d")
        _array(fh, "Density", sol[:, 0], 1)
        _array(fh, "Momentum", sol[:, 1:4], 3)
        _array(fh, "Energy", sol[:, 4] / sol[:, 0], 1)  # per unit mass, as the reference writes it
        _array(fh, "Velocity", w[:, 1:4], 3)
        _array(fh, "Pressure", w[:, 4], 1)
        _array(fh, "Mach", speed / a, 1)
        _array(fh, "Temperature", temp, 1)
        fh.write("</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")


def restart_from_vtu(path: str) -> np.ndarray:
    """The state of restart_from_vtu_file (ns_io_module.F90:1333-1382) from a
    grid this module wrote: Density, Momentum and Energy (per unit mass, so
    multiplied back by the density) of every cell, as sol (C, 5)."""
    cells = {a.get("Name"): a for a in ET.parse(path).getroot().iter("DataArray")}
    rho = np.array(cells["Density"].text.split(), np.float64)
    mom = np.array(cells["Momentum"].text.split(), np.float64).reshape(-1, 3)
    e = np.array(cells["Energy"].text.split(), np.float64)
    sol = np.empty((rho.size, 5))
    sol[:, 0] = rho
    sol[:, 1:4] = mom
    sol[:, 4] = e * rho
    return sol


def solver_snapshot(solver) -> dict:
    """Host copies of what the writers need from a context: the state and
    the mesh arrays in build_mesh numbering."""
    return {"sol": solver.get_sol(), "elem_coord": solver.elem_coord, "vert_coord": solver.mesh_d("vert_coord"),
            "elem_kind": solver.mesh_i("elem_kind"), "elem_vert": solver.mesh_i("elem_vert"),
            "elem_tag": solver.mesh_i("elem_tag")}


class AsyncWriter:
    """Writes on host threads; submit() returns at once with a Future."""

    def __init__(self, workers: int = 2):
        self.ex = ThreadPoolExecutor(max_workers=workers)

    def dat(self, path: str, snap: dict, box=None) -> Future:
        return self.ex.submit(write_sol_dat, path, snap["elem_coord"], snap["sol"], box)

    def vtu(self, path: str, snap: dict) -> Future:
        return self.ex.submit(write_sol_vtu, path, snap["vert_coord"], snap["elem_kind"], snap["elem_vert"],
                              snap["sol"], snap.get("elem_coord"), snap.get("elem_tag"))

    def close(self):
        self.ex.shutdown(wait=True)
