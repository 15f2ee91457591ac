"""Solution output from host snapshots of the device state.

`write_sol_dat` writes the table of write_sol_dat (src/ns/ns_io_module.F90:
50-96): a blank first line, then one record per cell strictly inside the
box -- centre, conservative state, primitive state.  `write_sol_vtu` writes
an ASCII VTK unstructured grid with the point and cell data of the
reference's write_sol_vtu (:403-884), computed by `reference_fields` from a
snapshot of the device state, its nodal and cell gradients and the mesh
(the reference's Hessians are zero -- their least-squares fit is commented
out there -- and its nodal length h_p is zero on the explicit path), and
`restart_from_vtu` reads the state back as the reference's restart does.  `AsyncWriter` runs writes on
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


def _lists(counts, flat):
    """(owner, item) pairs of a flat list with per-owner counts, 0-based."""
    counts = np.asarray(counts, np.int64)
    owner = np.repeat(np.arange(counts.size), counts)
    return owner, np.asarray(flat, np.int64) - 1


def reference_fields(snap: dict, second_order: bool = False, boundary_2d: bool = True, bc_kinds=(),
                     mu_p: float = 1.329e-5, sutherland=None, delta_sol=None, cv: float = 720.19471) -> dict:
    """The point and cell data of write_sol_vtu (ns_io_module.F90:582-870) for
    the snapshot's state, as {"point": [(name, array, ncomp)], "cell": [...]}.

    Nodal_Velocity and Nodal_Pressure: compute_nodal_velocity_LS and
    compute_nodal_pressure_LS (ns_euler_zb_module.F90:9-83, 214-268) with the
    explicit path's h_p = 0: the sub-element-volume averages of the (second
    order: reconstructed) cell velocities and pressures, the velocity
    projected off the wall normal (wall_normal, ns_euler_primitives_module.F90:
    162-190) at boundary vertices.  Nodal_Grad_*: compute_nodal_grad (the
    device's nodal gradients).  Ducros: omega_ducros (ns_euler_recon_module.F90:
    535-558) with compute_vert_div_rot (:492-533), whose division reads the
    volume of vertex number n_sub_faces + 1 (its loop index after the loop),
    as written there.  bc_kinds: the Euler BC ids of the tags (1 = wall);
    sutherland: (mu0, T0) for the mu field, else mu_p."""
    sol = np.asarray(snap["sol"], np.float64).reshape(-1, 5)
    C = sol.shape[0]
    xc = np.asarray(snap["elem_coord"], np.float64).reshape(-1, 3)
    X = np.asarray(snap["vert_coord"], np.float64).reshape(-1, 3)
    V = X.shape[0]
    vvol = np.asarray(snap["vert_volume"], np.float64)
    sevol = np.asarray(snap["sub_elem_volume"], np.float64)
    se_elem = np.asarray(snap["sub_elem_elem"], np.int64) - 1
    se_vert = np.asarray(snap["sub_elem_vert"], np.int64) - 1
    w = primitives(sol)
    a = np.sqrt(GAMMA * w[:, 4] / w[:, 0])
    grad = np.asarray(snap["cell_grad"], np.float64).reshape(C, 5, 3) if second_order else None
    # nodal velocity and pressure: sub-element volume averages
    ws = w[se_elem].copy()
    if second_order:
        ws += np.einsum("skd,sd->sk", grad[se_elem], X[se_vert] - xc[se_elem])
    vp = np.zeros((V, 3))
    pp = np.zeros(V)
    np.add.at(vp, se_vert, sevol[:, None] * ws[:, 1:4])
    np.add.at(pp, se_vert, sevol * ws[:, 4])
    vp /= vvol[:, None]
    pp /= vvol
    # wall normal at boundary vertices
    sfv, sf = _lists(snap["vert_n_sub_faces"], snap["vert_sub_face"])
    n_sf = np.asarray(snap["sub_face_norm"], np.float64).reshape(-1, 3)
    a_sf = np.asarray(snap["sub_face_area"], np.float64)
    face = np.asarray(snap["sub_face_face"], np.int64) - 1
    fl = np.asarray(snap["face_left"], np.int64) - 1
    fr = np.asarray(snap["face_right"], np.int64)
    re = fr[face[sf]]
    kinds = np.concatenate([[1], np.asarray(bc_kinds, np.int64)]) if len(bc_kinds) else np.array([1])
    wall = (re == 0) | ((re < 0) & (kinds[np.clip(-re, 0, len(kinds) - 1)] == 1))
    if boundary_2d:
        wall &= ~(np.abs(n_sf[sf, 2]) > 1e-12)
    Bp = np.zeros((V, 3))
    np.add.at(Bp, sfv[wall], a_sf[sf[wall], None] * n_sf[sf[wall]])
    bound = np.asarray(snap["vert_is_bound"]) != 0
    nb = np.sqrt((Bp * Bp).sum(1))
    proj = bound & (nb > 1e-12)
    B = Bp[proj] / nb[proj, None]
    vp[proj] -= (vp[proj] * B).sum(1)[:, None] * B
    vp[~np.isfinite(vp).all(1)] = 0.0
    pp[~np.isfinite(pp)] = 0.0
    ng = np.asarray(snap["nodal_grad"], np.float64).reshape(V, 5, 3)
    # Ducros sensor
    inner = re > 0
    du = (w[re[inner] - 1, 1:4] - w[fl[face[sf[inner]]], 1:4])
    nn, aa = n_sf[sf[inner]], a_sf[sf[inner]]
    div = np.zeros(V)
    rot = np.zeros((V, 3))
    np.add.at(div, sfv[inner], aa * (nn * du).sum(1))
    np.add.at(rot, sfv[inner], aa[:, None] * np.cross(nn, du))
    nsfn = np.asarray(snap["vert_n_sub_faces"], np.int64)
    vq = vvol[np.minimum(nsfn, V - 1)]
    div /= vq
    rot /= vq[:, None]
    div[bound] = 0.0
    rot[bound] = 0.0
    ea = np.zeros(C)
    eb = np.zeros(C)
    np.add.at(ea, se_elem, sevol * np.abs(div[se_vert]) ** 2)
    np.add.at(eb, se_elem, sevol * (np.abs(div[se_vert]) ** 2 + 5 * (rot[se_vert] ** 2).sum(1)))
    ducros = ea / (eb + 1e-4)
    tau = 1.0 / sol[:, 0]
    temp = (sol[:, 4] * tau - 0.5 * (sol[:, 1] ** 2 + sol[:, 2] ** 2 + sol[:, 3] ** 2) * tau ** 2) / cv
    if sutherland is not None:
        mu0, T0 = sutherland
        mu = mu0 * (temp / T0) ** 1.5 * (T0 + 110.4) / (temp + 110.4)
    else:
        mu = np.full(C, mu_p)
    u = sol[:, 1:4] / sol[:, 0:1]
    speed = np.sqrt((u * u).sum(1))
    point = [("Nodal_Velocity", vp, 3), ("Nodal_Pressure", pp, 1), ("Nodal_Length", np.zeros((V, 3)), 3),
             ("Nodal_Grad_Density", ng[:, 0], 3), ("Nodal_Grad_Velocity_X", ng[:, 1], 3),
             ("Nodal_Grad_Velocity_Y", ng[:, 2], 3), ("Nodal_Grad_Velocity_Z", ng[:, 3], 3),
             ("Nodal_Grad_Pressure", ng[:, 4], 3), ("Hessian Mach", np.zeros((V, 9)), 9),
             ("Hessian Temp", np.zeros((V, 9)), 9)]
    cell = [("Internal_energy", sol[:, 4] / sol[:, 0] - 0.5 * speed ** 2, 1),
            ("H", sol[:, 4] / sol[:, 0] + w[:, 4] / sol[:, 0], 1), ("Mach", speed / a, 1),
            ("Temperature", temp, 1), ("mu", mu, 1), ("Ducros", ducros, 1),
            ("r", np.sqrt((xc * xc).sum(1)), 1), ("rxy", np.sqrt(xc[:, 0] ** 2 + xc[:, 1] ** 2), 1),
            ("Theta3", np.arctan2(-xc[:, 1], -xc[:, 0]), 1)]
    if second_order:
        for k, nm in enumerate(("Density_Grad", "Velocity_X_Grad", "Velocity_Y_Grad", "Velocity_Z_Grad",
                                "Pressure_Grad")):
            cell.append((nm, grad[:, k], 3))
    res = np.zeros(C) if delta_sol is None else np.sqrt((np.asarray(delta_sol).reshape(-1, 5) ** 2).sum(1))
    cell.append(("Residual", res, 1))
    return {"point": point, "cell": cell}


def write_sol_vtu(path: str, vert_coord: np.ndarray, elem_kind: np.ndarray, elem_vert: np.ndarray,
                  sol: np.ndarray, elem_coord: np.ndarray | None = None, elem_tag: np.ndarray | None = None,
                  cv: float = 720.19471, fields: dict | None = None) -> None:
    """vert_coord (V, 3); elem_kind gmsh kinds; elem_vert the elements'
    1-based vertex ids, concatenated; sol (C, 5) conservative; fields: the
    further point and cell data of reference_fields."""
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
        fh.write("</Cells>\n")
        if fields:
            fh.write("<PointData>\n")
            for name, data, nc in fields["point"]:
                _array(fh, name, data, nc)
            fh.write("</PointData>\n")
        fh.write("<CellData>\n")
        if elem_coord is not None:
            _array(fh, "Centroid", elem_coord, 3)
        if elem_tag is not None:
            _array(fh, "Tag", elem_tag, 1, "Int32", "%d")
            _array(fh, "MPI_Color", np.zeros(kind.size, np.int64), 1, "Int32", "%d")
        _array(fh, "Density", sol[:, 0], 1)
        _array(fh, "Momentum", sol[:, 1:4], 3)
        _array(fh, "Energy", sol[:, 4] / sol[:, 0], 1)  # per unit mass, as the reference writes it
        _array(fh, "Velocity", w[:, 1:4], 3)
        _array(fh, "Pressure", w[:, 4], 1)
        if fields:
            for name, data, nc in fields["cell"]:
                _array(fh, name, data, nc)
        else:
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


def solver_snapshot(solver, gradients: bool = False, method: int = 4) -> dict:
    """Host copies of what the writers need from a context: the state and
    the mesh arrays in build_mesh numbering; with gradients, also the nodal
    and cell gradients of the state (computed on the device) and the mesh
    lists reference_fields reads."""
    snap = {"sol": solver.get_sol(), "elem_coord": solver.elem_coord, "vert_coord": solver.mesh_d("vert_coord"),
            "elem_kind": solver.mesh_i("elem_kind"), "elem_vert": solver.mesh_i("elem_vert"),
            "elem_tag": solver.mesh_i("elem_tag")}
    if gradients:
        g, ng = solver.compute_cell_grad_from_nodal_grad(snap["sol"], method)
        snap["cell_grad"], snap["nodal_grad"] = g, ng
        for nm in ("vert_n_sub_faces", "vert_sub_face", "sub_face_face", "face_left", "face_right",
                   "sub_elem_elem", "sub_elem_vert", "vert_is_bound"):
            snap[nm] = solver.mesh_i(nm)
        for nm in ("vert_volume", "sub_elem_volume", "sub_face_norm", "sub_face_area"):
            snap[nm] = solver.mesh_d(nm)
    return snap


class AsyncWriter:
    """Writes on host threads; submit() returns at once with a Future."""

    def __init__(self, workers: int = 2):
        self.ex = ThreadPoolExecutor(max_workers=workers)

    def dat(self, path: str, snap: dict, box=None) -> Future:
        return self.ex.submit(write_sol_dat, path, snap["elem_coord"], snap["sol"], box)

    def vtu(self, path: str, snap: dict, **field_kw) -> Future:
        """field_kw: reference_fields' options; the fields are written when
        the snapshot holds the gradients (solver_snapshot(..., gradients=True))."""
        def job():
            fields = reference_fields(snap, **field_kw) if "nodal_grad" in snap else None
            write_sol_vtu(path, snap["vert_coord"], snap["elem_kind"], snap["elem_vert"], snap["sol"],
                          snap.get("elem_coord"), snap.get("elem_tag"), fields=fields)
        return self.ex.submit(job)

    def close(self):
        self.ex.shutdown(wait=True)
