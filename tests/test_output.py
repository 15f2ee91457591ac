"""Solution output (subfv_amd/output.py): write_sol_dat's table
(ns_io_module.F90:50-96) and the VTU grid with the reference's state names,
read back; the asynchronous writer."""
import xml.etree.ElementTree as ET

import numpy as np

from subfv_amd import meshgen as mg
from subfv_amd import output as OUT


def _state(n, seed=3):
    rng = np.random.default_rng(seed)
    w = np.stack([1 + rng.random(n), rng.random(n) - 0.5, rng.random(n) - 0.5, 0.1 * rng.random(n),
                  1 + rng.random(n)], 1)
    return mg.primit_to_conserv(w), w


def _centroids(mesh):
    off = mg.elem_offsets(mesh.kind)
    return np.array([mesh.coord[mesh.elem_vert[off[e]:off[e + 1]] - 1].mean(0) for e in range(mesh.n_elems)])


def test_dat_table_and_box(tmp_path):
    mesh = mg.tri_mesh(6)
    sol, w = _state(mesh.n_elems)
    xc = _centroids(mesh)
    n = OUT.write_sol_dat(str(tmp_path / "all.dat"), xc, sol)
    text = (tmp_path / "all.dat").read_text()
    assert text.startswith("\n") and n == mesh.n_elems
    rows = np.loadtxt(tmp_path / "all.dat").reshape(n, 13)
    assert np.array_equal(rows[:, :3], xc) and np.array_equal(rows[:, 3:8], sol)
    assert np.allclose(rows[:, 8:], w, rtol=1e-14, atol=1e-14)
    box = (0.0, 0.5, 0.25, 1.0, -1.0, 1.0)  # strict inequalities, as the reference
    n = OUT.write_sol_dat(str(tmp_path / "box.dat"), xc, sol, box)
    inside = (xc[:, 0] > 0.0) & (xc[:, 0] < 0.5) & (xc[:, 1] > 0.25) & (xc[:, 1] < 1.0)
    assert n == inside.sum()
    assert np.array_equal(np.loadtxt(tmp_path / "box.dat").reshape(n, 13)[:, 3:8], sol[inside])


def _arrays(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {a.get("Name"): np.array(a.text.split(), float) for a in root.iter("DataArray")}
    return piece, out


def test_vtu_mixed_kinds(tmp_path):
    for mesh in (mg.quad_mesh(3), mg.tri_mesh(3), mg.tet_mesh(2)):
        sol, w = _state(mesh.n_elems)
        path = str(tmp_path / f"{mesh.name}.vtu")
        OUT.write_sol_vtu(path, mesh.coord, mesh.kind, mesh.elem_vert, sol, _centroids(mesh))
        piece, a = _arrays(path)
        assert int(piece.get("NumberOfCells")) == mesh.n_elems and int(piece.get("NumberOfPoints")) == mesh.n_vert
        assert np.array_equal(a["Points"].reshape(-1, 3), mesh.coord)
        assert np.array_equal(a["connectivity"], mesh.elem_vert - 1)
        assert np.array_equal(a["offsets"], mg.elem_offsets(mesh.kind)[1:])
        assert set(a["types"]) == {OUT.VTK_TYPE[int(mesh.kind[0])]}
        assert np.array_equal(a["Density"], sol[:, 0]) and np.array_equal(a["Momentum"].reshape(-1, 3), sol[:, 1:4])
        assert np.allclose(a["Pressure"], w[:, 4], rtol=1e-14)
        assert np.allclose(a["Energy"], sol[:, 4] / sol[:, 0], rtol=1e-15)
        # restart_from_vtu_file multiplies the specific energy back by rho
        back = OUT.restart_from_vtu(path)
        assert np.array_equal(back[:, :4], sol[:, :4])
        assert np.allclose(back[:, 4], sol[:, 4], rtol=1e-15)


def test_async_writer(tmp_path):
    mesh = mg.quad_mesh(4)
    sol, _ = _state(mesh.n_elems)
    snap = {"sol": sol, "elem_coord": _centroids(mesh), "vert_coord": mesh.coord, "elem_kind": mesh.kind,
            "elem_vert": mesh.elem_vert}
    wr = OUT.AsyncWriter(workers=2)
    futs = [wr.dat(str(tmp_path / f"s{i}.dat"), snap) for i in range(4)] + [wr.vtu(str(tmp_path / "s.vtu"), snap)]
    wr.close()
    assert all(f.done() and f.exception() is None for f in futs)
    assert all((tmp_path / f"s{i}.dat").exists() for i in range(4)) and (tmp_path / "s.vtu").exists()


def _oracle_snapshot(om, sol, method=4):
    """reference_fields' inputs from the oracle (its numbering is build_mesh's)."""
    g, ng = om.compute_cell_grad(sol, method)
    snap = {"sol": sol, "elem_coord": om.elem_coord, "vert_coord": om.getd("vert_coord"), "cell_grad": g,
            "nodal_grad": ng}
    for nm in ("vert_n_sub_faces", "vert_sub_face", "sub_face_face", "face_left", "face_right", "sub_elem_elem",
               "sub_elem_vert", "vert_is_bound"):
        snap[nm] = om.geti(nm)
    for nm in ("vert_volume", "sub_elem_volume", "sub_face_norm", "sub_face_area"):
        snap[nm] = om.getd(nm)
    return snap


def test_reference_fields_closed_forms():
    """Uniform flow: the nodal velocity is the flow velocity at interior
    vertices and its wall-tangent part at wall vertices, the nodal pressure
    the pressure, the Ducros sensor zero; a pure rotation u = (-y, x) has
    zero divergence, so Ducros stays zero, while a pure expansion u = (x, y)
    gives a positive sensor below one; derived cell fields follow their
    definitions."""
    import oracle as O

    mesh = mg.tri_mesh(8)
    om = O.OracleMesh(mesh)
    xc = om.elem_coord
    n = om.n_elems
    w = np.tile([1.2, 0.3, -0.2, 0.0, 0.9], (n, 1))
    f = OUT.reference_fields(_oracle_snapshot(om, mg.primit_to_conserv(w)))
    pt = {k: v for k, v, _ in f["point"]}
    ce = {k: v for k, v, _ in f["cell"]}
    bound = om.geti("vert_is_bound") != 0
    assert np.allclose(pt["Nodal_Velocity"][~bound], [0.3, -0.2, 0.0], atol=1e-14)
    assert np.allclose(pt["Nodal_Pressure"], 0.9, atol=1e-14)
    # on the square's sides the wall-normal part is removed
    X = om.getd("vert_coord").reshape(-1, 3)
    side = bound & (np.abs(X[:, 0]) < 1e-12) & (X[:, 1] > 0.05) & (X[:, 1] < 0.95)
    assert side.any() and np.allclose(pt["Nodal_Velocity"][side, 0], 0.0, atol=1e-13)
    assert np.all(ce["Ducros"] == 0.0)
    assert np.allclose(ce["Mach"], np.sqrt(0.3 ** 2 + 0.2 ** 2) / np.sqrt(1.4 * 0.9 / 1.2))
    assert np.allclose(ce["H"], (0.9 / 0.4 + 0.5 * 1.2 * 0.13) / 1.2 + 0.9 / 1.2)
    for case, expect in (("rot", 0.0), ("exp", None)):
        u = np.stack([-xc[:, 1], xc[:, 0]], 1) if case == "rot" else xc[:, :2].copy()
        wv = np.column_stack([np.ones(n), 0.1 * u, np.zeros(n), np.ones(n)])
        d = {k: v for k, v, _ in OUT.reference_fields(_oracle_snapshot(om, mg.primit_to_conserv(wv)))["cell"]}
        if expect is not None:
            assert np.abs(d["Ducros"]).max() < 1e-12
        else:
            inner = d["Ducros"] > 0
            assert inner.sum() > n // 2 and d["Ducros"].max() < 1.0
