"""build_mesh's connectivity built on the device (subfv_amd/csrc/mesh_device.hip)
against the host build (mesh_build.cpp, itself held to the oracle's build):
every connectivity array the mesh handle exports must be identical -- vertex
and element renumbering (breadth-first order, including its restart on a
mesh in two pieces), faces and their neighbours and boundary tags, vertex-
major sub-faces, cell-major sub-elements, their lists, element tags and
neighbours by vertex -- on hex, prism, tet and mixed meshes, on meshes whose
file orders are shuffled, and on the 1024 x 1024 prism mesh."""
import time

import numpy as np
import pytest

from subfv_amd import meshgen as mg
from subfv_amd.api import HostMesh, Solver

pytestmark = pytest.mark.gpu

NAMES = ["vert_nto", "elem_nto", "vert_n_sub_faces", "vert_n_sub_elems", "vert_n_elems", "vert_sub_elem",
         "vert_elem", "elem_n_vert", "elem_kind", "elem_vert", "elem_face", "elem_tag", "elem_n_neigh_by_vert",
         "elem_neigh_by_vert", "face_n_vert", "face_vert", "face_left", "face_right", "face_sub_face",
         "sub_face_vert", "sub_face_face", "sub_face_loc", "sub_face_left_elem", "sub_face_right_elem",
         "sub_face_left_sub_elem", "sub_face_right_sub_elem", "sub_elem_loc", "sub_elem_vert", "sub_elem_elem",
         "sub_elem_sub_face"]


def _shuffled(mesh, seed):
    """The same mesh with its vertices and elements in another file order."""
    rng = np.random.default_rng(seed)
    nv, ne = mesh.n_vert, mesh.n_elems
    perm = rng.permutation(nv)           # new file id of old vertex i
    inv = np.argsort(perm)
    off = mg.elem_offsets(mesh.kind)
    order = rng.permutation(ne)
    ev = np.concatenate([perm[mesh.elem_vert[off[e]:off[e + 1]] - 1] + 1 for e in order]).astype(np.int32)
    out = mg.Mesh(mesh.coord[inv], mesh.kind[order].copy(), ev, boundary_2d=mesh.boundary_2d, name=mesh.name + "s")
    if mesh.bf_nvert.size:
        out.bf_nvert = mesh.bf_nvert.copy()
        out.bf_vert = (perm[mesh.bf_vert - 1] + 1).astype(np.int32)
        out.bf_tag = mesh.bf_tag.copy()
    return out


def _two_pieces():
    """Two quad meshes side by side that share no vertex: the breadth-first
    numbering restarts at the lowest unvisited vertex."""
    a, b = mg.quad_mesh(4), mg.quad_mesh(3, x0=2.0)
    coord = np.concatenate([a.coord, b.coord])
    ev = np.concatenate([a.elem_vert, b.elem_vert + a.n_vert]).astype(np.int32)
    return mg.Mesh(coord, np.concatenate([a.kind, b.kind]), ev, boundary_2d=True, name="two")


def _fan():
    k = 5
    ang = 2 * np.pi * np.arange(k) / k + 0.1
    pts = np.concatenate([[[0.0, 0.0]], np.stack([np.cos(ang), np.sin(ang)], 1)])
    coord = np.concatenate([np.column_stack([pts, np.full(len(pts), 0.4 * z)]) for z in range(3)])
    ev = []
    for lay in range(2):
        o, o2 = lay * len(pts), (lay + 1) * len(pts)
        for q in range(k):
            ev.append([o, o + 1 + q, o + 1 + (q + 1) % k, o2, o2 + 1 + q, o2 + 1 + (q + 1) % k])
    return mg.Mesh(coord, np.full(len(ev), 6, np.int32), (np.array(ev) + 1).astype(np.int32).ravel(),
                   boundary_2d=False, name="fan")


CASES = {
    "quad9": lambda: mg.quad_mesh(9),
    "quad8lr": lambda: mg.quad_mesh(8, tags="lr"),
    "quad12lrbt": lambda: mg.quad_mesh(12, tags="lrbt"),
    "tri7": lambda: mg.tri_mesh(7),
    "tri33": lambda: mg.tri_mesh(33),
    "tet3": lambda: mg.tet_mesh(3),
    "tet5": lambda: mg.tet_mesh(5),
    "hex4": lambda: mg.hex_mesh(4),
    "fan": _fan,
    "two_pieces": _two_pieces,
    "quad10lrbt_shuffled": lambda: _shuffled(mg.quad_mesh(10, tags="lrbt"), 1),
    "tri9_shuffled": lambda: _shuffled(mg.tri_mesh(9), 2),
    "tet4_shuffled": lambda: _shuffled(mg.tet_mesh(4), 3),
}


@pytest.fixture(scope="module")
def solver():
    s = Solver(0)
    yield s
    s.close()


def _compare(host, dev):
    for name in NAMES:
        a, b = host.geti(name), dev.geti(name)
        assert a.shape == b.shape and np.array_equal(a, b), name


@pytest.mark.parametrize("name", list(CASES))
def test_device_build_equals_host_build(solver, name):
    print(f"[gpu] device mesh build {name}", flush=True)
    mesh = CASES[name]()
    _compare(HostMesh(mesh), HostMesh(mesh, solver=solver))


def test_device_build_equals_host_build_tri1024(solver):
    print("[gpu] device mesh build tri1024", flush=True)
    mesh = mg.tri_mesh(1024)
    t0 = time.perf_counter()
    host = HostMesh(mesh)
    t1 = time.perf_counter()
    dev = HostMesh(mesh, solver=solver)
    t2 = time.perf_counter()
    print(f"  host build {t1 - t0:.3f} s, device build {t2 - t1:.3f} s", flush=True)
    _compare(host, dev)
