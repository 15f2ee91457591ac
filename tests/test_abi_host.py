"""CPU-side checks of the C-ABI library (no GPU needed).

- libsubfv_hip.so loads and exports every function include/subfv.h declares
  (nothing that needs a device is called here);
- the host connectivity build behind subfv_mesh_build (build_mesh,
  mesh_connectivity_module.F90:12-59) matches the oracle's integer arrays
  exactly on hex, prism and tet meshes.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from subfv_amd import _lib
from subfv_amd import meshgen as mg
from subfv_amd.api import HostMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "subfv.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(subfv_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_function():
    names = header_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    """Every declared function has a ctypes signature in the binding."""
    names = set(header_functions())
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))


MESHES = {
    "quad9": lambda: mg.quad_mesh(9),
    "quad8lr": lambda: mg.quad_mesh(8, tags="lr"),
    "tri7": lambda: mg.tri_mesh(7),
    "tet3": lambda: mg.tet_mesh(3),
    "quad17": lambda: mg.quad_mesh(17, seed=5),
    "tri12": lambda: mg.tri_mesh(12, seed=11),
}
HOST_NAMES = [n for n in O._INT_NAMES if n != "vert_is_bound"]  # is_bound needs the geometry


@pytest.mark.parametrize("name", list(MESHES))
def test_host_connectivity_matches_oracle(name):
    mesh = MESHES[name]()
    hm = HostMesh(mesh)
    om = O.OracleMesh(mesh)
    for n in HOST_NAMES:
        a, b = hm.geti(n), om.geti(n)
        assert a.shape == b.shape, (n, a.shape, b.shape)
        assert np.array_equal(a, b), n


def test_host_mesh_rejects_bad_input():
    mesh = mg.quad_mesh(3)
    ev = mesh.elem_vert.copy()
    ev[0] = mesh.n_vert + 5  # vertex id out of range
    bad = mg.Mesh(mesh.coord, mesh.kind, ev, mesh.bf_nvert, mesh.bf_vert, mesh.bf_tag, mesh.boundary_2d, "bad")
    with pytest.raises(Exception):
        HostMesh(bad)


def test_host_mesh_rejects_face_of_three_elements():
    """Three hexahedra on one quadrilateral (a non-manifold face) would leave
    an element's face unset: the build refuses the mesh."""
    coord = np.random.default_rng(0).random((16, 3))
    base = np.arange(8)
    ev = np.concatenate([base, [4, 5, 6, 7, 8, 9, 10, 11], [4, 5, 6, 7, 12, 13, 14, 15]]) + 1
    bad = mg.Mesh(coord, np.full(3, 5, np.int32), ev.astype(np.int32), name="three")
    with pytest.raises(Exception):
        HostMesh(bad)
