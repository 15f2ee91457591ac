"""The Fortran host (subfv_amd/fortran): its msh 2.2 reader (with the
reference's boundary-tag semantics) feeding the C-ABI host connectivity
build, compared with the same mesh built from Python.  No device needed."""
import os
import subprocess

import numpy as np
import pytest

from subfv_amd import build as B
from subfv_amd import meshgen as mg
from subfv_amd.api import HostMesh

EXE = os.path.join(B.FBUILD, "mesh_check")
pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built (no flang)")


@pytest.mark.parametrize("name,mesh,bcs", [
    ("quad", mg.quad_mesh(6), ()),
    ("quad_lr", mg.quad_mesh(5, tags="lr"), ("left", "right")),
    ("tri", mg.tri_mesh(5), ()),
    ("tet", mg.tet_mesh(2), ()),
])
def test_fortran_reader_and_host_build(tmp_path, name, mesh, bcs):
    path = str(tmp_path / f"{name}.msh")
    mg.write_msh(mesh, path, bcs)
    out = subprocess.run([EXE, path, *bcs], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    nv, ne, nbf, n1, n2 = (int(x) for x in out.stdout.split())
    assert (nv, ne, nbf) == (mesh.n_vert, mesh.n_elems, mesh.bf_nvert.size)
    raw = np.fromfile(tmp_path / "mesh_check.bin", dtype=np.int32)
    arrays, off = [], 0
    for _ in range(3):
        n = raw[off]
        arrays.append(raw[off + 1:off + 1 + n])
        off += 1 + n
    hm = HostMesh(mesh)
    # the Python reader gives the library the same mesh
    hm_py = HostMesh(mg.read_msh(path, bcs, mesh.boundary_2d))
    for k, name in enumerate(("sub_face_left_elem", "elem_vert", "face_right")):
        assert np.array_equal(arrays[k], hm_py.geti(name)), name
    assert np.array_equal(arrays[0], hm.geti("sub_face_left_elem"))
    assert np.array_equal(arrays[1], hm.geti("elem_vert"))
    # face_right carries minus the boundary index the reader gave each face
    assert np.array_equal(arrays[2], hm.geti("face_right"))
    assert (arrays[2] < 0).any() == bool(bcs)
