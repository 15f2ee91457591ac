"""The host msh 2.2 reader (meshgen.read_msh, the semantics of
read_mesh_msh_2, src/mesh/mesh_reading_module.F90:989-1324): round trip of
write_msh, boundary tags by BC name, the reference's stops."""
import numpy as np
import pytest

from subfv_amd import meshgen as mg


@pytest.mark.parametrize("mesh,bcs", [(mg.quad_mesh(5, tags="lr"), ("left", "right")), (mg.tri_mesh(4), ()),
                                      (mg.tet_mesh(2), ())])
def test_round_trip(tmp_path, mesh, bcs):
    path = str(tmp_path / "m.msh")
    mg.write_msh(mesh, path, bcs)
    r = mg.read_msh(path, bcs, mesh.boundary_2d)
    for f in ("coord", "kind", "elem_vert", "bf_nvert", "bf_vert", "bf_tag"):
        assert np.array_equal(getattr(r, f), getattr(mesh, f)), f
    assert r.boundary_2d == mesh.boundary_2d


def test_tags_follow_the_bc_names(tmp_path):
    mesh = mg.quad_mesh(5, tags="lr")
    path = str(tmp_path / "m.msh")
    mg.write_msh(mesh, path, ("left", "right"))
    # listed in the other order, the indices swap; an unlisted group is the wall
    r = mg.read_msh(path, ("right", "left"))
    assert np.array_equal(r.bf_tag, np.where(mesh.bf_tag == 1, 2, np.where(mesh.bf_tag == 2, 1, 0)))
    r = mg.read_msh(path, ("left",))
    assert np.array_equal(r.bf_tag, np.where(mesh.bf_tag == 1, 1, 0))


def test_reference_stops(tmp_path):
    path = tmp_path / "nophys.msh"
    path.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$EndNodes\n"
                    "$Elements\n0\n$EndElements\n")
    with pytest.raises(ValueError, match="No Physical names"):
        mg.read_msh(str(path), ("wall",))
    with pytest.raises(ValueError, match="No 3D elements"):
        mg.read_msh(str(path))
