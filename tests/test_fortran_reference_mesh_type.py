"""The reference-signature Fortran module (subfv_amd/fortran/
subfv_reference_api.f90) and the driver that calls it compile against the
reference's own mesh_module (the mesh_type of src/mesh/mesh_module.F90:17-88,
compiled into oracle/_ref by the oracle's recipe) in place of
mesh_module_counts.f90: inside the reference's build the procedures read
and fill the reference's mesh_type.  Compile only; skipped where the
reference's modules or flang are absent."""
import os
import shutil
import subprocess

import pytest

from subfv_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
FLANG = B.find_flang()
pytestmark = pytest.mark.skipif(
    FLANG is None or not os.path.exists(os.path.join(REF, "mesh_module.mod")),
    reason="the reference's compiled mesh_module (oracle/_ref) or flang is absent")


def test_reference_api_compiles_against_the_reference_mesh_type(tmp_path):
    for mod in ("mesh_module.mod", "precision_module.mod"):
        shutil.copy(os.path.join(REF, mod), tmp_path)
    for src in ("subfv_c.f90", "subfv_ns_gpu.f90", "subfv_reference_api.f90", "reference_api_check.f90"):
        out = subprocess.run([FLANG, "-c", "-J", str(tmp_path), os.path.join(B.FORTRAN, src),
                              "-o", str(tmp_path / (src + ".o"))], capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, src + "\n" + out.stdout + out.stderr
