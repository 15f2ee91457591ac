"""Builds libsubfv_hip.so (gfx950) in-tree with hipcc."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsubfv_hip.so")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # every expression is evaluated as written (bit-exact FP64 with the
    # reference arithmetic); no packed-FP32 instructions
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


FORTRAN = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FORTRAN, "build")
# the Fortran host: iso_c_binding module, the reference-named procedures and
# the subfvns driver; FP contraction off as for the kernels
F_FLAGS = ["-O2", "-ffp-contract=off"]
F_MODULES = ["subfv_c.f90", "subfv_ns_gpu.f90", "mesh_module_counts.f90", "subfv_reference_api.f90"]
F_PROGRAMS = ["subfvns_gpu.f90", "mesh_check.f90", "reference_api_check.f90"]


def find_flang():
    for c in ("flang", "amdflang", "/opt/rocm/bin/amdflang", "/opt/rocm/lib/llvm/bin/flang"):
        p = shutil.which(c)
        if p:
            return p
    return None


def build_fortran(verbose: bool = False):
    """Compiles the Fortran host against libsubfv_hip.so; returns the
    executables built (none when no Fortran compiler is installed)."""
    fc = find_flang()
    if fc is None:
        return []
    os.makedirs(FBUILD, exist_ok=True)
    objs = []
    for m in F_MODULES:
        o = os.path.join(FBUILD, m.replace(".f90", ".o"))
        cmd = [fc, *F_FLAGS, "-J", FBUILD, "-c", os.path.join(FORTRAN, m), "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(o)
    exes = []
    for p in F_PROGRAMS:
        exe = os.path.join(FBUILD, p.replace(".f90", ""))
        cmd = [fc, *F_FLAGS, "-J", FBUILD, os.path.join(FORTRAN, p), *objs, "-o", exe, "-L", HERE,
               "-lsubfv_hip", "-Wl,-rpath,$ORIGIN/../.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        exes.append(exe)
    return exes


# flux_fast.hip (the node sweep of the SUBFV_ARITH_FAST mode) is compiled on
# its own: products contracted into fused multiply-adds, divisions as products
# with reciprocals, no NaN and no signed-zero semantics
FAST_FLAGS = [f for f in HIP_FLAGS if f not in ("-shared", "-ffp-contract=off")] + [
    "-ffp-contract=fast", "-freciprocal-math", "-fno-signed-zeros", "-fno-honor-nans"]
OBJ = os.path.join(HERE, "_obj")


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    fast_obj = os.path.join(OBJ, "flux_fast.o")
    cmd = ["hipcc", *FAST_FLAGS, "-c", "-o", fast_obj, os.path.join(CSRC, "flux_fast.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    srcs = [os.path.join(CSRC, "capi.hip"), os.path.join(CSRC, "mesh_device.hip"), os.path.join(CSRC, "mesh_build.cpp")]
    tmp = LIB + ".tmp"
    cmd = ["hipcc", *HIP_FLAGS, "-o", tmp, *srcs, "-x", "none", fast_obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
