"""Writes tests/golden/sedov_2d_analytic.txt from the analytic Sedov profile
that the reference's own Sedov test holds (test/sedov_tri/analytic_sedov_2D.dat,
columns: index, r, density, ...): the (r, density) pairs, as data.  Run in
the container that has the reference tree; the fixture is committed."""
import os
import sys

import numpy as np

SRC = "/root/reference/test/sedov_tri/analytic_sedov_2D.dat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sedov_2d_analytic.txt")

if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    d = np.loadtxt(src)
    np.savetxt(OUT, d[:, 1:3], fmt="%.6e", header="r density (analytic Sedov blast, 2-D, E = 0.984042, t = 1)")
