"""subfv_bcsr_gmres_diag_noalloc (Krylov basis in the library's fixed device
scratch array) against subfv_bcsr_gmres_diag on the same systems: identical
kernels in the same order, so the solutions are equal bit for bit.  Two
calls in a row, the second on a smaller system, so the second runs over the
first call's leftovers in the scratch array; a system too large for the
array is refused with a message."""
import numpy as np
import pytest

from subfv_amd.api import Bcsr, Solver, SubfvError

pytestmark = pytest.mark.gpu


def random_bcsr(m, b, seed, per_row=5, diag_shift=2.0):
    rng = np.random.default_rng(seed)
    row = [1]
    cols, vals = [], []
    for i in range(m):
        nb = sorted(set([i] + list(rng.integers(0, m, per_row - 1))))
        for j in nb:
            blk = rng.standard_normal((b, b)) * 0.3
            if j == i:
                blk = blk + diag_shift * b * np.eye(b)
            cols.append(j + 1)
            vals.append(blk.T.reshape(-1))
        row.append(row[-1] + len(nb))
    return np.array(row, np.int32), np.array(cols, np.int32), np.concatenate(vals)


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_gmres_equals_gmres_diag_twice(precision):
    print(f"[gpu] gmres noalloc precision={precision}", flush=True)
    s = Solver(0)
    try:
        for m, b, restart, seed in ((20000, 4, 30, 5), (701, 5, 12, 6)):  # n = 3505: odd
            row, col, v = random_bcsr(m, b, seed, diag_shift=0.6)
            A = Bcsr(s, m, b, row, col, v)
            rhs = np.random.default_rng(seed + 1).standard_normal(m * b)
            tol = 1e-10 * np.linalg.norm(rhs)
            x_ref, conv_ref, st_ref = A.gmres_diag(rhs, tol, restart, precision)
            x_na, conv_na, st_na = A.gmres_diag_noalloc(rhs, tol, restart, precision)
            A.close()
            assert conv_ref and conv_na
            assert st_na == st_ref
            assert st_ref["spmv"] > restart  # more than one cycle: restarts run over the scratch basis
            assert np.array_equal(x_na, x_ref), (m, b)
    finally:
        s.close()


def test_noalloc_gmres_rejects_a_system_beyond_the_scratch_array():
    print("[gpu] gmres noalloc size limit", flush=True)
    s = Solver(0)
    try:
        m, b, restart = 40000, 4, 30  # 160,000 * 31 doubles > 4 Mi
        row, col, v = random_bcsr(m, b, 9)
        A = Bcsr(s, m, b, row, col, v)
        with pytest.raises(SubfvError, match="scratch array"):
            A.gmres_diag_noalloc(np.ones(m * b), 1e-8, restart)
        A.close()
    finally:
        s.close()
