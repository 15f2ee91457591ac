"""BCSR SpMV, block-Jacobi and GMRES on the device against the oracle
(sparse_bcsr_linear_module.F90:13-29, 1414-1436, 1568-1691).

SpMV and block-Jacobi in FP64 evaluate the oracle's expressions in its order
and are compared bit for bit.  GMRES reduces its dot products in a fixed
two-level order that differs from the sequential one, so its iterates are
compared within tolerance: both must reach ||b - A x|| < tol, and the
solutions agree to the conditioning of the system.  The FP32-inner mode
(SpMV and block-Jacobi in FP32 inside each cycle, FP64 restart residuals)
must reach the same FP64 tolerance; its error against the FP64 solution is
stated in the test.
"""
import numpy as np
import pytest

import oracle as O
from subfv_amd import lowmach as lm
from subfv_amd.api import Bcsr, Solver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    s = Solver(0)
    yield s
    s.close()


def random_bcsr(m, b, seed, per_row=5, diag_shift=2.0):
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(m):
        nb = sorted(set([i] + list(rng.integers(0, m, per_row - 1))))
        for j in nb:
            blk = rng.standard_normal((b, b)) * 0.3
            if j == i:
                blk = blk + diag_shift * b * np.eye(b)
            rows.append(i)
            cols.append(j)
            vals.append(blk)
    row = np.zeros(m + 1, np.int64)
    for i in rows:
        row[i + 1] += 1
    row = (np.cumsum(row) + 1).astype(np.int32)
    v = np.stack([blk.T for blk in vals]).reshape(-1)
    return row, np.array(cols, np.int32) + 1, v


@pytest.mark.parametrize("b", [3, 4, 5])
def test_spmv_and_block_jacobi_bit_exact(solver, b):
    print(f"[gpu] bcsr spmv/diag b={b}", flush=True)
    m = 300
    row, col, v = random_bcsr(m, b, 10 + b)
    A = Bcsr(solver, m, b, row, col, v)
    x = np.random.default_rng(b).standard_normal(m * b)
    assert np.array_equal(A.matmul(x), O.bcsr_matmul(m, b, row, col, v, x))
    assert np.array_equal(A.diag(x), O.bcsr_diag(m, b, row, col, v, x))
    A.close()


def test_spmv_bit_exact_on_low_mach_system(solver):
    print("[gpu] bcsr low-Mach spmv", flush=True)
    s = lm.build_system(64)
    A = Bcsr(solver, s["m"], 4, s["row"], s["col"], s["v"])
    x = np.random.default_rng(3).standard_normal(s["m"] * 4)
    assert np.array_equal(A.matmul(x), O.bcsr_matmul(s["m"], 4, s["row"], s["col"], s["v"], x))
    assert np.array_equal(A.diag(x), O.bcsr_diag(s["m"], 4, s["row"], s["col"], s["v"], x))
    A.close()


def test_gmres_fp64_matches_oracle(solver):
    print("[gpu] bcsr gmres fp64", flush=True)
    s = lm.build_system(48)
    rhs = s["rhs"]
    tol = 1e-10 * np.linalg.norm(rhs)
    A = Bcsr(solver, s["m"], 4, s["row"], s["col"], s["v"])
    x, conv, st = A.gmres_diag(rhs, tol, 30, 0)
    xo, convo, sto = O.bcsr_gmres_diag(s["m"], 4, s["row"], s["col"], s["v"], rhs, tol, 30)
    assert conv and convo
    res = np.linalg.norm(rhs - O.bcsr_matmul(s["m"], 4, s["row"], s["col"], s["v"], x))
    assert res < tol
    # both solutions satisfy the same residual bound; they agree to ~cond * tol
    assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)
    A.close()


def test_gmres_fp32_inner_reaches_fp64_tolerance(solver):
    print("[gpu] bcsr gmres fp32 inner", flush=True)
    s = lm.build_system(48)
    rhs = s["rhs"]
    tol = 1e-10 * np.linalg.norm(rhs)
    A = Bcsr(solver, s["m"], 4, s["row"], s["col"], s["v"])
    x64, c64, _ = A.gmres_diag(rhs, tol, 30, 0)
    x32, c32, st32 = A.gmres_diag(rhs, tol, 30, 1)
    assert c64 and c32, st32
    res = np.linalg.norm(rhs - O.bcsr_matmul(s["m"], 4, s["row"], s["col"], s["v"], x32))
    assert res < tol
    # stated bound: the FP32-inner solution is within 1e-7 (relative) of the FP64 one,
    # of the device and of the oracle (both FP64 solutions meet the same residual bound)
    xo, convo, _ = O.bcsr_gmres_diag(s["m"], 4, s["row"], s["col"], s["v"], rhs, tol, 30)
    assert convo
    e_dev = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    e_ora = np.linalg.norm(x32 - xo) / np.linalg.norm(xo)
    print(f"  FP32-inner vs FP64: device {e_dev:.2e}, oracle {e_ora:.2e}", flush=True)
    assert e_dev <= 1e-7 and e_ora <= 1e-7
    A.close()


def test_gmres_on_random_system_matches_direct_solve(solver):
    print("[gpu] bcsr gmres random", flush=True)
    m, b = 200, 4
    row, col, v = random_bcsr(m, b, 77, diag_shift=3.0)
    rhs = np.random.default_rng(8).standard_normal(m * b)
    tol = 1e-11 * np.linalg.norm(rhs)
    A = Bcsr(solver, m, b, row, col, v)
    x, conv, _ = A.gmres_diag(rhs, tol, 20, 0)
    assert conv
    D = np.zeros((m * b, m * b))
    rr = row - 1
    for i in range(m):
        for j in range(rr[i], rr[i + 1]):
            c = col[j] - 1
            D[i * b:(i + 1) * b, c * b:(c + 1) * b] = v[j * b * b:(j + 1) * b * b].reshape(b, b).T
    assert np.allclose(x, np.linalg.solve(D, rhs), rtol=1e-8, atol=1e-10)
    A.close()


@pytest.mark.parametrize("m,b,precision", [(201, 3, 0), (201, 3, 1), (77, 5, 0), (77, 5, 1), (150, 4, 1)])
def test_gmres_odd_and_even_lengths_match_direct_solve(solver, m, b, precision):
    """Odd n = m b: the Gram-Schmidt passes take two entries per thread, the
    last entry of an odd n goes to one thread, and the basis columns are n
    rounded up to even apart.  FP64 and FP32-inner (FP32 basis) against a
    direct solve, both to the FP64 residual bound."""
    print(f"[gpu] bcsr gmres m={m} b={b} precision={precision}", flush=True)
    row, col, v = random_bcsr(m, b, 91 + m + b, diag_shift=3.0)
    rhs = np.random.default_rng(m + b).standard_normal(m * b)
    tol = 1e-11 * np.linalg.norm(rhs)
    A = Bcsr(solver, m, b, row, col, v)
    x, conv, st = A.gmres_diag(rhs, tol, 20, precision)
    assert conv, st
    D = np.zeros((m * b, m * b))
    rr = row - 1
    for i in range(m):
        for j in range(rr[i], rr[i + 1]):
            c = col[j] - 1
            D[i * b:(i + 1) * b, c * b:(c + 1) * b] = v[j * b * b:(j + 1) * b * b].reshape(b, b).T
    assert np.linalg.norm(rhs - D @ x) < tol
    assert np.allclose(x, np.linalg.solve(D, rhs), rtol=1e-8, atol=1e-10)
    A.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_gmres_cycle_graph_replays_are_deterministic(solver, precision):
    """Each Arnoldi cycle is a captured graph replayed at every restart: the
    same solve twice (the second replays the first's graph), a solve with
    another restart length (recaptured), then the first again (recaptured
    back) give the same bits, with eager host work between the replays."""
    print(f"[gpu] bcsr gmres graph replays precision={precision}", flush=True)
    m, b = 300, 4
    row, col, v = random_bcsr(m, b, 55, diag_shift=0.6)
    rhs = np.random.default_rng(56).standard_normal(m * b)
    tol = 1e-10 * np.linalg.norm(rhs)
    A = Bcsr(solver, m, b, row, col, v)
    x1, c1, st1 = A.gmres_diag(rhs, tol, 20, precision)
    assert c1 and st1["spmv"] > 20  # several cycles: the graph is replayed
    x2, c2, _ = A.gmres_diag(rhs, tol, 20, precision)
    x3, c3, _ = A.gmres_diag(rhs, tol, 12, precision)
    x4, c4, _ = A.gmres_diag(rhs, tol, 20, precision)
    assert c2 and c3 and c4
    assert np.array_equal(x1, x2) and np.array_equal(x1, x4)
    assert np.linalg.norm(x3 - x1) <= 1e-7 * np.linalg.norm(x1)
    A.close()


def test_gmres_reports_non_convergence(solver):
    print("[gpu] bcsr gmres not converged", flush=True)
    s = lm.build_system(32)
    A = Bcsr(solver, s["m"], 4, s["row"], s["col"], s["v"])
    x, conv, st = A.gmres_diag(s["rhs"], 1e-300, 2, 0)  # unreachable tolerance
    assert not conv
    assert st["restarts"] == 40
    A.close()
