"""The other BCSR solvers on the device against the oracle
(sparse_bcsr_linear_module.F90: bcsr_lusgs :1268-1335, bcsr_parallel_lusgs
:1438-1502, bcsr_gmres :194-290, bcsr_gmres_lusgs :447-548,
bcsr_gmres_parallel_lusgs :885-1037, bcsr_pcgm :1110-1169, add_block_to_bcsr
:56-71).

The sweeps evaluate every row in the oracle's order (the rows of one
dependency level, or of one colour set, run together), so LU-SGS and the
coloured LU-SGS are compared bit for bit.  The Krylov solvers reduce their
dot products in a fixed two-level order, so, as for bcsr_gmres_diag, their
solutions are compared within tolerance and both must converge."""
import numpy as np
import pytest

import oracle as O
from subfv_amd import lowmach as lm
from subfv_amd.api import Bcsr, Solver, SubfvError
from test_bcsr_solvers_oracle import dense, greedy_colours, random_bcsr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    s = Solver(0)
    yield s
    s.close()


@pytest.mark.parametrize("b", [3, 4, 5])
def test_lusgs_bit_exact(solver, b):
    print(f"[gpu] bcsr lusgs b={b}", flush=True)
    nb = 400
    row, col, v = random_bcsr(nb, b, seed=100 + b)
    A = Bcsr(solver, nb, b, row, col, v)
    rhs = np.random.default_rng(b).standard_normal(nb * b)
    assert np.array_equal(A.lusgs(rhs), O.bcsr_lusgs(nb, b, row, col, v, rhs))
    A.close()


@pytest.mark.parametrize("b", [3, 4, 5])
def test_parallel_lusgs_bit_exact_with_given_and_greedy_colours(solver, b):
    print(f"[gpu] bcsr parallel lusgs b={b}", flush=True)
    nb = 400
    row, col, v = random_bcsr(nb, b, seed=200 + b)
    A = Bcsr(solver, nb, b, row, col, v)
    rhs = np.random.default_rng(b).standard_normal(nb * b)
    c = greedy_colours(nb, row, col)
    ref = O.bcsr_parallel_lusgs(nb, b, row, col, v, rhs, c)
    assert A.set_colors(c) == c.max()
    assert np.array_equal(A.parallel_lusgs(rhs), ref)
    # the library's greedy colouring is the same rule
    assert A.set_colors(None) == c.max()
    assert np.array_equal(A.parallel_lusgs(rhs), ref)
    A.close()


def test_sweeps_on_the_low_mach_system(solver):
    print("[gpu] bcsr sweeps low-Mach 48x48", flush=True)
    sysd = lm.build_system(n=48)
    m, b, row, col, v = sysd["m"], sysd["b"], sysd["row"], sysd["col"], sysd["v"]
    A = Bcsr(solver, m, b, row, col, v)
    rhs = sysd["rhs"]
    assert np.array_equal(A.lusgs(rhs), O.bcsr_lusgs(m, b, row, col, v, rhs))
    c = greedy_colours(m, row, col)
    A.set_colors(c)
    assert np.array_equal(A.parallel_lusgs(rhs), O.bcsr_parallel_lusgs(m, b, row, col, v, rhs, c))
    A.close()


@pytest.mark.parametrize("precond", ["none", "lusgs", "parallel_lusgs"])
def test_gmres_variants_match_the_oracle(solver, precond):
    print(f"[gpu] bcsr gmres {precond}", flush=True)
    nb, b = 300, 4
    row, col, v = random_bcsr(nb, b, seed=300)
    rhs = np.random.default_rng(7).standard_normal(nb * b)
    c = greedy_colours(nb, row, col)
    A = Bcsr(solver, nb, b, row, col, v)
    A.set_colors(c)
    tol = 1e-10
    x, conv, info = A.gmres(rhs, tol, restart=20, precond=precond)
    xo, convo, infoo = O.bcsr_gmres(nb, b, row, col, v, rhs, tol, restart=20, precond=precond, color=c)
    assert conv and convo
    Ad = dense(nb, b, row, col, v)
    assert np.linalg.norm(rhs - Ad @ x) < tol
    assert abs(info["restarts"] - infoo["restarts"]) <= 1
    assert np.allclose(x, xo, rtol=1e-8, atol=1e-10)
    A.close()


def test_gmres_lusgs_rejects_fp32_and_missing_colours(solver):
    nb, b = 20, 3
    row, col, v = random_bcsr(nb, b, seed=301)
    A = Bcsr(solver, nb, b, row, col, v)
    with pytest.raises(SubfvError):
        A.gmres(np.ones(nb * b), 1e-8, precond="lusgs", precision=1)
    with pytest.raises(SubfvError):
        A.gmres(np.ones(nb * b), 1e-8, precond="parallel_lusgs")
    A.close()


def test_pcgm_matches_the_oracle(solver):
    print("[gpu] bcsr pcgm", flush=True)
    nb, b = 300, 3
    row, col, v = random_bcsr(nb, b, seed=400, sym=True, extra=2)
    rhs = np.random.default_rng(8).standard_normal(nb * b)
    x0 = np.random.default_rng(9).standard_normal(nb * b) * 0.1
    A = Bcsr(solver, nb, b, row, col, v)
    x, conv, info = A.pcgm(rhs, 1e-11, x0=x0)
    xo, convo, infoo = O.bcsr_pcgm(nb, b, row, col, v, rhs, 1e-11, x0=x0)
    assert conv and convo == 1
    assert abs(info["iterations"] - infoo["iterations"]) <= 1
    assert np.allclose(x, xo, rtol=1e-8, atol=1e-10)
    # singular diagonal block: lu_solve's stop is reported
    v2 = v.copy()
    k = [q for q in range(row[0] - 1, row[1] - 1) if col[q] == 1][0]
    v2[k * 9:(k + 1) * 9] = np.diag([1.0, 1.0, 0.0]).ravel()
    A2 = Bcsr(solver, nb, b, row, col, v2)
    with pytest.raises(SubfvError):
        A2.pcgm(rhs, 1e-11)
    A.close()
    A2.close()


def test_add_block_matches_the_oracle(solver):
    print("[gpu] bcsr add_block", flush=True)
    nb, b = 50, 4
    row, col, v = random_bcsr(nb, b, seed=500)
    A = Bcsr(solver, nb, b, row, col, v)
    rng = np.random.default_rng(10)
    v_ref = v.copy()
    for (i, j) in [(3, int(col[row[2] - 1])), (7, 7), (1, 1)]:
        M = rng.standard_normal((b, b))
        A.add_block(i, j, M)
        v_ref = O.bcsr_add_block(nb, b, row, col, v_ref, i, j, M)
    x = rng.standard_normal(nb * b)
    assert np.array_equal(A.matmul(x), O.bcsr_matmul(nb, b, row, col, v_ref, x))
    assert np.array_equal(A.diag(x), O.bcsr_diag(nb, b, row, col, v_ref, x))  # refreshed factors
    A.close()
