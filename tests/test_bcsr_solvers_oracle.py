"""The oracle's restatements of the other BCSR solvers of
src/linear_algebra/sparse_bcsr_linear_module.F90 (bcsr_lusgs :1268-1335,
bcsr_parallel_lusgs :1438-1502, bcsr_gmres :194-290, bcsr_gmres_lusgs
:447-548, bcsr_gmres_parallel_lusgs :885-1037, bcsr_pcgm :1110-1169,
add_block_to_bcsr :56-71) pinned against closed forms built with numpy:
the sweeps are triangular solves of dense matrices assembled from the blocks,
the Krylov solvers converge to the direct solution."""
import numpy as np
import pytest

import oracle as O


def random_bcsr(nb, b, seed, sym=False, extra=3):
    """Block-sparse matrix with sorted columns, a diagonal block per row and
    `extra` random off-diagonal blocks per row (mirrored when sym)."""
    rng = np.random.default_rng(seed)
    pattern = [set([i]) for i in range(nb)]
    for i in range(nb):
        for j in rng.choice(nb, size=extra, replace=False):
            pattern[i].add(int(j))
            if sym:
                pattern[int(j)].add(i)
    blocks = {}
    for i in range(nb):
        for j in sorted(pattern[i]):
            if (i, j) in blocks:
                continue
            if i == j:
                R = rng.standard_normal((b, b))
                blocks[(i, i)] = R @ R.T + 4.0 * b * np.eye(b) if sym else R + 4.0 * b * np.eye(b)
            else:
                Bm = 0.3 * rng.standard_normal((b, b))
                blocks[(i, j)] = Bm
                if sym:
                    blocks[(j, i)] = Bm.T.copy()
    row = [1]
    col, v = [], []
    for i in range(nb):
        for j in sorted(pattern[i]):
            col.append(j + 1)
            v.append(blocks[(i, j)].T.ravel())  # column-major block
        row.append(len(col) + 1)
    return np.array(row, np.int32), np.array(col, np.int32), np.concatenate(v)


def dense(nb, b, row, col, v):
    A = np.zeros((nb * b, nb * b))
    for i in range(nb):
        for k in range(row[i] - 1, row[i + 1] - 1):
            j = col[k] - 1
            A[i * b:(i + 1) * b, j * b:(j + 1) * b] = v[k * b * b:(k + 1) * b * b].reshape(b, b).T
    return A


def block_mask(nb, b, pred):
    M = np.zeros((nb * b, nb * b), bool)
    for i in range(nb):
        for j in range(nb):
            if pred(i, j):
                M[i * b:(i + 1) * b, j * b:(j + 1) * b] = True
    return M


def greedy_colours(nb, row, col):
    c = np.zeros(nb, np.int32)
    for i in range(nb):
        used = {c[col[k] - 1] for k in range(row[i] - 1, row[i + 1] - 1) if col[k] - 1 != i}
        # symmetric closure: rows that list i
        k = 1
        while k in used:
            k += 1
        c[i] = k
    # fix conflicts from non-symmetric patterns by a second pass over both directions
    adj = [set() for _ in range(nb)]
    for i in range(nb):
        for k in range(row[i] - 1, row[i + 1] - 1):
            j = col[k] - 1
            if j != i:
                adj[i].add(j)
                adj[j].add(i)
    c[:] = 0
    for i in range(nb):
        used = {c[j] for j in adj[i]}
        k = 1
        while k in used:
            k += 1
        c[i] = k
    return c


@pytest.mark.parametrize("b", [3, 4, 5])
def test_lusgs_is_two_triangular_solves(b):
    nb = 40
    row, col, v = random_bcsr(nb, b, seed=b)
    A = dense(nb, b, row, col, v)
    rhs = np.random.default_rng(1).standard_normal(nb * b)
    Lo = np.where(block_mask(nb, b, lambda i, j: j < i), A, 0.0)
    D = np.where(block_mask(nb, b, lambda i, j: j == i), A, 0.0)
    y = np.linalg.solve(Lo + np.eye(nb * b), rhs)
    ref = np.linalg.solve(Lo + D, y)
    got = O.bcsr_lusgs(nb, b, row, col, v, rhs)
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("b", [3, 4])
def test_parallel_lusgs_is_coloured_symmetric_sweep(b):
    nb = 50
    row, col, v = random_bcsr(nb, b, seed=10 + b)
    A = dense(nb, b, row, col, v)
    c = greedy_colours(nb, row, col)
    assert c.max() >= 3
    rhs = np.random.default_rng(2).standard_normal(nb * b)
    Lc = np.where(block_mask(nb, b, lambda i, j: c[j] < c[i]), A, 0.0)
    Uc = np.where(block_mask(nb, b, lambda i, j: c[j] > c[i]), A, 0.0)
    D = np.where(block_mask(nb, b, lambda i, j: j == i), A, 0.0)
    y = np.linalg.solve(Lc + np.eye(nb * b), rhs)
    ref = np.linalg.solve(D + Uc, y)
    got = O.bcsr_parallel_lusgs(nb, b, row, col, v, rhs, c)
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("precond", ["none", "lusgs", "parallel_lusgs"])
def test_gmres_variants_reach_the_direct_solution(precond):
    nb, b = 60, 4
    row, col, v = random_bcsr(nb, b, seed=20)
    A = dense(nb, b, row, col, v)
    rhs = np.random.default_rng(3).standard_normal(nb * b)
    c = greedy_colours(nb, row, col)
    x, conv, info = O.bcsr_gmres(nb, b, row, col, v, rhs, tol=1e-10, restart=20, precond=precond, color=c)
    assert conv, info
    assert np.linalg.norm(rhs - A @ x) < 1e-10
    assert np.allclose(x, np.linalg.solve(A, rhs), rtol=1e-8, atol=1e-10)


def test_gmres_without_preconditioner_adds_the_basis_to_x():
    # restart length 1 with one restart limit: the plain variant must still
    # reduce the residual monotonically (minimal residual property)
    nb, b = 30, 3
    row, col, v = random_bcsr(nb, b, seed=21)
    A = dense(nb, b, row, col, v)
    rhs = np.random.default_rng(4).standard_normal(nb * b)
    _, _, info1 = O.bcsr_gmres(nb, b, row, col, v, rhs, tol=1e-30, restart=2, precond="none")
    assert info1["restarts"] == 20  # maxit of bcsr_gmres
    x, conv, info = O.bcsr_gmres(nb, b, row, col, v, rhs, tol=1e-9, restart=40, precond="none")
    assert conv and np.linalg.norm(rhs - A @ x) < 1e-9


def test_pcgm_on_spd_blocks_of_three():
    nb, b = 50, 3
    row, col, v = random_bcsr(nb, b, seed=30, sym=True, extra=2)
    A = dense(nb, b, row, col, v)
    assert np.allclose(A, A.T) and np.linalg.eigvalsh(A).min() > 0
    rhs = np.random.default_rng(5).standard_normal(nb * b)
    x, conv, info = O.bcsr_pcgm(nb, b, row, col, v, rhs, tol=1e-11)
    assert conv == 1, info
    assert np.allclose(x, np.linalg.solve(A, rhs), rtol=1e-8, atol=1e-10)
    assert 0 < info["iterations"] < nb * b


def test_pcgm_reports_a_singular_diagonal_block():
    nb, b = 10, 3
    row, col, v = random_bcsr(nb, b, seed=31, sym=True, extra=1)
    v = v.copy()
    k = [q for q in range(row[0] - 1, row[1] - 1) if col[q] == 1][0]
    v[k * 9:(k + 1) * 9] = np.diag([1.0, 1.0, 0.0]).ravel()  # lu_solve stops: |C(3,3)| < 1e-12
    _, conv, _ = O.bcsr_pcgm(nb, b, row, col, v, np.ones(nb * b), tol=1e-10)
    assert conv == -1


def test_add_block_adds_to_the_first_matching_block_only():
    nb, b = 8, 3
    row, col, v = random_bcsr(nb, b, seed=40)
    M = np.arange(9.0).reshape(3, 3)
    i = 3
    j = int(col[row[i - 1] - 1])  # first column of row i
    out = O.bcsr_add_block(nb, b, row, col, v, i, j, M)
    A0, A1 = dense(nb, b, row, col, v), dense(nb, b, row, col, out)
    D = A1 - A0
    assert np.allclose(D[(i - 1) * b:i * b, (j - 1) * b:j * b], M)
    D[(i - 1) * b:i * b, (j - 1) * b:j * b] = 0.0
    assert not D.any()
    # a block that is not stored is ignored
    missing = [jj for jj in range(1, nb + 1) if jj not in set(col[row[i - 1] - 1:row[i] - 1])][0]
    assert np.array_equal(O.bcsr_add_block(nb, b, row, col, v, i, missing, M), v)


def test_pcgm_first_iterate_is_the_block_jacobi_step():
    """One PCG step from x = 0 with the block-diagonal preconditioner:
    x1 = alpha M^-1 b, alpha = (M^-1 b . b) / (M^-1 b . A M^-1 b); a tolerance
    between the first and second preconditioned residuals stops there."""
    nb, b = 30, 3
    row, col, v = random_bcsr(nb, b, seed=32, sym=True, extra=1)
    v = v.copy()
    v *= 0.05  # weak coupling: one step already reduces the residual a lot
    for i in range(nb):
        k = [q for q in range(row[i] - 1, row[i + 1] - 1) if col[q] == i + 1][0]
        v[k * 9:(k + 1) * 9] = v[k * 9:(k + 1) * 9] / 0.05
    A = dense(nb, b, row, col, v)
    Mb = np.where(block_mask(nb, b, lambda i, j: i == j), A, 0.0)
    rhs = np.random.default_rng(6).standard_normal(nb * b)
    p0 = np.linalg.solve(Mb, rhs)
    rdr0 = p0 @ rhs
    alpha = rdr0 / (p0 @ (A @ p0))
    x1 = alpha * p0
    r1 = rhs - A @ x1
    rp1 = np.linalg.solve(Mb, r1) @ r1
    lo, hi = max(np.sqrt(rp1), np.linalg.norm(r1)), np.sqrt(rdr0)
    assert lo < hi
    x, conv, info = O.bcsr_pcgm(nb, b, row, col, v, rhs, tol=np.sqrt(lo * hi))
    assert conv == 1 and info["iterations"] == 1
    assert np.allclose(x, x1, rtol=1e-12, atol=1e-14)
