"""The oracle against the reference's own Fortran routines, compiled from
the reference tree by `make -C oracle ref` (only possible where the
reference sources are present; skipped elsewhere).

- lu_solve_noswp, conserv_to_primit, primit_to_conserv, sound_speed_w,
  temp_u and max_omega_for_positivity are plain Fortran: the oracle must
  match them bit for bit;
- lu_solve_inplace_lapack, pseudo_inverse_inplace_lapack and inv_lapack
  go through LAPACK (here the OpenBLAS bundled with scipy; the reference
  leaves the implementation unpinned): the oracle's restatements must agree
  to rounding (relative 1e-13, scaled by the condition number for the
  pseudo-inverse and the inverse).
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libsubfv_ref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference routines not built (make -C oracle ref)")

dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


@pytest.fixture(scope="module")
def ref():
    L = C.CDLL(REF)
    L.ref_lu_solve_noswp.argtypes = [C.c_int, dp, dp, dp]
    L.ref_lu_solve_lapack.argtypes = [C.c_int, dp, dp, dp]
    L.ref_pinv.argtypes = [C.c_int, dp]
    L.ref_inv.argtypes = [C.c_int, dp]
    L.ref_inverse_3_by_3.argtypes = [dp, dp]
    L.ref_primitives.argtypes = [dp, dp, dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.ref_max_omega.argtypes = [dp, dp]
    L.ref_max_omega.restype = C.c_double
    L.ref_norm2.argtypes = [C.c_int, dp]
    L.ref_norm2.restype = C.c_double
    return L


def olib():
    L = O.lib()
    L.om_test_lu_solve_noswp.argtypes = [C.c_int, dp, dp, dp]
    L.om_test_lu_solve.argtypes = [C.c_int, dp, dp, dp]
    L.om_test_pinv3.argtypes = [dp]
    L.om_test_primitives.argtypes = [dp, dp, dp, C.POINTER(C.c_double)]
    L.om_test_temp_u.argtypes = [dp, C.c_double]
    L.om_test_temp_u.restype = C.c_double
    L.om_max_omega_for_positivity.argtypes = [dp, dp]
    L.om_max_omega_for_positivity.restype = C.c_double
    return L


def colmajor(A):
    return np.ascontiguousarray(np.asarray(A, np.float64).T).ravel()


@pytest.mark.parametrize("n", [3, 4, 5])
def test_lu_solve_noswp_bit_exact(ref, n):
    L = olib()
    rng = np.random.default_rng(n)
    for _ in range(50):
        A = rng.standard_normal((n, n)) + n * np.eye(n)
        b = rng.standard_normal(n)
        x1, x2 = np.zeros(n), np.zeros(n)
        ref.ref_lu_solve_noswp(n, colmajor(A), x1, b)
        L.om_test_lu_solve_noswp(n, colmajor(A), b, x2)
        assert np.array_equal(x1, x2)


def test_primitives_bit_exact(ref):
    L = olib()
    rng = np.random.default_rng(3)
    for _ in range(200):
        w = np.array([0.1 + rng.random(), rng.standard_normal(), rng.standard_normal(), rng.standard_normal(),
                      0.1 + rng.random()])
        u = np.array([w[0], w[0] * w[1], w[0] * w[2], w[0] * w[3], w[4] / 0.4 + 0.5 * w[0] * (w[1:4] ** 2).sum()])
        rw, ru, ra, rt = np.zeros(5), np.zeros(5), C.c_double(), C.c_double()
        ref.ref_primitives(u, rw, ru, C.byref(ra), C.byref(rt))
        ow, ou, oa = np.zeros(5), np.zeros(5), C.c_double()
        L.om_test_primitives(u, ow, ou, C.byref(oa))
        assert np.array_equal(rw, ow) and np.array_equal(ru, ou) and ra.value == oa.value
        assert rt.value == L.om_test_temp_u(u, 720.19471)


def test_max_omega_bit_exact(ref):
    L = olib()
    rng = np.random.default_rng(11)
    for _ in range(200):
        u = np.array([1.0 + rng.random(), rng.standard_normal(), rng.standard_normal(), 0.0, 3.0 + rng.random()])
        du = rng.standard_normal(5) * rng.choice([0.1, 1.0, 5.0])
        assert ref.ref_max_omega(u, du) == L.om_max_omega_for_positivity(u, du)


@pytest.mark.parametrize("n", [3, 5])
def test_lapack_lu_solve_agrees(ref, n):
    L = olib()
    rng = np.random.default_rng(20 + n)
    exact = 0
    for _ in range(100):
        A = rng.standard_normal((n, n))
        b = rng.standard_normal(n)
        x1, x2 = np.zeros(n), np.zeros(n)
        ref.ref_lu_solve_lapack(n, colmajor(A), x1, b)
        L.om_test_lu_solve(n, colmajor(A), b, x2)
        scale = np.linalg.cond(A) * np.abs(x1).max()
        assert np.abs(x1 - x2).max() <= 1e-13 * scale
        exact += np.array_equal(x1, x2)
    print(f"bit-identical LU solves: {exact}/100")


def test_pseudo_inverse_agrees(ref):
    L = olib()
    rng = np.random.default_rng(5)
    for k in range(100):
        A = rng.standard_normal((3, 3))
        if k % 3 == 1:  # rank 2, as on one-layer meshes
            A[:, 2] = 0.0
            A[2, :] = 0.0
        a_ref, a_om = colmajor(A), colmajor(A)
        ref.ref_pinv(3, a_ref)
        L.om_test_pinv3(a_om)
        s = np.linalg.svd(A, compute_uv=False)
        cond = s[0] / s[s > 3e-12 * s[0]][-1]
        assert np.abs(a_ref - a_om).max() <= 1e-13 * cond * np.abs(a_ref).max()


def test_inverse_agrees(ref):
    rng = np.random.default_rng(9)
    for n in (3, 12, 36):
        A = rng.standard_normal((n, n)) + 2 * np.eye(n)
        a_ref = colmajor(A)
        ref.ref_inv(n, a_ref)
        Xr = a_ref.reshape(n, n).T
        Xo = O.inv(A)
        assert np.abs(Xr - Xo).max() <= 1e-13 * np.linalg.cond(A) * np.abs(Xr).max()


def test_inverse_3_by_3_bit_exact(ref):
    rng = np.random.default_rng(12)
    for _ in range(200):
        A = rng.standard_normal((3, 3)) * 10.0 ** rng.uniform(-3, 3)
        out = np.zeros(9)
        ref.ref_inverse_3_by_3(colmajor(A), out)
        assert np.array_equal(out.reshape(3, 3).T, O.inverse_3_by_3(A))


def test_norm2_agrees_with_the_compiler_intrinsic_to_two_ulp(ref):
    """The oracle evaluates the intrinsic norm2 of short vectors as
    sqrt(x1^2 + x2^2 + ...).  Compilers evaluate NORM2 with scaled sums
    (this image's flang as well as gfortran), so the reference's own value
    depends on the compiler it is built with; the two agree to within two
    units in the last place (about 60% of the cases bit for bit, most of the
    rest one unit apart)."""
    rng = np.random.default_rng(13)
    equal = total = 0
    for n in (2, 3, 5):
        for _ in range(300):
            x = rng.standard_normal(n) * 10.0 ** rng.uniform(-6, 6)
            got = ref.ref_norm2(n, np.ascontiguousarray(x))
            s = 0.0
            for v in x:
                s = s + v * v
            mine = np.sqrt(s)
            assert abs(got - mine) <= 2 * np.spacing(mine)
            equal += got == mine
            total += 1
    print(f"norm2 bit-identical in {equal}/{total} cases")
