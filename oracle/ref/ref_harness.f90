! TEST INFRASTRUCTURE ONLY -- C-callable entry points into the reference's
! own Fortran modules (compiled from the sources under the reference tree by
! oracle/Makefile's `ref` target, LAPACK/BLAS from the OpenBLAS bundled with
! scipy).  Used by tests/test_oracle_vs_reference.py to pin the oracle's
! dense solves and primitive conversions against the reference code itself:
!   lu_solve_noswp, lu_solve_inplace_lapack, pseudo_inverse_inplace_lapack,
!   inv_lapack, inverse_3_by_3  (src/linear_algebra/linear_solver_module.F90)
!   primit_to_conserv, conserv_to_primit, sound_speed_w, temp_u,
!   max_omega_for_positivity  (src/ns/ns_euler_primitives_module.F90)
module ref_harness
  use iso_c_binding
  use precision_module
  use linear_solver_module
  use ns_euler_primitives_module
  implicit none
contains
  subroutine ref_lu_solve_noswp(n, A, x, b) bind(C, name="ref_lu_solve_noswp")
    integer(c_int), value :: n
    real(c_double), intent(in) :: A(n, n), b(n)
    real(c_double), intent(out) :: x(n)
    real(kind=DOUBLE) :: Cw(n, n)
    x = 0.0_DOUBLE
    call lu_solve_noswp(n, A, x, b, Cw)
  end subroutine

  subroutine ref_lu_solve_lapack(n, A, x, b) bind(C, name="ref_lu_solve_lapack")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: A(n, n)
    real(c_double), intent(in) :: b(n)
    real(c_double), intent(out) :: x(n)
    call lu_solve_inplace_lapack(n, A, x, b)
  end subroutine

  subroutine ref_pinv(n, A) bind(C, name="ref_pinv")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: A(n, n)
    call pseudo_inverse_inplace_lapack(n, A)
  end subroutine

  subroutine ref_inv(n, A) bind(C, name="ref_inv")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: A(n, n)
    call inv_lapack(n, A)
  end subroutine

  subroutine ref_inverse_3_by_3(A, Ainv) bind(C, name="ref_inverse_3_by_3")
    real(c_double), intent(in) :: A(3, 3)
    real(c_double), intent(out) :: Ainv(3, 3)
    Ainv = 0.0_DOUBLE
    call inverse_3_by_3(A, Ainv)
  end subroutine

  subroutine ref_primitives(u, w, u_back, a, theta) bind(C, name="ref_primitives")
    real(c_double), intent(in) :: u(5)
    real(c_double), intent(out) :: w(5), u_back(5), a, theta
    w = conserv_to_primit(u)
    u_back = primit_to_conserv(w)
    a = sound_speed_w(w)
    theta = temp_u(u)
  end subroutine

  real(c_double) function ref_max_omega(u, du) bind(C, name="ref_max_omega")
    real(c_double), intent(in) :: u(5), du(5)
    ref_max_omega = max_omega_for_positivity(u, du)
  end function
  ! the intrinsic norm2 as the compiler of the reference evaluates it
  real(c_double) function ref_norm2(n, x) bind(C, name="ref_norm2")
    integer(c_int), value :: n
    real(c_double), intent(in) :: x(n)
    ref_norm2 = norm2(x)
  end function
end module ref_harness
