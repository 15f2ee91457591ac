! The reference's module procedures of the explicit path and of its BCSR
! solver, with the reference's own names and argument lists, running on the
! device context this module holds:
!   compute_euler_flux(mesh, sol, rhs, grad, sum_lambda, second_order, vp, h_p, mat_h_p)
!       (src/ns/ns_euler_module.F90:151-174)
!   compute_cell_grad_from_nodal_grad(mesh, sol, grad, all_nodal_grad, method)
!       (src/ns/ns_euler_recon_module.F90:8-15)
!   flux_limiting(mesh, sol, rhs, rhs_o2, sum_lambda, cfl)
!       (src/ns/ns_euler_recon_module.F90:560-569)
!   add_timedisc_and_solve(mesh, sol, rhs, diag, sum_lambda, delta_sol, res,
!       res_vect, num_procs, me, a1, a2, cfl, t, dt)  (src/ns/ns_euler_module.F90:521-540)
!   bcsr_matrix_type, y = bcsr_matmul(A, x), bcsr_diag(A, x, b),
!   bcsr_gmres_diag(A, x, b, tol, m, conv, H, V, r, y, c, s, z, zp)
!       (src/linear_algebra/sparse_bcsr_linear_module.F90:6-29,1414,1592)
!
! The mesh argument is read for its counts only (mesh%n_elems, mesh%n_vert):
! the device context given to subfv_ref_use_context holds the mesh itself
! (built from the same msh file with build_mesh of subfv_ns_gpu, or
! subfv_mesh_build).  The scheme and its options come from the context's
! configuration (subfv_set_config), as the reference takes them from
! ns_global_data_module.  A failure prints the reference's message and
! stops, as the reference does.
module subfv_reference_api
  use iso_c_binding
  use subfv_c
  use subfv_ns_gpu, only: check, get_named, put_named
  use mesh_module, only: mesh_type
  implicit none
  private
  public :: mesh_type, bcsr_matrix_type, subfv_ref_use_context, subfv_ref_release
  public :: compute_euler_flux, compute_cell_grad_from_nodal_grad, flux_limiting, add_timedisc_and_solve
  public :: bcsr_matmul, bcsr_diag, bcsr_gmres_diag

  integer, parameter :: DP = c_double

  ! sparse_bcsr_linear_module.F90:6-11
  type :: bcsr_matrix_type
    integer :: m, n, nnz
    integer :: blc_size
    integer, dimension(:), allocatable :: col, row
    real(DP), dimension(:, :, :), allocatable :: v
  end type bcsr_matrix_type

  type(c_ptr), save :: ctx = c_null_ptr
  ! the device copy of the last matrix: its pattern is kept while the row and
  ! col arrays and the sizes stay the same; the values are sent on every call,
  ! since the reference refills v in place between solves
  type(c_ptr), save :: dmat = c_null_ptr
  integer, save :: dm_m = -1, dm_b = -1, dm_nnz = -1
  integer(c_intptr_t), save :: dm_row = 0, dm_col = 0

contains

  subroutine subfv_ref_use_context(c)
    type(c_ptr), intent(in) :: c
    ctx = c
  end subroutine

  ! drops the device copy of the last matrix
  subroutine subfv_ref_release()
    if (c_associated(dmat)) call subfv_bcsr_destroy(dmat)
    dmat = c_null_ptr
    dm_m = -1
  end subroutine

  subroutine need_mesh(mesh)
    type(mesh_type), intent(in) :: mesh
    if (.not. c_associated(ctx)) error stop "subfv_reference_api: no device context (subfv_ref_use_context)"
    if (subfv_mesh_count(ctx, "n_elems" // c_null_char) /= mesh%n_elems .or. &
        subfv_mesh_count(ctx, "n_vert" // c_null_char) /= mesh%n_vert) &
      error stop "subfv_reference_api: the device context holds another mesh"
  end subroutine

  ! rhs and sum_lambda (from 1e-12) of the first- or second-order fluxes;
  ! vp receives the nodal velocities (3, n_vert) when it has room for them.
  ! h_p and mat_h_p belong to schemes off the explicit path and are unused.
  subroutine compute_euler_flux(mesh, sol, rhs, grad, sum_lambda, second_order, vp, h_p, mat_h_p)
    type(mesh_type), intent(in) :: mesh
    real(DP), dimension(5, mesh%n_elems), intent(in) :: sol
    real(DP), dimension(:, :, :), intent(in) :: grad
    real(DP), dimension(5, mesh%n_elems), intent(inout) :: rhs
    real(DP), dimension(:, :), intent(inout) :: vp
    real(DP), dimension(mesh%n_elems), intent(inout) :: sum_lambda
    real(DP), dimension(mesh%n_vert), intent(in) :: h_p
    real(DP), dimension(3, 3, mesh%n_vert), intent(in) :: mat_h_p
    logical, intent(in) :: second_order
    integer(c_int32_t) :: want_vp
    call need_mesh(mesh)
    want_vp = merge(1, 0, size(vp, 1) == 3 .and. size(vp, 2) >= mesh%n_vert)
    call check(ctx, subfv_state_put(ctx, sol))
    if (second_order) then
      call put_named(ctx, "cell_grad", grad, 15 * mesh%n_elems)
      call check(ctx, subfv_compute_euler_flux(ctx, 1, want_vp))
      call get_named(ctx, "rhs_o2", rhs, 5 * mesh%n_elems)
      call get_named(ctx, "sum_lambda_o2", sum_lambda, mesh%n_elems)
    else
      call check(ctx, subfv_compute_euler_flux(ctx, 0, want_vp))
      call get_named(ctx, "rhs", rhs, 5 * mesh%n_elems)
      call get_named(ctx, "sum_lambda", sum_lambda, mesh%n_elems)
    end if
    if (want_vp == 1) call get_named(ctx, "vp", vp, 3 * mesh%n_vert)
  end subroutine

  subroutine compute_cell_grad_from_nodal_grad(mesh, sol, grad, all_nodal_grad, method)
    type(mesh_type), intent(in) :: mesh
    real(DP), dimension(5, mesh%n_elems), intent(in) :: sol
    real(DP), dimension(3, 5, mesh%n_elems), intent(inout) :: grad
    real(DP), dimension(3, 5, mesh%n_vert), intent(inout) :: all_nodal_grad
    integer, intent(in) :: method
    call need_mesh(mesh)
    call check(ctx, subfv_state_put(ctx, sol))
    call check(ctx, subfv_compute_cell_grad(ctx, int(method, c_int32_t)))
    call get_named(ctx, "cell_grad", grad, 15 * mesh%n_elems)
    call get_named(ctx, "nodal_grad", all_nodal_grad, 15 * mesh%n_vert)
  end subroutine

  subroutine flux_limiting(mesh, sol, rhs, rhs_o2, sum_lambda, cfl)
    type(mesh_type), intent(in) :: mesh
    real(DP), dimension(5, mesh%n_elems), intent(in) :: sol
    real(DP), dimension(5, mesh%n_elems), intent(inout) :: rhs
    real(DP), dimension(5, mesh%n_elems), intent(in) :: rhs_o2
    real(DP), dimension(mesh%n_elems), intent(in) :: sum_lambda
    real(DP), intent(in) :: cfl
    call need_mesh(mesh)
    call check(ctx, subfv_state_put(ctx, sol))
    call put_named(ctx, "rhs", rhs, 5 * mesh%n_elems)
    call put_named(ctx, "rhs_o2", rhs_o2, 5 * mesh%n_elems)
    call put_named(ctx, "sum_lambda", sum_lambda, mesh%n_elems)
    call check(ctx, subfv_flux_limiting(ctx, cfl))
    call get_named(ctx, "rhs", rhs, 5 * mesh%n_elems)
  end subroutine

  ! The stage update: sol advanced on the device; on return rhs, diag and
  ! sum_lambda hold what the reference leaves in them (rhs + c sol, diag + c I
  ! with c = a1 (vol / dt + sum_lambda), sum_lambda zeroed without local time
  ! stepping), res_vect the norms of the incoming rhs, delta_sol and res the
  ! increment and the sum of its norms (recovered from the new state, to
  ! rounding), dt the reference's dt.
  subroutine add_timedisc_and_solve(mesh, sol, rhs, diag, sum_lambda, delta_sol, res, res_vect, num_procs, me, &
      a1, a2, cfl, t, dt)
    type(mesh_type), intent(in) :: mesh
    real(DP), dimension(5, mesh%n_elems), intent(inout) :: sol, rhs
    real(DP), dimension(5, 5, mesh%n_elems), intent(inout) :: diag
    real(DP), dimension(mesh%n_elems), intent(inout) :: sum_lambda
    real(DP), dimension(5, mesh%n_elems), intent(inout) :: delta_sol
    real(DP), intent(inout) :: res
    real(DP), dimension(5), intent(inout) :: res_vect
    integer, intent(in) :: num_procs, me
    real(DP), intent(in) :: a1, a2, cfl, t
    real(DP), intent(inout) :: dt
    real(DP) :: omega, dts(1), lt(1), cc, vol(mesh%n_elems)
    integer :: i, k
    call need_mesh(mesh)
    res_vect = 0.0_DP
    do i = 1, mesh%n_elems
      do k = 1, 5
        res_vect(k) = res_vect(k) + rhs(k, i)**2
      end do
    end do
    res_vect = sqrt(res_vect)
    call check(ctx, subfv_state_put(ctx, sol))
    call put_named(ctx, "rhs", rhs, 5 * mesh%n_elems)
    call put_named(ctx, "sum_lambda", sum_lambda, mesh%n_elems)
    if (any(diag /= 0.0_DP)) call put_named(ctx, "diag", diag, 25 * mesh%n_elems)
    call check(ctx, subfv_add_timedisc_and_solve(ctx, a1, a2, cfl, t, dt, omega))
    if (subfv_get_f64(ctx, "dt_solve" // c_null_char, dts, 1_c_int64_t) /= 1) error stop "dt_solve"
    if (subfv_mesh_get_f64(ctx, "elem_volume" // c_null_char, vol, int(mesh%n_elems, c_int64_t)) /= mesh%n_elems) &
      error stop "elem_volume"
    if (subfv_get_f64(ctx, "local_time_step" // c_null_char, lt, 1_c_int64_t) /= 1) error stop "local_time_step"
    ! without local time stepping the reference zeroes sum_lambda before the solve
    if (lt(1) == 0.0_DP) sum_lambda = 0.0_DP
    delta_sol = sol
    call check(ctx, subfv_state_get(ctx, sol))
    res = 0.0_DP
    do i = 1, mesh%n_elems
      cc = a1 * (vol(i) / dts(1) + sum_lambda(i))
      do k = 1, 5
        diag(k, k, i) = diag(k, k, i) + cc
        rhs(k, i) = rhs(k, i) + cc * delta_sol(k, i)
      end do
      delta_sol(:, i) = (sol(:, i) - delta_sol(:, i)) / (a2 * omega)
      res = res + norm2(delta_sol(:, i))
    end do
    if (omega < 1.0_DP .and. me == 0 .and. num_procs >= 1) print *, "OMEGA", omega
  end subroutine

  ! the device matrix for A: pattern uploaded when A's arrays or sizes
  ! change, values every call
  type(c_ptr) function device_matrix(A) result(h)
    type(bcsr_matrix_type), intent(in), target :: A
    integer(c_intptr_t) :: pr, pc
    integer(c_int32_t), allocatable :: row32(:), col32(:)
    if (.not. c_associated(ctx)) error stop "subfv_reference_api: no device context (subfv_ref_use_context)"
    pr = transfer(c_loc(A%row), pr)
    pc = transfer(c_loc(A%col), pc)
    if (.not. (c_associated(dmat) .and. dm_m == A%m .and. dm_b == A%blc_size .and. dm_nnz == A%nnz .and. &
        dm_row == pr .and. dm_col == pc)) then
      call subfv_ref_release()
      row32 = int(A%row(1:A%m + 1), c_int32_t)
      col32 = int(A%col(1:A%nnz), c_int32_t)
      call check(ctx, subfv_bcsr_create(ctx, dmat, int(A%m, c_int32_t), int(A%blc_size, c_int32_t), &
        int(A%nnz, c_int32_t), row32, col32, A%v))
      dm_m = A%m
      dm_b = A%blc_size
      dm_nnz = A%nnz
      dm_row = pr
      dm_col = pc
    else
      call check(ctx, subfv_bcsr_set_values(dmat, A%v))
    end if
    h = dmat
  end function

  function bcsr_matmul(A, x) result(y)
    type(bcsr_matrix_type), intent(in), target :: A
    real(DP), dimension(A%blc_size * A%n), intent(in) :: x
    real(DP), dimension(A%blc_size * A%m) :: y
    call check(ctx, subfv_bcsr_matmul(device_matrix(A), x, y))
  end function

  subroutine bcsr_diag(A, x, b)
    type(bcsr_matrix_type), intent(in), target :: A
    real(DP), dimension(A%blc_size * A%m), intent(inout) :: x
    real(DP), dimension(A%blc_size * A%m), intent(in) :: b
    call check(ctx, subfv_bcsr_diag(device_matrix(A), x, b))
  end subroutine

  ! GMRES(m) from x = 0 with absolute tolerance tol, block-Jacobi on the
  ! right, up to 40 restarts; the Krylov work arrays H, V, r, y, c, s, z, zp
  ! of the reference's interface live on the device and are not touched.
  subroutine bcsr_gmres_diag(A, x, b, tol, m, conv, H, V, r, y, c, s, z, zp)
    integer, intent(in) :: m
    real(DP), intent(in) :: tol
    type(bcsr_matrix_type), intent(in), target :: A
    real(DP), dimension(A%blc_size * A%m), intent(in) :: b
    real(DP), dimension(A%blc_size * A%m), intent(inout) :: x
    logical, intent(inout) :: conv
    real(DP), dimension(m + 1, m) :: H
    real(DP), dimension(A%blc_size * A%m, m + 1) :: V
    real(DP), dimension(A%blc_size * A%m) :: r
    real(DP), dimension(m + 1) :: y, c, s
    real(DP), dimension(A%blc_size * A%m) :: z, zp
    integer(c_int32_t) :: cv
    integer(c_int) :: st
    real(DP) :: stats(3)
    st = subfv_bcsr_gmres_diag(device_matrix(A), x, b, tol, int(m, c_int32_t), 0_c_int32_t, cv, stats)
    if (st /= SUBFV_OK .and. st /= SUBFV_E_NOT_CONVERGED) call check(ctx, st)
    conv = cv /= 0
  end subroutine
end module subfv_reference_api
