! iso_c_binding interface to libsubfv_hip.so (include/subfv.h).
!
! Every C function is declared with bind(C) and value/reference arguments
! matching the header; arrays are passed as assumed-size arguments, so a
! Fortran caller hands over its own column-major arrays (sol(5, n_elems),
! grad(3, 5, n_elems), vp(3, n_vert), BCSR v(b, b, nnzb)) unchanged.
module subfv_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SUBFV_OK = 0, SUBFV_E_INVALID = 1, SUBFV_E_DEVICE = 2, &
    SUBFV_E_BAD_OMEGA = 3, SUBFV_E_NEG_STAR = 4, SUBFV_E_MESH = 5, SUBFV_E_BAD_BC = 6, &
    SUBFV_E_NOT_CONVERGED = 7
  integer, parameter :: SUBFV_MNBC = 10

  type, bind(C) :: subfv_config
    integer(c_int32_t) :: scheme_id = 1
    integer(c_int32_t) :: bc_style = 1
    integer(c_int32_t) :: boundary_2d = 0
    integer(c_int32_t) :: exclude_bound_vert = 0
    integer(c_int32_t) :: local_time_step = 0
    integer(c_int32_t) :: second_order = 0
    integer(c_int32_t) :: method = 4
    integer(c_int32_t) :: n_bc = 0
    integer(c_int32_t) :: bc_euler_id(SUBFV_MNBC) = 1
    real(c_double) :: bc_val(5, SUBFV_MNBC) = 0.0_c_double
    real(c_double) :: cfl = 0.4_c_double
    real(c_double) :: tmax = 1.0_c_double
  end type subfv_config

  interface
    integer(c_int) function subfv_create(ctx, device) bind(C, name="subfv_create")
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function
    subroutine subfv_destroy(ctx) bind(C, name="subfv_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function subfv_error(ctx) bind(C, name="subfv_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function subfv_mesh_build(ctx, n_vert, coord, n_elems, kind, elem_vert, n_bfaces, &
        bf_nvert, bf_vert, bf_tag, boundary_2d) bind(C, name="subfv_mesh_build")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n_vert, n_elems, n_bfaces, boundary_2d
      real(c_double), intent(in) :: coord(*)
      integer(c_int32_t), intent(in) :: kind(*), elem_vert(*), bf_nvert(*), bf_vert(*), bf_tag(*)
    end function
    integer(c_int64_t) function subfv_mesh_count(ctx, what) bind(C, name="subfv_mesh_count")
      import :: c_ptr, c_int64_t, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: what(*)
    end function
    integer(c_int64_t) function subfv_mesh_get_i32(ctx, name, out, cap) bind(C, name="subfv_mesh_get_i32")
      import :: c_ptr, c_int64_t, c_int32_t, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int32_t), intent(out) :: out(*)
      integer(c_int64_t), value :: cap
    end function
    integer(c_int64_t) function subfv_mesh_get_f64(ctx, name, out, cap) bind(C, name="subfv_mesh_get_f64")
      import :: c_ptr, c_int64_t, c_double, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int64_t), value :: cap
    end function
    integer(c_int) function subfv_host_mesh_build(hm, n_vert, coord, n_elems, kind, elem_vert, n_bfaces, &
        bf_nvert, bf_vert, bf_tag) bind(C, name="subfv_host_mesh_build")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), intent(out) :: hm
      integer(c_int32_t), value :: n_vert, n_elems, n_bfaces
      real(c_double), intent(in) :: coord(*)
      integer(c_int32_t), intent(in) :: kind(*), elem_vert(*), bf_nvert(*), bf_vert(*), bf_tag(*)
    end function
    integer(c_int64_t) function subfv_host_mesh_get_i32(hm, name, out, cap) &
        bind(C, name="subfv_host_mesh_get_i32")
      import :: c_ptr, c_int64_t, c_int32_t, c_char
      type(c_ptr), value :: hm
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int32_t), intent(out) :: out(*)
      integer(c_int64_t), value :: cap
    end function
    subroutine subfv_host_mesh_free(hm) bind(C, name="subfv_host_mesh_free")
      import :: c_ptr
      type(c_ptr), value :: hm
    end subroutine
    integer(c_int) function subfv_init_1drp(ctx, x1drp, w_left, w_right) bind(C, name="subfv_init_1drp")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: x1drp
      real(c_double), intent(in) :: w_left(5), w_right(5)
    end function
    integer(c_int) function subfv_host_mesh_make_periodic(hm) bind(C, name="subfv_host_mesh_make_periodic")
      import :: c_ptr, c_int
      type(c_ptr), value :: hm
    end function
    integer(c_int) function subfv_mesh_upload(ctx, hm, boundary_2d) bind(C, name="subfv_mesh_upload")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, hm
      integer(c_int32_t), value :: boundary_2d
    end function
    integer(c_int) function subfv_set_config(ctx, cfg) bind(C, name="subfv_set_config")
      import :: c_ptr, c_int, subfv_config
      type(c_ptr), value :: ctx
      type(subfv_config), intent(in) :: cfg
    end function
    ! the node sweep's arithmetic: 0 exact (bit-identical), 1 fast (within 1e-12)
    integer(c_int) function subfv_set_arith(ctx, mode) bind(C, name="subfv_set_arith")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode
    end function
    integer(c_int) function subfv_set_time(ctx, t) bind(C, name="subfv_set_time")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t
    end function
    integer(c_int) function subfv_state_put(ctx, sol) bind(C, name="subfv_state_put")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: sol(*)
    end function
    integer(c_int) function subfv_state_get(ctx, sol) bind(C, name="subfv_state_get")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sol(*)
    end function
    integer(c_int64_t) function subfv_get_f64(ctx, name, out, cap) bind(C, name="subfv_get_f64")
      import :: c_ptr, c_int64_t, c_double, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int64_t), value :: cap
    end function
    integer(c_int64_t) function subfv_put_f64(ctx, name, in, n) bind(C, name="subfv_put_f64")
      import :: c_ptr, c_int64_t, c_double, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(in) :: in(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function subfv_compute_euler_flux(ctx, second_order, want_vp) &
        bind(C, name="subfv_compute_euler_flux")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: second_order, want_vp
    end function
    integer(c_int) function subfv_compute_cell_grad(ctx, method) bind(C, name="subfv_compute_cell_grad")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: method
    end function
    integer(c_int) function subfv_flux_limiting(ctx, cfl) bind(C, name="subfv_flux_limiting")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cfl
    end function
    integer(c_int) function subfv_add_timedisc_and_solve(ctx, a1, a2, cfl, t, dt, omega) &
        bind(C, name="subfv_add_timedisc_and_solve")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: a1, a2, cfl, t
      real(c_double), intent(out) :: dt, omega
    end function
    integer(c_int) function subfv_compute_rhs_ns(ctx) bind(C, name="subfv_compute_rhs_ns")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function subfv_set_sutherland(ctx, activate, mu0, t0) bind(C, name="subfv_set_sutherland")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: activate
      real(c_double), value :: mu0, t0
    end function
    integer(c_int) function subfv_compute_mu(ctx, sol, mu_arr, mu0corr) bind(C, name="subfv_compute_mu")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: sol(*)
      real(c_double), intent(out) :: mu_arr(*)
      real(c_double), intent(out) :: mu0corr
    end function
    integer(c_int) function subfv_add_diffusion(ctx, sol, mu_arr, mu0corr, parts, rhs, diag) &
        bind(C, name="subfv_add_diffusion")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: sol(*), mu_arr(*)
      real(c_double), value :: mu0corr
      integer(c_int32_t), value :: parts
      real(c_double), intent(inout) :: rhs(*), diag(*)
    end function
    integer(c_int) function subfv_set_diffusion(ctx, activate, mu_p, cv_p, prandtl, n_bc, bc_v_id, bc_val_v, &
        bc_t_id, bc_val_t) bind(C, name="subfv_set_diffusion")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: activate, n_bc
      real(c_double), value :: mu_p, cv_p, prandtl
      integer(c_int32_t), intent(in) :: bc_v_id(*), bc_t_id(*)
      real(c_double), intent(in) :: bc_val_v(*), bc_val_t(*)
    end function
    integer(c_int) function subfv_run(ctx, n_steps, time_scheme, t, dt) bind(C, name="subfv_run")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n_steps, time_scheme
      real(c_double), intent(out) :: t, dt
    end function
    integer(c_int) function subfv_profile_enable(ctx, on) bind(C, name="subfv_profile_enable")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function subfv_profile_read(ctx, phase, ms_total, count) bind(C, name="subfv_profile_read")
      import :: c_ptr, c_int, c_char, c_double, c_int64_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: phase(*)
      real(c_double), intent(out) :: ms_total
      integer(c_int64_t), intent(out) :: count
    end function
    integer(c_int) function subfv_compute_rhs_lagrange(ctx, sol, mass, b2d, rhs, vp) &
        bind(C, name="subfv_compute_rhs_lagrange")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: sol(*), mass(*)
      integer(c_int32_t), value :: b2d
      real(c_double), intent(out) :: rhs(*), vp(*)
    end function
    integer(c_int) function subfv_acoustic_run(ctx, sol, n_steps, cfl, n_bc, bc_kind, bc_val, t, dt_last) &
        bind(C, name="subfv_acoustic_run")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: sol(*)
      integer(c_int32_t), value :: n_steps, n_bc
      real(c_double), value :: cfl
      integer(c_int32_t), intent(in) :: bc_kind(*)
      real(c_double), intent(in) :: bc_val(*)
      real(c_double), intent(inout) :: t
      real(c_double), intent(out) :: dt_last
    end function
    integer(c_int) function subfv_bcsr_create(ctx, A, m, b, nnzb, row, col, v) bind(C, name="subfv_bcsr_create")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: A
      integer(c_int32_t), value :: m, b, nnzb
      integer(c_int32_t), intent(in) :: row(*), col(*)
      real(c_double), intent(in) :: v(*)
    end function
    subroutine subfv_bcsr_destroy(A) bind(C, name="subfv_bcsr_destroy")
      import :: c_ptr
      type(c_ptr), value :: A
    end subroutine
    integer(c_int) function subfv_bcsr_set_values(A, v) bind(C, name="subfv_bcsr_set_values")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: A
      real(c_double), intent(in) :: v(*)
    end function
    integer(c_int) function subfv_bcsr_matmul(A, x, y) bind(C, name="subfv_bcsr_matmul")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: A
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function
    integer(c_int) function subfv_bcsr_diag(A, x, rhs) bind(C, name="subfv_bcsr_diag")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: A
      real(c_double), intent(out) :: x(*)
      real(c_double), intent(in) :: rhs(*)
    end function
    integer(c_int) function subfv_bcsr_gmres_diag(A, x, rhs, tol, restart, precision, conv, stats) &
        bind(C, name="subfv_bcsr_gmres_diag")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: A
      real(c_double), intent(out) :: x(*)
      real(c_double), intent(in) :: rhs(*)
      real(c_double), value :: tol
      integer(c_int32_t), value :: restart, precision
      integer(c_int32_t), intent(out) :: conv
      real(c_double), intent(out) :: stats(3)
    end function
    integer(c_int) function subfv_bcsr_gmres(A, x, rhs, tol, restart, precond, precision, conv, stats) &
        bind(C, name="subfv_bcsr_gmres")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: A
      real(c_double), intent(out) :: x(*)
      real(c_double), intent(in) :: rhs(*)
      real(c_double), value :: tol
      integer(c_int32_t), value :: restart, precond, precision
      integer(c_int32_t), intent(out) :: conv
      real(c_double), intent(out) :: stats(3)
    end function
    integer(c_int) function subfv_bcsr_lusgs(A, x, rhs) bind(C, name="subfv_bcsr_lusgs")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: A
      real(c_double), intent(out) :: x(*)
      real(c_double), intent(in) :: rhs(*)
    end function
    integer(c_int) function subfv_bcsr_set_colors(A, color, n_sets) bind(C, name="subfv_bcsr_set_colors")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: A
      integer(c_int32_t), intent(in) :: color(*)
      integer(c_int32_t), intent(out) :: n_sets
    end function
    integer(c_int) function subfv_bcsr_parallel_lusgs(A, x, rhs) bind(C, name="subfv_bcsr_parallel_lusgs")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: A
      real(c_double), intent(out) :: x(*)
      real(c_double), intent(in) :: rhs(*)
    end function
    integer(c_int) function subfv_bcsr_pcgm(A, x, rhs, tol, stats) bind(C, name="subfv_bcsr_pcgm")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: A
      real(c_double), intent(inout) :: x(*)
      real(c_double), intent(in) :: rhs(*)
      real(c_double), value :: tol
      real(c_double), intent(out) :: stats(2)
    end function
    integer(c_int) function subfv_bcsr_add_block(A, i, j, M) bind(C, name="subfv_bcsr_add_block")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: A
      integer(c_int32_t), value :: i, j
      real(c_double), intent(in) :: M(*)
    end function
  end interface

contains

  ! Fortran string -> NUL-terminated C string
  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s)//c_null_char
  end function

  ! Message of the last error of a context
  function subfv_message(ctx) result(msg)
    type(c_ptr), value :: ctx
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: i
    msg = ""
    cp = subfv_error(ctx)
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function
end module subfv_c
