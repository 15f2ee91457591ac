! Drives one second-order residual and one stage update of compute_rhs_ns
! (src/ns/ns_main.F90:130-179) and the BCSR routines through the
! reference-signature procedures of subfv_reference_api, as the reference's
! own driver calls them, and writes what they return (stream files) for
! comparison with the oracle.
!
! Usage: reference_api_check mesh.msh bcsr.bin
!   mesh.msh  msh 2.2 mesh, all walls; init_1drp state (x < 0.5: rho 1, p 1;
!             else rho 0.125, p 0.1), multi_point, bc_style 2, boundary_2d,
!             second order with method 4, cfl 0.4
!   bcsr.bin  int32 m, b, nnz, row(m + 1), col(nnz), then doubles v(b, b, nnz),
!             x(b m), tol
! Writes residual.bin (rhs after flux_limiting, rhs_o2, grad, nodal grad,
! sol after add_timedisc_and_solve, dt) and bcsr_out.bin (bcsr_matmul(A, x),
! bcsr_diag(A, ., x), bcsr_gmres_diag(A, ., x), conv).
program reference_api_check
  use iso_c_binding
  use subfv_c
  use subfv_ns_gpu, only: raw_mesh, read_mesh_msh_2, build_mesh, elem_geometry, check, mesh_count
  use subfv_reference_api
  implicit none

  integer, parameter :: DP = c_double
  character(len=512) :: mfile, bfile
  type(c_ptr) :: ctx
  type(raw_mesh) :: raw
  type(mesh_type) :: mesh
  type(subfv_config) :: cfg
  type(bcsr_matrix_type) :: A
  real(DP), allocatable :: sol(:, :), rhs(:, :), rhs_o2(:, :), grad(:, :, :), ngrad(:, :, :), vp(:, :)
  real(DP), allocatable :: sl(:), sl2(:), diag(:, :, :), dsol(:, :), xc(:, :), vol(:), h_p(:), mat_h_p(:, :, :)
  real(DP), allocatable :: x(:), y(:), xd(:), xg(:), H(:, :), V(:, :), r(:), yy(:), cc(:), ss(:), z(:), zp(:)
  real(DP) :: res, res_vect(5), dt, tol
  integer :: i, u, mm
  integer(c_int32_t) :: m, b, nnz
  logical :: conv
  character(len=255) :: names(1)

  call get_command_argument(1, mfile)
  call get_command_argument(2, bfile)
  call check(c_null_ptr, subfv_create(ctx, 0))
  names = ""
  call read_mesh_msh_2(raw, "", mfile, 0, names)
  call build_mesh(ctx, raw, .true.)
  mesh%n_elems = mesh_count(ctx, "n_elems")
  mesh%n_vert = mesh_count(ctx, "n_vert")
  cfg%scheme_id = 1
  cfg%bc_style = 2
  cfg%boundary_2d = 1
  cfg%second_order = 1
  cfg%method = 4
  cfg%cfl = 0.4_DP
  cfg%tmax = 1e30_DP
  call check(ctx, subfv_set_config(ctx, cfg))
  call subfv_ref_use_context(ctx)

  associate (ne => mesh%n_elems, nv => mesh%n_vert)
    allocate (sol(5, ne), rhs(5, ne), rhs_o2(5, ne), grad(3, 5, ne), ngrad(3, 5, nv), vp(3, nv))
    allocate (sl(ne), sl2(ne), diag(5, 5, ne), dsol(5, ne), xc(3, ne), vol(ne), h_p(nv), mat_h_p(3, 3, nv))
    call elem_geometry(ctx, ne, xc, vol)
    do i = 1, ne
      if (xc(1, i) < 0.5_DP) then
        sol(:, i) = (/1.0_DP, 0.0_DP, 0.0_DP, 0.0_DP, 1.0_DP / 0.4_DP/)
      else
        sol(:, i) = (/0.125_DP, 0.0_DP, 0.0_DP, 0.0_DP, 0.1_DP / 0.4_DP/)
      end if
    end do
    grad = 0.0_DP
    h_p = 0.0_DP
    mat_h_p = 0.0_DP
    ! compute_rhs_ns: first-order fluxes, gradients, second-order fluxes, limiting
    call compute_euler_flux(mesh, sol, rhs, grad, sl, .false., vp, h_p, mat_h_p)
    call compute_cell_grad_from_nodal_grad(mesh, sol, grad, ngrad, 4)
    call compute_euler_flux(mesh, sol, rhs_o2, grad, sl2, .true., vp, h_p, mat_h_p)
    call flux_limiting(mesh, sol, rhs, rhs_o2, sl, 0.4_DP)
    open (newunit=u, file="residual.bin", access="stream", form="unformatted", status="replace")
    write (u) int(ne, c_int32_t), int(nv, c_int32_t)
    write (u) rhs, rhs_o2, grad, ngrad
    ! the forward-Euler stage update (a1 = a2 = 1, t = 0)
    diag = 0.0_DP
    dt = 0.0_DP
    call add_timedisc_and_solve(mesh, sol, rhs, diag, sl, dsol, res, res_vect, 1, 0, 1.0_DP, 1.0_DP, &
      0.4_DP, 0.0_DP, dt)
    write (u) sol, dt
    close (u)
  end associate

  ! ---- BCSR
  open (newunit=u, file=trim(bfile), access="stream", form="unformatted", status="old")
  read (u) m, b, nnz
  A%m = m
  A%n = m
  A%nnz = nnz
  A%blc_size = b
  allocate (A%row(m + 1), A%col(nnz), A%v(b, b, nnz), x(b * m))
  block
    integer(c_int32_t) :: row32(m + 1), col32(nnz)
    read (u) row32, col32
    A%row = row32
    A%col = col32
  end block
  read (u) A%v, x, tol
  close (u)
  mm = 30
  allocate (y(b * m), xd(b * m), xg(b * m), H(mm + 1, mm), V(b * m, mm + 1), r(b * m), yy(mm + 1), cc(mm + 1), &
    ss(mm + 1), z(b * m), zp(b * m))
  y = bcsr_matmul(A, x)
  xd = 0.0_DP
  call bcsr_diag(A, xd, x)
  ! values refilled in place (doubled) between solves: the next call must see them
  A%v = 2.0_DP * A%v
  xg = 0.0_DP
  conv = .false.
  call bcsr_gmres_diag(A, xg, x, tol, mm, conv, H, V, r, yy, cc, ss, z, zp)
  open (newunit=u, file="bcsr_out.bin", access="stream", form="unformatted", status="replace")
  write (u) y, xd, xg, merge(1_c_int32_t, 0_c_int32_t, conv)
  close (u)
  call subfv_ref_release()
  call subfv_destroy(ctx)
end program reference_api_check
