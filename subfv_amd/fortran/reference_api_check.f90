! Drives the mesh set-up (read_mesh_msh_2, build_mesh, compute_geometry_mesh),
! one second-order residual and one stage update of compute_rhs_ns
! (src/ns/ns_main.F90:130-179), the viscous builders (compute_mu,
! build_momentum_diff_mat_into_global_mat_diag,
! build_diff_mat_into_global_mat_diag) and the BCSR routines through the
! reference-signature procedures of subfv_reference_api, as the reference's
! own driver calls them, and writes what they return (stream files) for
! comparison with the oracle.
!
! Usage: reference_api_check mesh.msh bcsr.bin [periodic]
!   mesh.msh  msh 2.2 mesh, all walls; init_1drp state (x < 0.5: rho 1, p 1;
!             else rho 0.125, p 0.1), multi_point, bc_style 2, boundary_2d,
!             second order with method 4, cfl 0.4
!   bcsr.bin  int32 m, b, nnz, row(m + 1), col(nnz), then doubles v(b, b, nnz),
!             x(b m), tol
! Writes mesh.bin (the counts, then the vertices, elements and boundary
! faces of the built mesh; with "periodic" the mesh is built with
! periodic_mesh = .true. and mesh.bin is the only output), residual.bin (rhs after flux_limiting, rhs_o2,
! grad, nodal grad, sol after add_timedisc_and_solve, dt), viscous.bin (a
! smooth state, cell viscosities, compute_mu's result, the two builders'
! sums into rhs and diag, a warm state and compute_mu's Sutherland
! viscosities for it) and bcsr_out.bin (bcsr_matmul(A, x),
! bcsr_diag(A, ., x), bcsr_gmres_diag(A, ., x), conv).
program reference_api_check
  use iso_c_binding
  use subfv_c
  use subfv_ns_gpu, only: check
  use subfv_reference_api
  implicit none

  integer, parameter :: DP = c_double
  character(len=512) :: mfile, bfile, mode
  type(c_ptr) :: ctx
  type(mesh_type) :: mesh
  type(subfv_config) :: cfg
  type(bcsr_matrix_type) :: A
  real(DP), allocatable :: sol(:, :), rhs(:, :), rhs_o2(:, :), grad(:, :, :), ngrad(:, :, :), vp(:, :)
  real(DP), allocatable :: sl(:), sl2(:), diag(:, :, :), dsol(:, :), xc(:, :), vol(:), h_p(:), mat_h_p(:, :, :)
  real(DP), allocatable :: x(:), y(:), xd(:), xg(:), H(:, :), V(:, :), r(:), yy(:), cc(:), ss(:), z(:), zp(:)
  real(DP) :: res, res_vect(5), dt, tol
  integer :: i, j, u, mm, msr
  integer(c_int32_t) :: m, b, nnz
  logical :: conv
  character(len=255) :: names(1)

  call get_command_argument(1, mfile)
  call get_command_argument(2, bfile)
  call get_command_argument(3, mode)
  names = ""
  call read_mesh_msh_2(mesh, "", mfile, 0, names)
  msr = 0
  call build_mesh(mesh, 1, msr, .true., .true., periodic_mesh=(trim(mode) == "periodic"))
  call compute_geometry_mesh(mesh, .true., .true.)
  ctx = subfv_ref_context()
  if (.not. c_associated(ctx)) error stop "no device context after build_mesh"
  open (newunit=u, file="mesh.bin", access="stream", form="unformatted", status="replace")
  write (u) int((/mesh%n_vert, mesh%n_elems, mesh%n_faces, mesh%n_sub_faces, mesh%n_sub_elems, &
    mesh%n_max_sub_faces_around_node, mesh%n_boundary_faces/), c_int32_t)
  write (u) (mesh%vert(i)%coord, i = 1, mesh%n_vert), (mesh%vert(i)%volume, i = 1, mesh%n_vert)
  write (u) (merge(1_c_int32_t, 0_c_int32_t, mesh%vert(i)%is_bound), i = 1, mesh%n_vert)
  write (u) (int(mesh%elem(i)%elem_kind, c_int32_t), i = 1, mesh%n_elems), &
    (int(mesh%elem(i)%tag, c_int32_t), i = 1, mesh%n_elems), &
    ((int(mesh%elem(i)%vert(j), c_int32_t), j = 1, mesh%elem(i)%n_vert), i = 1, mesh%n_elems)
  write (u) (mesh%elem(i)%coord, i = 1, mesh%n_elems), (mesh%elem(i)%volume, i = 1, mesh%n_elems)
  write (u) ((int(mesh%boundary_face(i)%vert(j), c_int32_t), j = 1, mesh%boundary_face(i)%n_vert), &
    i = 1, mesh%n_boundary_faces)
  close (u)
  if (trim(mode) == "periodic") then
    call subfv_ref_release()
    stop
  end if
  cfg%scheme_id = 1
  cfg%bc_style = 2
  cfg%boundary_2d = 1
  cfg%second_order = 1
  cfg%method = 4
  cfg%cfl = 0.4_DP
  cfg%tmax = 1e30_DP
  call check(ctx, subfv_set_config(ctx, cfg))

  associate (ne => mesh%n_elems, nv => mesh%n_vert)
    allocate (sol(5, ne), rhs(5, ne), rhs_o2(5, ne), grad(3, 5, ne), ngrad(3, 5, nv), vp(3, nv))
    allocate (sl(ne), sl2(ne), diag(5, 5, ne), dsol(5, ne), xc(3, ne), vol(ne), h_p(nv), mat_h_p(3, 3, nv))
    do i = 1, ne
      xc(:, i) = mesh%elem(i)%coord
    end do
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
    call viscous(ne)
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

contains

  ! compute_rhs_ns's viscous block (ns_main.F90: compute_mu, then the
  ! momentum and heat builders into the same rhs and diag) on a smooth state
  ! and varying cell viscosities, mu_p = 2e-3, Cv 720.19471, Pr 0.71, walls;
  ! then compute_mu with Sutherland's law (mu0 1.716e-5, T0 273.15)
  subroutine viscous(ne)
    integer, intent(in) :: ne
    real(DP) :: sv(5, ne), sv1(5, ne), mu(ne), mus(ne), m0, m0s, r(5, ne), d(5, 5, ne), rm(5, ne), dm(5, 5, ne)
    real(DP) :: x, y, rho, uu, vv, p, mmu(ne), m0c
    integer(c_int32_t) :: ib(1), ii(1)
    real(DP) :: vb(3), tb(1)
    integer :: k, uo
    do k = 1, ne
      x = mesh%elem(k)%coord(1)
      y = mesh%elem(k)%coord(2)
      rho = 1.0_DP + 0.2_DP * sin(3.0_DP * x + 1.0_DP) * cos(2.0_DP * y)
      uu = 0.3_DP * sin(2.0_DP * y + 0.5_DP)
      vv = -0.2_DP * cos(3.0_DP * x)
      p = 1.0_DP + 0.3_DP * cos(2.0_DP * x + y)
      sv(:, k) = (/rho, rho * uu, rho * vv, 0.0_DP, p / 0.4_DP + 0.5_DP * rho * (uu * uu + vv * vv)/)
      mu(k) = 2e-3_DP * (1.0_DP + 0.5_DP * sin(5.0_DP * x) * cos(3.0_DP * y))
    end do
    ib = 1
    ii = 1
    vb = 0.0_DP
    tb = 0.0_DP
    call check(ctx, subfv_set_diffusion(ctx, 1, 2e-3_DP, 720.19471_DP, 0.71_DP, 0, ib, vb, ii, tb))
    call compute_mu(mesh, sv, 1, mmu, m0c)
    ! the builders, one after the other into the same arrays, from 1 and I
    m0 = minval(mu)
    r = 1.0_DP
    d = 0.0_DP
    do k = 1, 5
      d(k, k, :) = 1.0_DP
    end do
    call build_momentum_diff_mat_into_global_mat_diag(mesh, d, r, sv, mu, m0)
    rm = r
    dm = d
    call build_diff_mat_into_global_mat_diag(mesh, d, r, sv, mu)
    sv1 = sv
    ! warm air for Sutherland's law: T = 300 + 40 x, u as above
    do k = 1, ne
      x = mesh%elem(k)%coord(1)
      rho = sv(1, k)
      uu = 30.0_DP * sv(2, k) / rho
      vv = 20.0_DP * sv(3, k) / rho
      sv(5, k) = rho * 720.19471_DP * (300.0_DP + 40.0_DP * x) + 0.5_DP * rho * (uu * uu + vv * vv)
      sv(2, k) = rho * uu
      sv(3, k) = rho * vv
    end do
    call check(ctx, subfv_set_sutherland(ctx, 1, 1.716e-5_DP, 273.15_DP))
    call compute_mu(mesh, sv, 1, mus, m0s)
    call check(ctx, subfv_set_sutherland(ctx, 0, 1.716e-5_DP, 273.15_DP))
    call check(ctx, subfv_set_diffusion(ctx, 0, 2e-3_DP, 720.19471_DP, 0.71_DP, 0, ib, vb, ii, tb))
    open (newunit=uo, file="viscous.bin", access="stream", form="unformatted", status="replace")
    write (uo) mmu, m0c, mu, sv1, rm, dm, r, d, sv, mus, m0s
    close (uo)
  end subroutine
end program reference_api_check
