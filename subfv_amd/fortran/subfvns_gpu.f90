! subfvns on one MI355X: the driver of src/ns/ns_main.F90 over the device
! path.  It reads the reference's namelist file (&INPUT_PARAM,
! ns_global_data_module.F90:177-213; every entry is accepted, the ones the
! explicit path and its viscous terms use are honoured) and its msh 2.2 mesh,
! builds the mesh on
! the device, sets the initial state (init_uniform, init_1drp,
! init_isentropic_vortex, init_gresho, init_sedov or init_double_mach,
! ns_euler_module.F90:20-138), runs the time loop of ns_main.F90:83-107 with
! the state resident in HBM, prints the final time and the wall time of the
! loop (as ns_main.F90:113-126 does) and writes the final sol(5, n_elems) to
! sol_final.bin (stream: int32 n_elems, then the doubles).
!
! Usage: subfvns_gpu [namelist file, default input_data.f]
program subfvns_gpu
  use iso_c_binding
  use subfv_c
  use subfv_ns_gpu
  implicit none

  integer, parameter :: mnbc = 10
  ! ---- &INPUT_PARAM entries with the reference's names, types and defaults
  real(DP) :: Cv_p = 720.19471_DP, mu_p = 1.329e-5_DP, mu0 = 1.716e-5_DP, T0 = 273.15_DP
  logical :: use_sutherland = .false.
  real(DP) :: cfl = 1e5_DP, tmax = 1.0_DP
  character(len=255) :: meshfile_path = "", meshfile = ""
  logical :: rescale = .false., periodic_mesh = .false., use_cylinder_map = .false.
  real(DP) :: rescale_factor = 1.0_DP
  integer :: dim_cylinder_map = 2
  integer :: n_bc = 0
  character(len=255) :: bc_name(mnbc) = "", bc_type(mnbc) = "", bc_type_V(mnbc) = "", bc_type_T(mnbc) = ""
  real(DP) :: bc_val(5, mnbc) = 0.0_DP, bc_val_V(3, mnbc) = 0.0_DP, bc_val_T(mnbc) = 0.0_DP
  integer :: n_iter_print = 1, n_iter_write_sol = 100, method = 4, bc_style = 1
  logical :: second_order = .false., exclude_bound_vert = .false., activate_diffusion = .false.
  logical :: local_time_step = .true., boundary_2d = .false.
  character(len=255) :: scheme = ""
  logical :: write_residual = .true.
  integer :: n_iter_residual = 100, n_max_iter = 10000
  real(DP) :: final_res = 1e-4_DP
  logical :: init_uniform = .false., init_sedov = .false., init_isentropic_vortex = .false.
  logical :: init_gresho = .false., init_potential_flow_2d = .false., init_potential_flow_3d = .false.
  logical :: init_kelvin = .false., init_double_mach = .false., init_1drp = .false.
  real(DP) :: sol_uniform(5) = 0.0_DP, r_sedov = 0.0_DP, x1drp = 0.5_DP
  real(DP) :: sol_w_1drp_l(5) = 0.0_DP, sol_w_1drp_r(5) = 0.0_DP
  logical :: error_2d = .false., thermal_couette = .false.
  real(DP) :: error_2d_h = 0.025_DP
  logical :: init_restart = .false.
  character(len=255) :: restart_file = ""
  integer :: id_vtk_restart = 0
  logical :: compute_coeffs = .false.
  real(DP) :: pinf = 0.0_DP, rhoinf = 0.0_DP, vinf = 0.0_DP
  integer :: coeffs_surf = 0
  logical :: plot_solution_dat = .false.
  real(DP) :: xmin_dat = 0, xmax_dat = 0, ymin_dat = 0, ymax_dat = 0, zmin_dat = 0, zmax_dat = 0
  logical :: write_cell_size = .false.
  real(DP) :: delta_mach_imp = 1.0_DP

  namelist /INPUT_PARAM/ &
    meshfile_path, meshfile, rescale, rescale_factor, periodic_mesh, &
    use_cylinder_map, dim_cylinder_map, &
    n_bc, bc_name, bc_type, bc_val, bc_type_V, bc_val_V, bc_type_T, bc_val_T, &
    n_iter_print, n_iter_write_sol, second_order, method, &
    scheme, exclude_bound_vert, activate_diffusion, local_time_step, &
    bc_style, boundary_2d, write_residual, n_iter_residual, &
    final_res, n_max_iter, init_uniform, sol_uniform, init_sedov, r_sedov, &
    init_isentropic_vortex, init_gresho, init_potential_flow_2d, init_potential_flow_3d, &
    init_kelvin, init_double_mach, init_1drp, sol_w_1drp_l, sol_w_1drp_r, x1drp, &
    error_2d, error_2d_h, init_restart, restart_file, id_vtk_restart, &
    compute_coeffs, pinf, rhoinf, vinf, coeffs_surf, plot_solution_dat, &
    xmin_dat, xmax_dat, ymin_dat, ymax_dat, zmin_dat, zmax_dat, &
    write_cell_size, delta_mach_imp, mu_p, Cv_p, use_sutherland, mu0, T0, &
    cfl, tmax, thermal_couette

  real(DP), parameter :: gamma = 1.4_DP
  character(len=512) :: nlfile
  integer :: funit, i, n_elems, n_vert
  type(c_ptr) :: ctx
  type(raw_mesh) :: mesh
  type(subfv_config) :: cfg
  real(DP), allocatable :: sol(:, :), xc(:, :), vol(:)
  real(DP) :: t, dt, w(5), wall_time, rmin, r
  integer :: imin
  integer(8) :: start_count, end_count, count_rate
  integer(c_int) :: st
  integer(c_int32_t) :: bc_v_id(mnbc) = 1, bc_t_id(mnbc) = 1

  nlfile = "input_data.f"
  if (command_argument_count() >= 1) call get_command_argument(1, nlfile)
  open (newunit=funit, file=trim(nlfile), status="old")
  read (funit, nml=INPUT_PARAM)
  close (funit)

  ! ---- init_flags (ns_global_data_module.F90:220-330), the explicit-path subset
  select case (trim(adjustl(scheme)))
  case ("multi_point"); cfg%scheme_id = 1
  case ("multi_point_iso"); cfg%scheme_id = 2
  case ("three_wave"); cfg%scheme_id = 3
  case ("two_wave"); cfg%scheme_id = 4
  case ("modified_three_wave"); cfg%scheme_id = 5
  case ("multi_point_pressure"); cfg%scheme_id = 6
  case ("multi_point_pressure_ph"); cfg%scheme_id = 8
  case default
    print *, "SCHEME NOT RECOGNIZED !"
    error stop
  end select
  if (n_bc > mnbc) error stop "n_bc > mnbc"
  cfg%n_bc = n_bc
  do i = 1, n_bc
    select case (trim(adjustl(bc_type(i))))
    case ("wall", ""); cfg%bc_euler_id(i) = 1
    case ("adherence_wall"); cfg%bc_euler_id(i) = 2
    case ("freestream"); cfg%bc_euler_id(i) = 3
    case ("outflowsupersonic"); cfg%bc_euler_id(i) = 4
    case ("inflow_pond"); cfg%bc_euler_id(i) = 5
    case ("inout_double_mach"); cfg%bc_euler_id(i) = 6
    case ("double_mach_bottom"); cfg%bc_euler_id(i) = 7
    case ("potential_flow_2d"); cfg%bc_euler_id(i) = 8
    case ("potential_flow_3d"); cfg%bc_euler_id(i) = 9
    case default
      print *, "BC TYPE NOT RECOGNIZED !"
      error stop
    end select
    cfg%bc_val(:, i) = bc_val(:, i)
  end do
  cfg%bc_style = bc_style
  cfg%boundary_2d = merge(1, 0, boundary_2d)
  cfg%exclude_bound_vert = merge(1, 0, exclude_bound_vert)
  cfg%local_time_step = merge(1, 0, local_time_step)
  cfg%second_order = merge(1, 0, second_order)
  cfg%method = method
  cfg%cfl = cfl
  cfg%tmax = tmax
  if (rescale .or. use_cylinder_map) error stop "rescaled/mapped meshes are not supported"

  ! ---- mesh and initial state
  call check(c_null_ptr, subfv_create(ctx, 0))
  call read_mesh_msh_2(mesh, meshfile_path, meshfile, n_bc, bc_name)
  call build_mesh(ctx, mesh, boundary_2d, periodic_mesh)
  call check(ctx, subfv_set_config(ctx, cfg))
  if (activate_diffusion) then
    ! viscous terms with constant viscosity; Prandtl is 0.71 in the reference
    ! (ns_global_data_module.F90:10); bc_type_V / bc_type_T per BC index
    do i = 1, n_bc
      bc_v_id(i) = merge(2, 1, trim(adjustl(bc_type_V(i))) == "dirichlet")
      bc_t_id(i) = merge(2, 1, trim(adjustl(bc_type_T(i))) == "dirichlet")
    end do
    call check(ctx, subfv_set_sutherland(ctx, merge(1, 0, use_sutherland), mu0, T0))  ! compute_mu
    call check(ctx, subfv_set_diffusion(ctx, 1, mu_p, Cv_p, 0.71_DP, int(n_bc, c_int32_t), bc_v_id, bc_val_V, &
      bc_t_id, bc_val_T))
  end if
  n_elems = mesh_count(ctx, "n_elems")
  n_vert = mesh_count(ctx, "n_vert")
  allocate (sol(5, n_elems), xc(3, n_elems), vol(n_elems))
  call elem_geometry(ctx, n_elems, xc, vol)
  if (init_1drp .and. .not. init_uniform) then
    ! init_sol's init_1drp branch on the device (ns_euler_module.F90:36-43)
    call check(ctx, subfv_init_1drp(ctx, x1drp, sol_w_1drp_l, sol_w_1drp_r))
  else if (init_sedov .and. .not. (init_uniform .or. init_isentropic_vortex .or. init_gresho)) then
    ! init_sedov (ns_euler_module.F90:64-111, one rank): rho = 1 at rest with
    ! E = 1e-16 / (gamma - 1), and the blast energy in the first cell of
    ! smallest centroid norm (the x-y norm with boundary_2d)
    rmin = huge(1.0_DP)
    imin = -1
    do i = 1, n_elems
      if (boundary_2d) then
        r = sqrt(xc(1, i)**2 + xc(2, i)**2)
      else
        r = sqrt(xc(1, i)**2 + xc(2, i)**2 + xc(3, i)**2)
      end if
      if (r < rmin) then
        rmin = r
        imin = i
      end if
    end do
    sol(1, :) = 1.0_DP
    sol(2:4, :) = 0.0_DP
    sol(5, :) = 1e-16_DP / (gamma - 1.0_DP)
    if (boundary_2d) then
      sol(5, imin) = 0.984042_DP / vol(imin)
    else
      sol(5, imin) = 0.851072_DP / vol(imin)
    end if
    call check(ctx, subfv_state_put(ctx, sol))
  else
    ! the other init_sol branches in the reference's order (ns_euler_module.F90:32-138)
    do i = 1, n_elems
      if (init_uniform) then
        w = sol_uniform
      else if (init_isentropic_vortex) then
        w = isentropic_vortex(xc(:, i))
      else if (init_gresho) then
        w = gresho_state(xc(:, i), 1.0e-5_DP)
      else if (init_double_mach) then
        ! the shock of the double Mach reflection through (0.1667, 0) at 60
        ! degrees; tmax is set to 0.2 as init_sol does (:127-138)
        if (xc(2, i) < 1.732_DP * (xc(1, i) - 0.1667_DP)) then
          w = (/1.4_DP, 0.0_DP, 0.0_DP, 0.0_DP, 1.0_DP/)
        else
          w = (/8.0_DP, 7.145_DP, -4.125_DP, 0.0_DP, 116.5_DP/)
        end if
      else
        error stop "supported initial states: init_uniform, init_1drp, init_isentropic_vortex, init_gresho, " // &
          "init_sedov, init_double_mach"
      end if
      sol(1, i) = w(1)
      sol(2:4, i) = w(1) * w(2:4)
      sol(5, i) = w(5) / (gamma - 1.0_DP) + 0.5_DP * w(1) * (w(2)**2 + w(3)**2 + w(4)**2)
    end do
    if (init_double_mach .and. .not. (init_uniform .or. init_isentropic_vortex .or. init_gresho)) then
      tmax = 0.2_DP
      cfg%tmax = tmax
      print *, "T_max =", tmax
      call check(ctx, subfv_set_config(ctx, cfg))
    end if
    call check(ctx, subfv_state_put(ctx, sol))
  end if
  call check(ctx, subfv_set_time(ctx, 0.0_DP))
  if (final_res > 0.0_DP) print *, "note: the device loop runs n_max_iter+1 steps or to tmax (final_res unused)"

  ! ---- time loop (do while iter <= n_max_iter .and. t < tmax)
  call system_clock(count_rate=count_rate)
  call system_clock(start_count)
  st = subfv_run(ctx, int(n_max_iter + 1, c_int32_t), 0, t, dt)
  call check(ctx, st)
  call check(ctx, subfv_state_get(ctx, sol))
  call system_clock(end_count)

  wall_time = real(end_count - start_count, DP) / real(count_rate, DP)
  print *, "  --------------------------------------------------------"
  print *, "  Final time                  =", t
  print *, "  Total CPU time (wall)       =", wall_time
  print *, "  --------------------------------------------------------"
  open (newunit=funit, file="sol_final.bin", access="stream", form="unformatted", status="replace")
  write (funit) int(n_elems, c_int32_t)
  write (funit) sol
  close (funit)
  ! the final write_sol of ns_main.F90 (iaff = -1): write_sol_dat
  ! (ns_io_module.F90:50-96), the cells strictly inside the *_dat box
  if (plot_solution_dat) then
    open (newunit=funit, file="output_-1.dat", status="replace")
    write (funit, *) ""
    do i = 1, n_elems
      if ((xc(1, i) < xmax_dat .and. xc(1, i) > xmin_dat) .and. &
          (xc(2, i) < ymax_dat .and. xc(2, i) > ymin_dat) .and. &
          (xc(3, i) < zmax_dat .and. xc(3, i) > zmin_dat)) then
        w(1) = sol(1, i)
        w(2:4) = sol(2:4, i) / sol(1, i)
        w(5) = (gamma - 1) * (sol(5, i) - 0.5_DP * w(1) * (w(2)**2 + w(3)**2 + w(4)**2))
        write (funit, *) xc(:, i), sol(:, i), w(:)
      end if
    end do
    close (funit)
  end if
  call subfv_destroy(ctx)

contains

  ! sol_gresho_mach (ns_euler_exact_sol_module.F90:7-44): the Gresho vortex
  ! of radius 0.4 around the origin, background pressure 1/(gamma Mach^2)
  function gresho_state(x, mach) result(w)
    real(DP), intent(in) :: x(3), mach
    real(DP) :: w(5), p0, r
    p0 = 1.0_DP / (gamma * mach**2)
    r = norm2(x(1:2))
    w(1) = 1.0_DP
    w(4) = 0.0_DP
    if (r < 0.2_DP) then
      w(2) = -5.0_DP * x(2)
      w(3) = 5.0_DP * x(1)
      w(5) = p0 + 12.5_DP * r**2
    else if (r > 0.2_DP .and. r < 0.4_DP) then
      w(2) = (5.0_DP - 2.0_DP / r) * x(2)
      w(3) = (-5.0_DP + 2.0_DP / r) * x(1)
      w(5) = p0 + 12.5_DP * r**2 + 4.0_DP - 20.0_DP * r + 4.0_DP * log(5.0_DP * r)
    else
      w(2) = 0.0_DP
      w(3) = 0.0_DP
      w(5) = p0 - 2.0_DP + 4.0_DP * log(2.0_DP)
    end if
  end function

  ! sol_isentropic_vortex at t = 0 (ns_euler_exact_sol_module.F90:47-74):
  ! strength 5 at the origin, no advection
  function isentropic_vortex(x) result(w)
    real(DP), intent(in) :: x(3)
    real(DP) :: w(5), r
    real(DP), parameter :: beta = 5.0_DP, pi = 4.0_DP * atan(1.0_DP)
    r = x(1)**2 + x(2)**2
    w(1) = (1.0_DP - ((gamma - 1) * beta**2) / (8.0_DP * gamma * pi**2) * exp(1.0_DP - r))**(1.0_DP / (gamma - 1.0_DP))
    w(2) = -x(2) * beta / (2.0_DP * pi) * exp(0.5_DP * (1.0_DP - r))
    w(3) = x(1) * beta / (2.0_DP * pi) * exp(0.5_DP * (1.0_DP - r))
    w(4) = 0.0_DP
    w(5) = w(1)**gamma
  end function
end program subfvns_gpu
