! Fortran surface of the explicit residual path on the device.
!
! The procedures keep the names, argument order and argument meaning of the
! reference's module procedures, with the mesh replaced by the device
! context that holds it:
!   read_mesh_msh_2 / build_mesh / compute_geometry_mesh
!       (src/mesh/mesh_reading_module.F90:989-1324,
!        src/mesh/mesh_connectivity_module.F90:12-59,
!        src/mesh/mesh_geometry_module.F90:16-57)
!   compute_euler_flux                 (src/ns/ns_euler_module.F90:151-174)
!   compute_cell_grad_from_nodal_grad  (src/ns/ns_euler_recon_module.F90:8-15)
!   flux_limiting                      (src/ns/ns_euler_recon_module.F90:560-569)
!   add_timedisc_and_solve             (src/ns/ns_euler_module.F90:521-540)
!   bcsr_matmul / bcsr_diag / bcsr_gmres_diag
!       (src/linear_algebra/sparse_bcsr_linear_module.F90:13,1414,1592)
! Host arrays go in and come out in the reference's layouts.  A failure
! prints the reference's message and stops, as the reference does.
module subfv_ns_gpu
  use iso_c_binding
  use subfv_c
  implicit none

  integer, parameter :: DP = c_double

  type :: raw_mesh
    integer(c_int32_t) :: n_vert = 0, n_elems = 0, n_bfaces = 0
    real(c_double), allocatable :: coord(:, :)
    integer(c_int32_t), allocatable :: kind(:), elem_vert(:)
    integer(c_int32_t), allocatable :: bf_nvert(:), bf_vert(:), bf_tag(:)
  end type raw_mesh

contains

  ! vertices of a gmsh element kind (gmsh format 2.2 element types 1-31:
  ! lines, triangles, quadrangles, tetrahedra, hexahedra, prisms, pyramids and
  ! their higher-order forms, points)
  pure integer function verts_of_kind(kind) result(n)
    integer, intent(in) :: kind
    select case (kind)
    case (15); n = 1
    case (1); n = 2
    case (2, 8); n = 3
    case (3, 4, 26); n = 4
    case (7, 27); n = 5
    case (6, 9, 28); n = 6
    case (5, 16); n = 8
    case (10, 20); n = 9
    case (11, 21); n = 10
    case (22); n = 12
    case (19); n = 13
    case (14); n = 14
    case (18, 23, 24); n = 15
    case (13); n = 18
    case (17, 29); n = 20
    case (25); n = 21
    case (12); n = 27
    case (30); n = 35
    case (31); n = 56
    case default; n = 0
    end select
  end function

  subroutine check(ctx, st)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: st
    if (st /= SUBFV_OK) then
      print *, trim(subfv_message(ctx))
      error stop
    end if
  end subroutine

  ! msh 2.2 reader (gmsh ASCII format 2.2: sections "$Name" ... "$EndName";
  ! $PhysicalNames lines "dim tag name", $Nodes lines "id x y z", $Elements
  ! lines "id kind ntags tag(1:ntags) vertex(1:n)"), with the reference's
  ! semantics (mesh_reading_module.F90:989-1324): nodes in file order (their
  ! ids are not used), volume elements of kinds 4-7 in file order, boundary
  ! faces of kinds 2 and 3 tagged with the bc index whose name carries the
  ! face's first (physical) tag, the last such index when several match, 0
  ! (default wall) when none does.  With n_bc > 0 the file must name its
  ! physical groups before its nodes.
  subroutine read_mesh_msh_2(mesh, meshfile_path, meshfile, n_bc, bc_name)
    type(raw_mesh), intent(out) :: mesh
    character(len=*), intent(in) :: meshfile_path, meshfile
    integer, intent(in) :: n_bc
    character(len=*), dimension(:), intent(in) :: bc_name
    character(len=512) :: line
    character(len=255) :: gname
    integer :: u, ios, count, k, m, gdim, gtag, head(3), nv, ntag
    integer :: ne, nb, lv, lb
    integer :: fld(96)
    integer, allocatable :: phys_of_bc(:)
    logical :: have_names

    allocate (phys_of_bc(max(n_bc, 1)))
    phys_of_bc = -huge(1)
    have_names = .false.
    ne = 0; nb = 0; lv = 0; lb = 0
    open (newunit=u, file=trim(adjustl(meshfile_path))//trim(adjustl(meshfile)), status="old", action="read")
    do
      read (u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      select case (trim(adjustl(line)))
      case ("$PhysicalNames")
        have_names = .true.
        read (u, *) count
        do k = 1, count
          read (u, *) gdim, gtag, gname
          do m = 1, n_bc
            if (trim(adjustl(bc_name(m))) == trim(adjustl(gname))) phys_of_bc(m) = gtag
          end do
        end do
      case ("$Nodes")
        if (n_bc /= 0 .and. .not. have_names) then
          print *, "No Physical names inside mesh!!!"
          error stop
        end if
        read (u, *) mesh%n_vert
        allocate (mesh%coord(3, mesh%n_vert))
        do k = 1, mesh%n_vert
          read (u, *) gtag, mesh%coord(:, k)
        end do
      case ("$Elements")
        read (u, *) count
        ! first pass over the section: sizes only
        allocate (mesh%kind(count), mesh%elem_vert(8 * count))
        allocate (mesh%bf_nvert(count), mesh%bf_vert(4 * count), mesh%bf_tag(count))
        do k = 1, count
          read (u, '(a)') line
          read (line, *) head
          ntag = head(3)
          nv = verts_of_kind(head(2))
          if (3 + ntag + nv > size(fld)) error stop "msh element line too long"
          read (line, *) fld(1:3 + ntag + nv)
          select case (head(2))
          case (4:7)
            ne = ne + 1
            mesh%kind(ne) = head(2)
            mesh%elem_vert(lv + 1:lv + nv) = fld(4 + ntag:3 + ntag + nv)
            lv = lv + nv
          case (2:3)
            nb = nb + 1
            mesh%bf_nvert(nb) = nv
            mesh%bf_vert(lb + 1:lb + nv) = fld(4 + ntag:3 + ntag + nv)
            lb = lb + nv
            mesh%bf_tag(nb) = 0
            if (ntag >= 1) then
              do m = 1, n_bc
                if (phys_of_bc(m) == fld(4)) mesh%bf_tag(nb) = m
              end do
            end if
          end select
        end do
      end select
    end do
    close (u)
    if (ne < 1) then
      print *, "[-] No 3D elements in mesh !"
      error stop
    end if
    mesh%n_elems = ne
    mesh%n_bfaces = nb
    mesh%kind = mesh%kind(1:ne)
    mesh%elem_vert = mesh%elem_vert(1:lv)
    mesh%bf_nvert = mesh%bf_nvert(1:max(nb, 1))
    mesh%bf_vert = mesh%bf_vert(1:max(lb, 1))
    mesh%bf_tag = mesh%bf_tag(1:max(nb, 1))
  end subroutine

  ! build_mesh + compute_geometry_mesh on the device (one rank, sub-entities);
  ! periodic_mesh stitches x, y, z = -0.5 / 0.5 on the host connectivity first
  subroutine build_mesh(ctx, mesh, boundary_2d, periodic_mesh)
    type(c_ptr), intent(in) :: ctx
    type(raw_mesh), intent(in) :: mesh
    logical, intent(in) :: boundary_2d
    logical, intent(in), optional :: periodic_mesh
    integer(c_int32_t) :: b2d
    type(c_ptr) :: hm
    logical :: per
    b2d = merge(1, 0, boundary_2d)
    per = .false.
    if (present(periodic_mesh)) per = periodic_mesh
    if (.not. per) then
      call check(ctx, subfv_mesh_build(ctx, mesh%n_vert, mesh%coord, mesh%n_elems, mesh%kind, mesh%elem_vert, &
        mesh%n_bfaces, mesh%bf_nvert, mesh%bf_vert, mesh%bf_tag, b2d))
      return
    end if
    if (subfv_host_mesh_build(hm, mesh%n_vert, mesh%coord, mesh%n_elems, mesh%kind, mesh%elem_vert, &
        mesh%n_bfaces, mesh%bf_nvert, mesh%bf_vert, mesh%bf_tag) /= 0) error stop "build_mesh failed"
    if (subfv_host_mesh_make_periodic(hm) /= 0) then
      print *, "Error periodic boundaries !"
      error stop
    end if
    call check(ctx, subfv_mesh_upload(ctx, hm, b2d))
    call subfv_host_mesh_free(hm)
  end subroutine

  integer function mesh_count(ctx, what)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    mesh_count = int(subfv_mesh_count(ctx, cstr(what)))
  end function

  ! elem_coord(3, n_elems) and volumes, for initial conditions on the host
  subroutine elem_geometry(ctx, n_elems, coord, volume)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n_elems
    real(DP), intent(out) :: coord(3, n_elems), volume(n_elems)
    if (subfv_mesh_get_f64(ctx, cstr("elem_coord"), coord, int(3 * n_elems, c_int64_t)) /= 3 * n_elems) &
      error stop "elem_coord"
    if (subfv_mesh_get_f64(ctx, cstr("elem_volume"), volume, int(n_elems, c_int64_t)) /= n_elems) &
      error stop "elem_volume"
  end subroutine

  subroutine get_named(ctx, name, a, n)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real(DP), intent(out) :: a(n)
    if (subfv_get_f64(ctx, cstr(name), a, int(n, c_int64_t)) /= n) then
      print *, "device array ", trim(name), " unavailable"
      error stop
    end if
  end subroutine

  subroutine put_named(ctx, name, a, n)
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real(DP), intent(in) :: a(n)
    if (subfv_put_f64(ctx, cstr(name), a, int(n, c_int64_t)) /= n) then
      print *, "device array ", trim(name), " rejected"
      error stop
    end if
  end subroutine

  ! compute_euler_flux(mesh, sol, rhs, grad, sum_lambda, second_order, vp)
  subroutine compute_euler_flux(ctx, n_elems, n_vert, sol, rhs, grad, sum_lambda, second_order, vp)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n_elems, n_vert
    real(DP), intent(in) :: sol(5, n_elems), grad(3, 5, n_elems)
    real(DP), intent(out) :: rhs(5, n_elems), sum_lambda(n_elems)
    logical, intent(in) :: second_order
    real(DP), intent(out) :: vp(3, n_vert)
    call check(ctx, subfv_state_put(ctx, sol))
    if (second_order) then
      call put_named(ctx, "cell_grad", grad, 15 * n_elems)
      call check(ctx, subfv_compute_euler_flux(ctx, 1, 1))
      call get_named(ctx, "rhs_o2", rhs, 5 * n_elems)
      call get_named(ctx, "sum_lambda_o2", sum_lambda, n_elems)
    else
      call check(ctx, subfv_compute_euler_flux(ctx, 0, 1))
      call get_named(ctx, "rhs", rhs, 5 * n_elems)
      call get_named(ctx, "sum_lambda", sum_lambda, n_elems)
    end if
    call get_named(ctx, "vp", vp, 3 * n_vert)
  end subroutine

  ! compute_cell_grad_from_nodal_grad(mesh, sol, grad, all_nodal_grad, method)
  subroutine compute_cell_grad_from_nodal_grad(ctx, n_elems, n_vert, sol, grad, all_nodal_grad, method)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n_elems, n_vert, method
    real(DP), intent(in) :: sol(5, n_elems)
    real(DP), intent(out) :: grad(3, 5, n_elems), all_nodal_grad(3, 5, n_vert)
    call check(ctx, subfv_state_put(ctx, sol))
    call check(ctx, subfv_compute_cell_grad(ctx, int(method, c_int32_t)))
    call get_named(ctx, "cell_grad", grad, 15 * n_elems)
    call get_named(ctx, "nodal_grad", all_nodal_grad, 15 * n_vert)
  end subroutine

  ! flux_limiting(mesh, sol, rhs, rhs_o2, sum_lambda, cfl)
  subroutine flux_limiting(ctx, n_elems, sol, rhs, rhs_o2, sum_lambda, cfl)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n_elems
    real(DP), intent(in) :: sol(5, n_elems), rhs_o2(5, n_elems), sum_lambda(n_elems), cfl
    real(DP), intent(inout) :: rhs(5, n_elems)
    call check(ctx, subfv_state_put(ctx, sol))
    call put_named(ctx, "rhs", rhs, 5 * n_elems)
    call put_named(ctx, "rhs_o2", rhs_o2, 5 * n_elems)
    call put_named(ctx, "sum_lambda", sum_lambda, n_elems)
    call check(ctx, subfv_flux_limiting(ctx, cfl))
    call get_named(ctx, "rhs", rhs, 5 * n_elems)
  end subroutine

  ! add_timedisc_and_solve(mesh, sol, rhs, diag, sum_lambda, delta_sol, res,
  ! res_vect, num_procs, me, a1, a2, cfl, t, dt).  res_vect is the norm of
  ! rhs per variable on entry and res the sum of the cells' |sol_new - sol|
  ! / (a2 omega) norms, as the reference reports them.
  subroutine add_timedisc_and_solve(ctx, n_elems, sol, rhs, diag, sum_lambda, delta_sol, res, res_vect, &
      num_procs, me, a1, a2, cfl, t, dt)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: n_elems, num_procs, me
    real(DP), intent(inout) :: sol(5, n_elems), rhs(5, n_elems), diag(5, 5, n_elems), sum_lambda(n_elems)
    real(DP), intent(inout) :: delta_sol(5, n_elems), res, res_vect(5)
    real(DP), intent(in) :: a1, a2, cfl, t
    real(DP), intent(inout) :: dt
    real(DP) :: omega, old(5)
    integer :: i, k
    res_vect = 0.0_DP
    do i = 1, n_elems
      do k = 1, 5
        res_vect(k) = res_vect(k) + rhs(k, i)**2
      end do
    end do
    res_vect = sqrt(res_vect)
    call check(ctx, subfv_state_put(ctx, sol))
    call put_named(ctx, "rhs", rhs, 5 * n_elems)
    call put_named(ctx, "sum_lambda", sum_lambda, n_elems)
    if (any(diag /= 0.0_DP)) call put_named(ctx, "diag", diag, 25 * n_elems)
    call check(ctx, subfv_add_timedisc_and_solve(ctx, a1, a2, cfl, t, dt, omega))
    res = 0.0_DP
    do i = 1, n_elems
      old = sol(:, i)
      delta_sol(:, i) = old
    end do
    call check(ctx, subfv_state_get(ctx, sol))
    do i = 1, n_elems
      delta_sol(:, i) = (sol(:, i) - delta_sol(:, i)) / (a2 * omega)
      res = res + norm2(delta_sol(:, i))
    end do
    if (omega < 1.0_DP .and. me == 0 .and. num_procs >= 1) print *, "OMEGA", omega
  end subroutine

  ! ------------------------------------------------- lagrange / acoustic
  ! compute_rhs_lagrange(mesh, sol, vp, dt, rhs, n_bc, bc_type, bc_val, b2d, mass)
  ! on the context's mesh: sol(5, n_elems), mass(n_elems) -> rhs(5, n_elems),
  ! vp(3, n_vert)
  subroutine compute_rhs_lagrange(ctx, sol, vp, rhs, b2d, mass)
    type(c_ptr), intent(in) :: ctx
    real(DP), intent(in) :: sol(:, :), mass(:)
    real(DP), intent(out) :: vp(:, :), rhs(:, :)
    logical, intent(in) :: b2d
    call check(ctx, subfv_compute_rhs_lagrange(ctx, sol, mass, merge(1_c_int32_t, 0_c_int32_t, b2d), rhs, vp))
  end subroutine

  ! n_steps of the acoustic loop on sol(4, n_elems) = (p, u, v, w); bc_kind(t)
  ! for tag t: 0 wall, 1 freestream (bc_val(:, t)), 2 outflow
  subroutine acoustic_run(ctx, sol, n_steps, cfl, bc_kind, bc_val, t, dt)
    type(c_ptr), intent(in) :: ctx
    real(DP), intent(inout) :: sol(:, :), t
    integer, intent(in) :: n_steps
    real(DP), intent(in) :: cfl, bc_val(:, :)
    integer(c_int32_t), intent(in) :: bc_kind(:)
    real(DP), intent(out) :: dt
    call check(ctx, subfv_acoustic_run(ctx, sol, int(n_steps, c_int32_t), cfl, int(size(bc_kind), c_int32_t), &
      bc_kind, bc_val, t, dt))
  end subroutine

  ! ------------------------------------------------------------ BCSR
  type(c_ptr) function bcsr_create(ctx, m, b, nnzb, row, col, v) result(A)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: m, b, nnzb
    integer(c_int32_t), intent(in) :: row(m + 1), col(nnzb)
    real(DP), intent(in) :: v(b, b, nnzb)
    call check(ctx, subfv_bcsr_create(ctx, A, int(m, c_int32_t), int(b, c_int32_t), int(nnzb, c_int32_t), &
      row, col, v))
  end function

  ! y = bcsr_matmul(A, x)
  subroutine bcsr_matmul(ctx, A, x, y)
    type(c_ptr), intent(in) :: ctx, A
    real(DP), intent(in) :: x(:)
    real(DP), intent(out) :: y(:)
    call check(ctx, subfv_bcsr_matmul(A, x, y))
  end subroutine

  ! bcsr_diag(A, x, b)
  subroutine bcsr_diag(ctx, A, x, b)
    type(c_ptr), intent(in) :: ctx, A
    real(DP), intent(out) :: x(:)
    real(DP), intent(in) :: b(:)
    call check(ctx, subfv_bcsr_diag(A, x, b))
  end subroutine

  ! bcsr_gmres_diag(A, x, b, tol, m, conv): GMRES(m) from x = 0; conv is
  ! .false. when the restarts run out (not an error, as in the reference)
  subroutine bcsr_gmres_diag(ctx, A, x, b, tol, m, conv, fp32_inner)
    type(c_ptr), intent(in) :: ctx, A
    real(DP), intent(out) :: x(:)
    real(DP), intent(in) :: b(:), tol
    integer, intent(in) :: m
    logical, intent(out) :: conv
    logical, intent(in), optional :: fp32_inner
    integer(c_int32_t) :: c, prec
    integer(c_int) :: st
    real(DP) :: stats(3)
    prec = 0
    if (present(fp32_inner)) prec = merge(1, 0, fp32_inner)
    st = subfv_bcsr_gmres_diag(A, x, b, tol, int(m, c_int32_t), prec, c, stats)
    if (st /= SUBFV_OK .and. st /= SUBFV_E_NOT_CONVERGED) call check(ctx, st)
    conv = c /= 0
  end subroutine

  ! the GMRES variants on the same interface: precond 0 bcsr_gmres, 2
  ! bcsr_gmres_lusgs, 3 bcsr_gmres_parallel_lusgs (colour sets from
  ! bcsr_set_color_sets)
  subroutine bcsr_gmres_kind(ctx, A, x, b, tol, m, conv, precond)
    type(c_ptr), intent(in) :: ctx, A
    real(DP), intent(out) :: x(:)
    real(DP), intent(in) :: b(:), tol
    integer, intent(in) :: m, precond
    logical, intent(out) :: conv
    integer(c_int32_t) :: c
    integer(c_int) :: st
    real(DP) :: stats(3)
    st = subfv_bcsr_gmres(A, x, b, tol, int(m, c_int32_t), int(precond, c_int32_t), 0_c_int32_t, c, stats)
    if (st /= SUBFV_OK .and. st /= SUBFV_E_NOT_CONVERGED) call check(ctx, st)
    conv = c /= 0
  end subroutine

  ! bcsr_gmres(A, x, b, tol, m, conv, ...)
  subroutine bcsr_gmres(ctx, A, x, b, tol, m, conv)
    type(c_ptr), intent(in) :: ctx, A
    real(DP), intent(out) :: x(:)
    real(DP), intent(in) :: b(:), tol
    integer, intent(in) :: m
    logical, intent(out) :: conv
    call bcsr_gmres_kind(ctx, A, x, b, tol, m, conv, 0)
  end subroutine

  ! bcsr_gmres_lusgs(A, x, b, tol, m, conv, ...)
  subroutine bcsr_gmres_lusgs(ctx, A, x, b, tol, m, conv)
    type(c_ptr), intent(in) :: ctx, A
    real(DP), intent(out) :: x(:)
    real(DP), intent(in) :: b(:), tol
    integer, intent(in) :: m
    logical, intent(out) :: conv
    call bcsr_gmres_kind(ctx, A, x, b, tol, m, conv, 2)
  end subroutine

  ! the colour sets of bcsr_parallel_lusgs, given as a colour per block row
  subroutine bcsr_set_color_sets(ctx, A, color, n_color_sets)
    type(c_ptr), intent(in) :: ctx, A
    integer(c_int32_t), intent(in) :: color(:)
    integer, intent(out) :: n_color_sets
    integer(c_int32_t) :: ns
    call check(ctx, subfv_bcsr_set_colors(A, color, ns))
    n_color_sets = ns
  end subroutine

  ! bcsr_gmres_parallel_lusgs(A, x, b, tol, m, n_color_sets, color_set, color, conv, ...)
  subroutine bcsr_gmres_parallel_lusgs(ctx, A, x, b, tol, m, conv)
    type(c_ptr), intent(in) :: ctx, A
    real(DP), intent(out) :: x(:)
    real(DP), intent(in) :: b(:), tol
    integer, intent(in) :: m
    logical, intent(out) :: conv
    call bcsr_gmres_kind(ctx, A, x, b, tol, m, conv, 3)
  end subroutine

  ! bcsr_lusgs(A, x, b)
  subroutine bcsr_lusgs(ctx, A, x, b)
    type(c_ptr), intent(in) :: ctx, A
    real(DP), intent(out) :: x(:)
    real(DP), intent(in) :: b(:)
    call check(ctx, subfv_bcsr_lusgs(A, x, b))
  end subroutine

  ! bcsr_parallel_lusgs(A, x, b, n_color_sets, color_set, color)
  subroutine bcsr_parallel_lusgs(ctx, A, x, b)
    type(c_ptr), intent(in) :: ctx, A
    real(DP), intent(out) :: x(:)
    real(DP), intent(in) :: b(:)
    call check(ctx, subfv_bcsr_parallel_lusgs(A, x, b))
  end subroutine

  ! bcsr_pcgm(A, x, b, tol): x is the initial guess on entry
  subroutine bcsr_pcgm(ctx, A, x, b, tol)
    type(c_ptr), intent(in) :: ctx, A
    real(DP), intent(inout) :: x(:)
    real(DP), intent(in) :: b(:), tol
    real(DP) :: stats(2)
    call check(ctx, subfv_bcsr_pcgm(A, x, b, tol, stats))
  end subroutine

  ! add_block_to_bcsr(mat, i, j, M)
  subroutine add_block_to_bcsr(ctx, A, i, j, M)
    type(c_ptr), intent(in) :: ctx, A
    integer, intent(in) :: i, j
    real(DP), intent(in) :: M(:, :)
    call check(ctx, subfv_bcsr_add_block(A, int(i, c_int32_t), int(j, c_int32_t), M))
  end subroutine
end module subfv_ns_gpu
