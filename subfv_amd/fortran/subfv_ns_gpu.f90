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

  ! gmsh element kind -> number of vertices (mesh_reading_module.F90:1006-1009)
  integer, parameter :: n_vert_kind(31) = (/ &
    2, 3, 4, 4, 8, 6, 5, 3, 6, 9, 10, &
    27, 18, 14, 1, 8, 20, 15, 13, 9, 10, 12, &
    15, 15, 21, 4, 5, 6, 20, 35, 56/)

  type :: raw_mesh
    integer(c_int32_t) :: n_vert = 0, n_elems = 0, n_bfaces = 0
    real(c_double), allocatable :: coord(:, :)
    integer(c_int32_t), allocatable :: kind(:), elem_vert(:)
    integer(c_int32_t), allocatable :: bf_nvert(:), bf_vert(:), bf_tag(:)
  end type raw_mesh

contains

  subroutine check(ctx, st)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: st
    if (st /= SUBFV_OK) then
      print *, trim(subfv_message(ctx))
      error stop
    end if
  end subroutine

  ! msh 2.2 reader with the reference's semantics: boundary faces (gmsh kinds
  ! 2 and 3) get the index j of the bc_name whose physical tag they carry, 0
  ! (default wall) otherwise; volume elements are kinds 4-7 in file order.
  subroutine read_mesh_msh_2(mesh, meshfile_path, meshfile, n_bc, bc_name)
    type(raw_mesh), intent(out) :: mesh
    character(len=*), intent(in) :: meshfile_path, meshfile
    integer, intent(in) :: n_bc
    character(len=*), dimension(:), intent(in) :: bc_name
    integer :: funit, i, j, n_items, dummy, kind, ntag, nv, bc_tag(max(n_bc, 1)), n_bc_tmp, tag
    integer :: iv, ib, ie, nbv
    integer :: tags(16), verts(64)
    character(len=255) :: text
    integer, allocatable :: kinds(:), nvs(:), ptag(:), vlist(:, :)

    bc_tag = -huge(1)
    open (newunit=funit, file=trim(adjustl(meshfile_path))//trim(adjustl(meshfile)), status="old")
    if (n_bc /= 0) then
      read (funit, '(a)') text
      do while ("$PhysicalNames" /= text(1:14))
        read (funit, '(a)') text
        if ("$Nodes" == text(1:6)) then
          print *, "No Physical names inside mesh!!!"
          error stop
        end if
      end do
      read (funit, *) n_bc_tmp
      do i = 1, n_bc_tmp
        read (funit, *) dummy, tag, text
        do j = 1, n_bc
          if (trim(adjustl(bc_name(j))) == trim(adjustl(text))) bc_tag(j) = tag
        end do
      end do
    end if
    read (funit, '(a)') text
    do while ("$Nodes" /= text(1:6))
      read (funit, '(a)') text
    end do
    read (funit, *) mesh%n_vert
    allocate (mesh%coord(3, mesh%n_vert))
    do i = 1, mesh%n_vert
      read (funit, *) dummy, mesh%coord(1, i), mesh%coord(2, i), mesh%coord(3, i)
    end do
    read (funit, '(a)') text
    do while ("$Elements" /= text(1:9))
      read (funit, '(a)') text
    end do
    read (funit, *) n_items
    allocate (kinds(n_items), nvs(n_items), ptag(n_items), vlist(8, n_items))
    mesh%n_elems = 0
    mesh%n_bfaces = 0
    nbv = 0
    iv = 0
    do i = 1, n_items
      read (funit, '(a)') text
      read (text, *) dummy, kind, ntag
      nv = n_vert_kind(kind)
      read (text, *) dummy, kind, ntag, tags(1:ntag), verts(1:nv)
      kinds(i) = kind
      nvs(i) = nv
      ptag(i) = tags(1)
      vlist(:, i) = 0
      if (nv <= 8) vlist(1:nv, i) = verts(1:nv)
      if (kind == 2 .or. kind == 3) then
        mesh%n_bfaces = mesh%n_bfaces + 1
        nbv = nbv + nv
      else if (kind >= 4 .and. kind <= 7) then
        mesh%n_elems = mesh%n_elems + 1
        iv = iv + nv
      end if
    end do
    close (funit)
    if (mesh%n_elems < 1) then
      print *, "[-] No 3D elements in mesh !"
      error stop
    end if
    allocate (mesh%kind(mesh%n_elems), mesh%elem_vert(iv))
    allocate (mesh%bf_nvert(max(mesh%n_bfaces, 1)), mesh%bf_vert(max(nbv, 1)), mesh%bf_tag(max(mesh%n_bfaces, 1)))
    ie = 0
    ib = 0
    iv = 0
    nbv = 0
    do i = 1, n_items
      if (kinds(i) == 2 .or. kinds(i) == 3) then
        ib = ib + 1
        mesh%bf_nvert(ib) = nvs(i)
        mesh%bf_vert(nbv + 1:nbv + nvs(i)) = vlist(1:nvs(i), i)
        nbv = nbv + nvs(i)
        mesh%bf_tag(ib) = 0
        do j = 1, n_bc
          if (bc_tag(j) == ptag(i)) mesh%bf_tag(ib) = j
        end do
      else if (kinds(i) >= 4 .and. kinds(i) <= 7) then
        ie = ie + 1
        mesh%kind(ie) = kinds(i)
        mesh%elem_vert(iv + 1:iv + nvs(i)) = vlist(1:nvs(i), i)
        iv = iv + nvs(i)
      end if
    end do
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
