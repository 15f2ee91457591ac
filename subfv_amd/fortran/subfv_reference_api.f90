! The reference's module procedures of the explicit path, of its viscous
! terms, of its mesh set-up and of its BCSR solver, with the reference's own
! names and argument lists, running on the device context this module holds:
!   read_mesh_msh_2(mesh, meshfile_path, meshfile, n_bc, bc_name)
!       (src/mesh/mesh_reading_module.F90:989-1324)
!   build_mesh(mesh, num_procs, mpi_send_recv, use_sub_entities, b2d, periodic_mesh)
!       (src/mesh/mesh_connectivity_module.F90:12-59)
!   compute_geometry_mesh(mesh, use_sub_entities, b2d, me, num_procs)
!       (src/mesh/mesh_geometry_module.F90:16-57)
!   compute_euler_flux(mesh, sol, rhs, grad, sum_lambda, second_order, vp, h_p, mat_h_p)
!       (src/ns/ns_euler_module.F90:151-174)
!   compute_cell_grad_from_nodal_grad(mesh, sol, grad, all_nodal_grad, method)
!       (src/ns/ns_euler_recon_module.F90:8-15)
!   flux_limiting(mesh, sol, rhs, rhs_o2, sum_lambda, cfl)
!       (src/ns/ns_euler_recon_module.F90:560-569)
!   add_timedisc_and_solve(mesh, sol, rhs, diag, sum_lambda, delta_sol, res,
!       res_vect, num_procs, me, a1, a2, cfl, t, dt)  (src/ns/ns_euler_module.F90:521-540)
!   compute_mu(mesh, sol, num_procs, mu_arr, mu0corr)
!       (src/ns/ns_vectorial_diffusion_module.F90:322-347)
!   build_momentum_diff_mat_into_global_mat_diag(mesh, diag, rhs, sol, mu_arr, mu0corr)
!       (src/ns/ns_vectorial_diffusion_module.F90:200-320)
!   build_diff_mat_into_global_mat_diag(mesh, diag, rhs, sol, mu_arr)
!       (src/ns/ns_diffusion_module.F90:112-184)
!   bcsr_matrix_type, y = bcsr_matmul(A, x), bcsr_diag(A, x, b),
!   bcsr_gmres_diag(A, x, b, tol, m, conv, H, V, r, y, c, s, z, zp)
!       (src/linear_algebra/sparse_bcsr_linear_module.F90:6-29,1414,1592)
!
! build_mesh puts the mesh on the device context (the one given to
! subfv_ref_use_context, or one it makes on device 0); the mesh_type keeps
! the counts, the vertices, the elements and the boundary faces, and the
! residual procedures read its counts only.  The scheme and its options come
! from the context's configuration (subfv_set_config, subfv_set_diffusion),
! as the reference takes them from ns_global_data_module.  A failure prints
! the reference's message and stops, as the reference does.
module subfv_reference_api
  use iso_c_binding
  use subfv_c
  use subfv_ns_gpu, only: check, get_named, put_named, raw_mesh, verts_of_kind, read_raw_msh => read_mesh_msh_2
  use mesh_module, only: mesh_type
  implicit none
  private
  public :: mesh_type, bcsr_matrix_type, subfv_ref_use_context, subfv_ref_context, subfv_ref_release
  public :: read_mesh_msh_2, build_mesh, compute_geometry_mesh
  public :: compute_mu, build_momentum_diff_mat_into_global_mat_diag, build_diff_mat_into_global_mat_diag
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
  logical, save :: own_ctx = .false.  ! made by build_mesh: destroyed by subfv_ref_release
  integer, save :: built_b2d = -1     ! boundary_2d of the mesh build_mesh made
  ! the device copy of the last matrix: its pattern is kept while the row and
  ! col arrays and the sizes stay the same; the values are sent on every call,
  ! since the reference refills v in place between solves
  type(c_ptr), save :: dmat = c_null_ptr
  integer, save :: dm_m = -1, dm_b = -1, dm_nnz = -1
  integer(c_intptr_t), save :: dm_row = 0, dm_col = 0

contains

  ! the device context the procedures run on (a context build_mesh made
  ! before is destroyed, unless it is c itself: the caller then owns it)
  subroutine subfv_ref_use_context(c)
    type(c_ptr), intent(in) :: c
    if (own_ctx) then
      if (c_associated(ctx, c)) then
        own_ctx = .false.
      else
        call subfv_ref_release()
      end if
    end if
    ctx = c
  end subroutine

  type(c_ptr) function subfv_ref_context() result(c)
    c = ctx
  end function

  ! drops the device copy of the last matrix
  subroutine drop_matrix()
    if (c_associated(dmat)) call subfv_bcsr_destroy(dmat)
    dmat = c_null_ptr
    dm_m = -1
  end subroutine

  ! drops the device copy of the last matrix, and the context build_mesh made
  subroutine subfv_ref_release()
    call drop_matrix()
    if (own_ctx) then
      call subfv_destroy(ctx)
      ctx = c_null_ptr
      own_ctx = .false.
    end if
  end subroutine

  subroutine need_mesh(mesh)
    type(mesh_type), intent(in) :: mesh
    if (.not. c_associated(ctx)) error stop "subfv_reference_api: no device context (subfv_ref_use_context)"
    if (subfv_mesh_count(ctx, "n_elems" // c_null_char) /= mesh%n_elems .or. &
        subfv_mesh_count(ctx, "n_vert" // c_null_char) /= mesh%n_vert) &
      error stop "subfv_reference_api: the device context holds another mesh"
  end subroutine

  ! faces of a volume element kind: tetrahedron 4, hexahedron 6, prism and
  ! pyramid 5 (mesh_reading_module.F90 n_face_kind)
  pure integer function faces_of_kind(kind) result(n)
    integer, intent(in) :: kind
    select case (kind)
    case (4); n = 4
    case (5); n = 6
    case (6, 7); n = 5
    case default; n = 0
    end select
  end function

  ! read_mesh_msh_2: the vertices, the volume elements (kinds 4-7) and the
  ! boundary faces (kinds 2, 3; tag = index of the bc whose name carries the
  ! face's physical tag, 0 for none) of the msh 2.2 file, in file order.
  ! mesh%n_faces counts the elements' faces before build_mesh merges the
  ! shared ones, as the reference's count does.
  subroutine read_mesh_msh_2(mesh, meshfile_path, meshfile, n_bc, bc_name)
    type(mesh_type), intent(inout) :: mesh
    character(len=*), intent(in) :: meshfile, meshfile_path
    integer, intent(in) :: n_bc
    character(len=*), dimension(:), intent(in) :: bc_name
    type(raw_mesh) :: raw
    integer :: i, o, nv
    call read_raw_msh(raw, meshfile_path, meshfile, n_bc, bc_name)
    mesh%n_vert = raw%n_vert
    mesh%n_elems = raw%n_elems
    mesh%n_boundary_faces = raw%n_bfaces
    if (allocated(mesh%vert)) deallocate (mesh%vert)
    if (allocated(mesh%elem)) deallocate (mesh%elem)
    if (allocated(mesh%boundary_face)) deallocate (mesh%boundary_face)
    allocate (mesh%vert(raw%n_vert), mesh%elem(raw%n_elems), mesh%boundary_face(raw%n_bfaces))
    do i = 1, raw%n_vert
      mesh%vert(i)%coord = raw%coord(:, i)
    end do
    mesh%n_faces = 0
    o = 0
    do i = 1, raw%n_elems
      nv = verts_of_kind(raw%kind(i))
      mesh%elem(i)%elem_kind = raw%kind(i)
      mesh%elem(i)%n_vert = nv
      mesh%elem(i)%n_faces = faces_of_kind(raw%kind(i))
      mesh%elem(i)%vert = raw%elem_vert(o + 1:o + nv)
      mesh%n_faces = mesh%n_faces + mesh%elem(i)%n_faces
      o = o + nv
    end do
    o = 0
    do i = 1, raw%n_bfaces
      nv = raw%bf_nvert(i)
      mesh%boundary_face(i)%n_vert = nv
      mesh%boundary_face(i)%vert = raw%bf_vert(o + 1:o + nv)
      mesh%boundary_face(i)%tag = raw%bf_tag(i)
      o = o + nv
    end do
  end subroutine

  ! build_mesh on the device context: connectivity, sub-faces, sub-elements
  ! and geometry of the mesh read_mesh_msh_2 filled, for one rank
  ! (num_procs = 1; mpi_send_recv is not read).  The sub-entities are built
  ! whatever use_sub_entities says.  On return mesh%vert and mesh%elem are in
  ! the reference's renumbered order (renumber_vert_mesh,
  ! renumber_elem_mesh), the elements' and boundary faces' vertex ids
  ! renumbered, elem%tag set (compute_elem_tag) and the counts those of the
  ! built mesh.  periodic_mesh stitches the faces on x, y, z = -0.5 / 0.5
  ! (stitch_periodic_faces).
  subroutine build_mesh(mesh, num_procs, mpi_send_recv, use_sub_entities, b2d, periodic_mesh)
    type(mesh_type), intent(inout) :: mesh
    integer, intent(in) :: num_procs
    class(*), intent(inout) :: mpi_send_recv
    logical, intent(in) :: use_sub_entities, b2d
    logical, intent(in), optional :: periodic_mesh
    real(DP), allocatable :: coord(:, :)
    integer(c_int32_t), allocatable :: kind(:), ev(:), bfn(:), bfv(:), bft(:), nto(:), otn(:), etag(:)
    integer(c_int32_t) :: ib2d
    integer(c_int) :: st
    type(c_ptr) :: hm
    logical :: per
    integer :: i, j, o, nv, ne, nb, lv, lb
    if (num_procs > 1) error stop "build_mesh: one rank's mesh only (num_procs = 1)"
    if (.not. allocated(mesh%vert) .or. .not. allocated(mesh%elem)) error stop "build_mesh: no mesh read"
    if (.not. c_associated(ctx)) then
      st = subfv_create(ctx, 0)
      if (st /= SUBFV_OK) then
        print *, "build_mesh: no device context on device 0, status", st
        error stop
      end if
      own_ctx = .true.
    end if
    nv = mesh%n_vert
    ne = mesh%n_elems
    nb = mesh%n_boundary_faces
    lv = 0
    do i = 1, ne
      lv = lv + mesh%elem(i)%n_vert
    end do
    lb = 0
    do i = 1, nb
      lb = lb + mesh%boundary_face(i)%n_vert
    end do
    allocate (coord(3, nv), kind(ne), ev(max(lv, 1)), bfn(max(nb, 1)), bfv(max(lb, 1)), bft(max(nb, 1)))
    do i = 1, nv
      coord(:, i) = mesh%vert(i)%coord
    end do
    o = 0
    do i = 1, ne
      kind(i) = mesh%elem(i)%elem_kind
      ev(o + 1:o + mesh%elem(i)%n_vert) = mesh%elem(i)%vert(1:mesh%elem(i)%n_vert)
      o = o + mesh%elem(i)%n_vert
    end do
    o = 0
    do i = 1, nb
      bfn(i) = mesh%boundary_face(i)%n_vert
      bft(i) = mesh%boundary_face(i)%tag
      bfv(o + 1:o + bfn(i)) = mesh%boundary_face(i)%vert(1:bfn(i))
      o = o + bfn(i)
    end do
    ib2d = merge(1, 0, b2d)
    per = .false.
    if (present(periodic_mesh)) per = periodic_mesh
    if (per) then
      if (subfv_host_mesh_build(hm, nv, coord, ne, kind, ev, nb, bfn, bfv, bft) /= 0) then
        print *, "build_mesh failed"
        error stop
      end if
      if (subfv_host_mesh_make_periodic(hm) /= 0) then
        call subfv_host_mesh_free(hm)
        print *, "Error periodic boundaries !"
        error stop
      end if
      call check(ctx, subfv_mesh_upload(ctx, hm, ib2d))
      call subfv_host_mesh_free(hm)
    else
      call check(ctx, subfv_mesh_build(ctx, nv, coord, ne, kind, ev, nb, bfn, bfv, bft, ib2d))
    end if
    built_b2d = ib2d
    ! the renumbered vertices: coordinates in the new order, old ids to new
    allocate (nto(nv), otn(nv), etag(ne))
    if (subfv_mesh_get_i32(ctx, "vert_nto" // c_null_char, nto, int(nv, c_int64_t)) /= nv) error stop "vert_nto"
    do i = 1, nv
      otn(nto(i)) = i
    end do
    if (subfv_mesh_get_f64(ctx, "vert_coord" // c_null_char, coord, int(3 * nv, c_int64_t)) /= 3 * nv) &
      error stop "vert_coord"
    do i = 1, nv
      mesh%vert(i)%coord = coord(:, i)
    end do
    do i = 1, nb
      do j = 1, mesh%boundary_face(i)%n_vert
        mesh%boundary_face(i)%vert(j) = otn(mesh%boundary_face(i)%vert(j))
      end do
    end do
    ! the renumbered elements, their vertices as the device mesh holds them
    if (subfv_mesh_get_i32(ctx, "elem_kind" // c_null_char, kind, int(ne, c_int64_t)) /= ne) error stop "elem_kind"
    if (subfv_mesh_get_i32(ctx, "elem_vert" // c_null_char, ev, int(lv, c_int64_t)) /= lv) error stop "elem_vert"
    if (subfv_mesh_get_i32(ctx, "elem_tag" // c_null_char, etag, int(ne, c_int64_t)) /= ne) error stop "elem_tag"
    o = 0
    do i = 1, ne
      mesh%elem(i)%elem_kind = kind(i)
      mesh%elem(i)%n_vert = verts_of_kind(kind(i))
      mesh%elem(i)%n_faces = faces_of_kind(kind(i))
      mesh%elem(i)%vert = ev(o + 1:o + mesh%elem(i)%n_vert)
      mesh%elem(i)%tag = etag(i)
      o = o + mesh%elem(i)%n_vert
    end do
    mesh%n_faces = int(subfv_mesh_count(ctx, "n_faces" // c_null_char))
    mesh%n_sub_faces = int(subfv_mesh_count(ctx, "n_sub_faces" // c_null_char))
    mesh%n_sub_elems = int(subfv_mesh_count(ctx, "n_sub_elems" // c_null_char))
    mesh%n_max_sub_faces_around_node = int(subfv_mesh_count(ctx, "n_max_sub_faces" // c_null_char))
  end subroutine

  ! compute_geometry_mesh: build_mesh computed the geometry on the device;
  ! the elements' centroids and volumes, the vertices' volumes
  ! (compute_vert_volume) and boundary flags (find_if_vert_is_bound) are
  ! copied into mesh%elem and mesh%vert.  The face, sub-face and sub-element
  ! geometry stays on the device.  b2d must be the one given to build_mesh.
  subroutine compute_geometry_mesh(mesh, use_sub_entities, b2d, me, num_procs)
    type(mesh_type), intent(inout) :: mesh
    logical, intent(in) :: use_sub_entities, b2d
    integer, optional, intent(in) :: me, num_procs
    real(DP), allocatable :: xc(:, :), vol(:)
    integer(c_int32_t), allocatable :: ib(:)
    integer :: i, ne, nv
    call need_mesh(mesh)
    if (merge(1, 0, b2d) /= built_b2d) error stop "compute_geometry_mesh: b2d differs from build_mesh's"
    ne = mesh%n_elems
    nv = mesh%n_vert
    allocate (xc(3, ne), vol(max(ne, nv)), ib(nv))
    if (subfv_mesh_get_f64(ctx, "elem_coord" // c_null_char, xc, int(3 * ne, c_int64_t)) /= 3 * ne) &
      error stop "elem_coord"
    if (subfv_mesh_get_f64(ctx, "elem_volume" // c_null_char, vol, int(ne, c_int64_t)) /= ne) error stop "elem_volume"
    do i = 1, ne
      mesh%elem(i)%coord = xc(:, i)
      mesh%elem(i)%volume = vol(i)
    end do
    if (subfv_mesh_get_f64(ctx, "vert_volume" // c_null_char, vol, int(nv, c_int64_t)) /= nv) error stop "vert_volume"
    if (subfv_mesh_get_i32(ctx, "vert_is_bound" // c_null_char, ib, int(nv, c_int64_t)) /= nv) &
      error stop "vert_is_bound"
    do i = 1, nv
      mesh%vert(i)%volume = vol(i)
      mesh%vert(i)%is_bound = ib(i) /= 0
    end do
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

  ! compute_mu: mu_arr and mu0corr of sol, Sutherland's law per cell when
  ! subfv_set_sutherland activated it, else mu_p of subfv_set_diffusion
  ! (one rank: num_procs is not read)
  subroutine compute_mu(mesh, sol, num_procs, mu_arr, mu0corr)
    type(mesh_type), intent(in) :: mesh
    real(DP), dimension(5, mesh%n_elems), intent(in) :: sol
    integer, intent(in) :: num_procs
    real(DP), dimension(mesh%n_elems), intent(out) :: mu_arr
    real(DP), intent(out) :: mu0corr
    call need_mesh(mesh)
    call check(ctx, subfv_compute_mu(ctx, sol, mu_arr, mu0corr))
  end subroutine

  ! the momentum system's contributions (diag(2:4, 2:4), rhs(2:4) and the
  ! viscous work in rhs(5)) for the cell viscosities mu_arr and mu0corr,
  ! added to diag and rhs
  subroutine build_momentum_diff_mat_into_global_mat_diag(mesh, diag, rhs, sol, mu_arr, mu0corr)
    type(mesh_type), intent(in) :: mesh
    real(DP), dimension(5, 5, mesh%n_elems) :: diag
    real(DP), dimension(5, mesh%n_elems) :: rhs
    real(DP), dimension(5, mesh%n_elems) :: sol
    real(DP), dimension(mesh%n_elems), intent(in) :: mu_arr
    real(DP), intent(in) :: mu0corr
    call need_mesh(mesh)
    call check(ctx, subfv_add_diffusion(ctx, sol, mu_arr, mu0corr, 1_c_int32_t, rhs, diag))
  end subroutine

  ! the heat system's contributions (diag(5, :) and rhs(5)) for the cell
  ! viscosities mu_arr (kappa = mu Cv gamma / Pr), added to diag and rhs
  subroutine build_diff_mat_into_global_mat_diag(mesh, diag, rhs, sol, mu_arr)
    type(mesh_type), intent(in) :: mesh
    real(DP), dimension(5, 5, mesh%n_elems), intent(inout) :: diag
    real(DP), dimension(5, mesh%n_elems), intent(inout) :: rhs
    real(DP), dimension(5, mesh%n_elems), intent(in) :: sol
    real(DP), dimension(mesh%n_elems), intent(in) :: mu_arr
    call need_mesh(mesh)
    ! the momentum system, formed alongside, is not used: its mu0corr is
    ! given a value that keeps it regular
    call check(ctx, subfv_add_diffusion(ctx, sol, mu_arr, minval(mu_arr), 2_c_int32_t, rhs, diag))
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
      call drop_matrix()
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
