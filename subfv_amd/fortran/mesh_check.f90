! Reads a msh 2.2 mesh with the Fortran reader and builds its connectivity on
! the host through the C-ABI (no device): prints the entity counts and
! writes sub_face_left_elem, elem_vert and face_right (int32 stream, each
! preceded by its length) to mesh_check.bin.
! Usage: mesh_check <mesh file> [bc name ...]
program mesh_check
  use iso_c_binding
  use subfv_c
  use subfv_ns_gpu
  implicit none
  type(raw_mesh) :: mesh
  type(c_ptr) :: hm
  character(len=255) :: path, names(10)
  integer :: nbc, i, funit
  integer(c_int64_t) :: n1, n2, n3
  integer(c_int32_t), allocatable :: a(:), b(:), c(:)
  call get_command_argument(1, path)
  nbc = command_argument_count() - 1
  names = ""
  do i = 1, nbc
    call get_command_argument(i + 1, names(i))
  end do
  call read_mesh_msh_2(mesh, "", path, nbc, names)
  if (subfv_host_mesh_build(hm, mesh%n_vert, mesh%coord, mesh%n_elems, mesh%kind, mesh%elem_vert, &
      mesh%n_bfaces, mesh%bf_nvert, mesh%bf_vert, mesh%bf_tag) /= 0) error stop "host mesh build failed"
  allocate (a(64 * (mesh%n_vert + mesh%n_elems) + 1024), b(64 * (mesh%n_vert + mesh%n_elems) + 1024))
  allocate (c(64 * (mesh%n_vert + mesh%n_elems) + 1024))
  n1 = subfv_host_mesh_get_i32(hm, cstr("sub_face_left_elem"), a, int(size(a), c_int64_t))
  n2 = subfv_host_mesh_get_i32(hm, cstr("elem_vert"), b, int(size(b), c_int64_t))
  n3 = subfv_host_mesh_get_i32(hm, cstr("face_right"), c, int(size(c), c_int64_t))
  print *, mesh%n_vert, mesh%n_elems, mesh%n_bfaces, n1, n2
  open (newunit=funit, file="mesh_check.bin", access="stream", form="unformatted", status="replace")
  write (funit) int(n1, c_int32_t), a(1:n1), int(n2, c_int32_t), b(1:n2), int(n3, c_int32_t), c(1:n3)
  close (funit)
  call subfv_host_mesh_free(hm)
end program mesh_check
