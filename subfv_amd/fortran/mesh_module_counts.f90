! The components of the reference's mesh_type (src/mesh/mesh_module.F90:17-88)
! that the reference-signature procedures (subfv_reference_api) read or fill,
! under the reference's module, type and component names, for builds without
! the reference's own mesh module: the counts, and the vertices, elements and
! boundary faces of read_mesh_msh_2 / build_mesh / compute_geometry_mesh.
! The faces, sub-faces and sub-elements stay on the device.  Inside the
! reference's build this file is left out: subfv_reference_api then binds to
! the reference's mesh_type, which has components of all these names.
module mesh_module
  implicit none
  integer, parameter, private :: DP = kind(1.0d0)

  type :: vert_type
    logical :: is_bound = .false.
    real(DP), dimension(3) :: coord = 0.0_DP
    real(DP) :: volume = 0.0_DP
  end type vert_type

  type :: boundary_face_type
    integer :: n_vert = -1
    integer, dimension(:), allocatable :: vert
    integer :: tag = 0
  end type boundary_face_type

  type :: elem_type
    integer :: n_vert = 0, n_faces = 0, elem_kind = 0, tag = -1
    integer, dimension(:), allocatable :: vert
    real(DP) :: volume = 0.0_DP
    real(DP), dimension(3) :: coord = 0.0_DP
  end type elem_type

  type :: mesh_type
    integer :: n_vert = 0, n_elems = 0, n_faces = 0, n_interior_elems = 0
    integer :: n_sub_faces = 0, n_boundary_faces = 0, n_sub_elems = 0
    integer :: n_max_sub_faces_around_node = 0
    type(vert_type), dimension(:), allocatable :: vert
    type(elem_type), dimension(:), allocatable :: elem
    type(boundary_face_type), dimension(:), allocatable :: boundary_face
  end type mesh_type
end module mesh_module
