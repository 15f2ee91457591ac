! The counts of the reference's mesh_type (src/mesh/mesh_module.F90:72-76)
! under the reference's module and type names, for builds of the
! reference-signature procedures (subfv_reference_api) without the
! reference's own mesh module.  Inside the reference's build this file is
! left out: subfv_reference_api then binds to the reference's mesh_type,
! whose components of these names are the only ones it reads.
module mesh_module
  implicit none

  type :: mesh_type
    integer :: n_vert = 0, n_elems = 0, n_faces = 0, n_interior_elems = 0
    integer :: n_sub_faces = 0, n_boundary_faces = 0, n_sub_elems = 0
    integer :: n_max_sub_faces_around_node = 0
  end type mesh_type
end module mesh_module
