// Host-side connectivity of the subface mesh, numbered exactly as subfv's
// build_mesh numbers it (src/mesh/mesh_connectivity_module.F90:12-59, one rank;
// the periodic stitch is a separate step, make_periodic).  All ids are 0-based here; the reference's 1-based numbers
// are these plus one.  The right neighbour of a face keeps the reference
// encoding: >0 one-based cell id, 0 default wall, <0 minus the BC index.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace subfv {

struct HostMesh {
  int nv = 0, ne = 0, nf = 0, nsf = 0, nse = 0;
  int nmax_sfn = 0, nmax_sen = 0;
  std::vector<double> vcoord;             // 3*nv, renumbered
  std::vector<int> vert_nto, elem_nto;    // new -> file id (0-based)
  // elements
  std::vector<int> e_kind, e_voff, e_vert;   // e_vert: new vertex ids
  std::vector<int> e_foff, e_face;           // local face -> face id
  std::vector<int> e_sfoff, e_sf;            // sub-faces in build order
  // vertex -> element lists (reference order)
  std::vector<int> v_eoff, v_elem;
  // faces
  std::vector<int> f_voff, f_vert, f_left, f_leftf, f_right, f_rightf;
  std::vector<int> f_sfoff, f_sf;
  std::vector<int> v_foff, v_face;
  // sub-faces (vertex-major): those of vertex v are [sf_off[v], sf_off[v+1])
  std::vector<int> sf_off, sf_vert, sf_face, sf_le, sf_re, sf_lse, sf_rse;
  // sub-elements (cell-major: those of cell e are [e_voff[e], e_voff[e+1]))
  std::vector<int> se_elem, se_vert, se_loc, se_sfoff, se_sf;
  // vertex -> sub-element lists (ascending sub-element id)
  std::vector<int> v_seoff, v_se;
  // element tags (smallest BC tag on its boundary faces, -1 if none)
  std::vector<int> e_tag;
  // cells sharing a vertex with the cell (neigh_by_vert)
  std::vector<int> e_nbvoff, e_nbv;
  // periodic meshes (make_periodic): each vertex's sub-face list as the node
  // sweep reads it, its own sub-faces first, then those its periodic partners
  // bring (x_sf holds sub-face ids); and the sub-elements the partners bring,
  // after the vertex's own (x_se, cell-major ids).  Empty otherwise.
  bool periodic = false;
  std::vector<int> x_sfoff, x_sf, x_seoff, x_se;
};

// kind: gmsh kinds (4 tet, 5 hex, 6 prism, 7 pyramid); elem_vert: 1-based
// file vertex ids; boundary faces likewise.  Returns an empty string on
// success or an error message.
std::string build_host_mesh(HostMesh &m, int nv, const double *coord, int ne, const int32_t *kind,
                            const int32_t *elem_vert, int nbf, const int32_t *bf_nvert,
                            const int32_t *bf_vert, const int32_t *bf_tag);

// The same connectivity built on the device (mesh_device.hip): identical
// numbering, computed by data-parallel passes, stable sorts and a
// level-synchronous breadth-first search on `stream` (a hipStream_t), then
// copied into m.
std::string build_device_mesh(HostMesh &m, int nv, const double *coord, int ne, const int32_t *kind,
                              const int32_t *elem_vert, int nbf, const int32_t *bf_nvert, const int32_t *bf_vert,
                              const int32_t *bf_tag, void *stream);

// The periodic part of build_mesh (src/mesh/mesh_connectivity_module.F90:44-54)
// on a built connectivity: faces on x, y, z = -0.5 and 0.5 are stitched to
// their partners, and the vertices there take their partners' sub-faces and
// sub-elements into their lists.  Returns an empty string or an error.
std::string make_periodic(HostMesh &m);

int gmsh_n_vert(int kind);
int gmsh_n_face(int kind);

}  // namespace subfv
