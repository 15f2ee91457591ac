// Device-resident mesh: structure-of-arrays geometry and the node-sorted
// sub-face / sub-element CSR the kernels stream through.  0-based ids.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfv {

// Vertex groups of the node sweep: one 256-thread workgroup per group of
// consecutive vertices.  Group-local sub-face indices fit a byte (0xFF ends a
// slot's contributor list) and the group's vertices fit one wavefront.
constexpr int kGrpThreads = 256;
constexpr int kGrpSub = 255;    // sub-faces per group at most
constexpr int kGrpVert = 64;    // vertices per group at most
constexpr int kGrpSlots = 256;  // sub-element slots per group at most
constexpr int kGrpHdr = 8;      // ints per group header: vb, nv, sf0, nsf, se0, nse, -, -
constexpr int kSlotSrc = 4;     // sub-faces of one cell meeting at a vertex (pyramid apex: 4)
constexpr int kRedStride = 257; // padded LDS row (doubles) against bank conflicts
// node-sorted slot rows of the node sweep's partials: 6 values in an 8-double
// (64-byte, aligned) row, so the gather's scattered row reads touch one
// 128-byte line each instead of straddling two
constexpr int kSlotStride = 8;

struct DevMesh {
  long V = 0, C = 0, F = 0, SF = 0, SE = 0;
  int ngroups = 0;
  // vertices
  double *vcoord = nullptr;   // [V][3]
  double *vvol = nullptr;     // [V]
  uint8_t *vbound = nullptr;  // [V]
  int *sf_off = nullptr;      // [V+1] sub-faces of v: [sf_off[v], sf_off[v+1])
  // Periodic meshes list a vertex's own sub-faces first, then those of its
  // periodic partners; every per-sub-face array below is then indexed by
  // list entry (a partner's sub-face appears once more, as a copy).
  int *sf_n0 = nullptr;       // [V] periodic: the vertex's own sub-faces (else null)
  int *vx_seoff = nullptr, *vx_se = nullptr;  // [V+1] periodic: partners' sub-elements
  int *se_off = nullptr;      // [V+1] node-sorted sub-element slots of v
  int *v_eoff = nullptr, *v_elem = nullptr;  // vertex -> cells (reference order)
  int *v_seid = nullptr;      // [SE] node-sorted slot -> cell-major sub-element id
  double *Bp = nullptr;       // [3][V] wall normal (depends on the BC set)
  int *grp = nullptr;         // [ngroups][kGrpHdr] vertex-group headers
  uint32_t *se_src = nullptr; // [SE] node-sorted slot -> its contributing group-local
                              // sub-faces, one byte each, increasing, 0xFF = none
  // faces
  int *f_voff = nullptr, *f_vert = nullptr, *f_left = nullptr, *f_right = nullptr;
  int *f_sfoff = nullptr, *f_sf = nullptr;
  double *fcoord = nullptr;   // [F][3]
  double *fnorm = nullptr;    // [F][3]
  double *farea = nullptr;    // [F]
  // sub-faces (vertex-major)
  int *sf_vert = nullptr, *sf_face = nullptr;
  int *sf_le = nullptr;       // left cell
  int *sf_re = nullptr;       // right code: >=0 cell, -1 wall, <=-2 BC (-code-1)
  double *sfn = nullptr;      // [4][SF]: nx, ny, nz, area
  // sub-elements
  int *se_sfoff = nullptr, *se_sf = nullptr, *se_elem = nullptr;  // cell-major ids
  double *se_vol = nullptr;   // [SE] cell-major
  int *se_cell = nullptr;     // [SE] node-sorted slot -> cell
  // cells
  int *e_voff = nullptr, *e_vert = nullptr, *e_foff = nullptr, *e_face = nullptr;
  int *c_slot = nullptr;      // [SE] cell-major: the cell's node-sorted slots in increasing vertex id
  double *ccoord = nullptr;   // [C][3]
  double *cvol = nullptr;     // [C]
};

}  // namespace sfv
