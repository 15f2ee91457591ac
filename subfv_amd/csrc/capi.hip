// C-ABI of libsubfv_hip: mesh upload, configuration, device-resident state
// and the stage orchestration of the explicit residual path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/subfv.h"
#include "device_mesh.hpp"
#include "kernels_batch.hpp"
#include "kernels_bcsr.hpp"
#include "kernels_cell.hpp"
#include "kernels_diff.hpp"
#include "kernels_flux.hpp"
#include "kernels_geom.hpp"
#include "kernels_sweeps.hpp"
#include "mesh_build.hpp"

using namespace sfv;

namespace sfv {
// flux_fast.hip: the node sweep compiled for SUBFV_ARITH_FAST
void launch_node_flux_fast(int scheme, unsigned grid, hipStream_t s, const FluxArgs &a);
}

#include <mutex>

// Workspace of subfv_bcsr_gmres_diag_noalloc: the Krylov basis V and the
// Hessenberg matrix H of one GMRES call live in this file-scope device array
// instead of memory the library allocates.  Each GPU holds its own copy for
// the life of the process; host code must use the per-device address that
// hipGetSymbolAddress returns (noalloc_ptr), never the array's name.
constexpr size_t kNoAllocDoubles = (size_t)4 << 20;  // 32 MiB
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex noalloc_mu;
double *noalloc_ptr[kMaxDevices] = {};

// once per device, with that device current
hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (noalloc_ptr[device]) return hipSuccess;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) noalloc_ptr[device] = (double *)p;
  return e;
}

double *noalloc_get(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  return noalloc_ptr[device];
}
}  // namespace

// Device-only tables derived from a host mesh (right-neighbour codes, slot
// cells, the cells' slots in vertex order, vertex groups and their slot
// sources), built once with the host mesh so an upload only copies.
struct DevTables {
  std::vector<int> sf_re, se_cell, c_slot, grp;
  std::vector<uint32_t> se_src;
  int ngroups = 0;
  std::string err;
  // periodic meshes: the per-sub-face arrays by list entry (sf_re above is
  // then by entry too), the face and sub-element sub-face lists as entries,
  // each sub-face's first entry and each vertex's own sub-face count
  std::vector<int> ent_vert, ent_face, ent_le, ent_fsf, ent_sesf, prim, n0;
};

struct subfv_host_mesh {
  std::shared_ptr<subfv::HostMesh> hm;
  std::shared_ptr<DevTables> tab;
};

// A batch of contexts advanced together by multi-instance launches.
struct subfv_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // members with the generation each had when the tables were built
  std::vector<std::pair<subfv_ctx *, uint64_t>> members;
  int scheme = 0, total_g = 0, total_c = 0;
  sfv::InstDesc *d_desc = nullptr;
  int *d_gtab = nullptr, *d_ctab = nullptr;
  size_t cap_desc = 0, cap_gtab = 0, cap_ctab = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_ok = false;
  // persistent form: workgroups per CU (0: as many as fit) and the
  // queue / arrival counters of one launch
  int wgs_per_cu = 0;
  unsigned *d_q = nullptr;
  size_t cap_q = 0;
};

struct subfv_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // recorded on the stream when a batch takes the context, so the batch's
  // stream waits for the context's pending uploads, state puts and inits
  hipEvent_t ev = nullptr;
  // bumped whenever the mesh, configuration, time or state changes; a batch
  // rebuilds its tables and its captured graph when a member's differs
  uint64_t gen = 0;
  std::string err;
  bool has_mesh = false;
  bool b2d = false;
  int arith = SUBFV_ARITH_EXACT;  // the node sweep's arithmetic (subfv_set_arith)
  std::shared_ptr<const subfv::HostMesh> hmp;  // shared with the host mesh handle it came from
  std::shared_ptr<const DevTables> tab;
  DevMesh dm;
  BCParams bc{};
  subfv_config cfg{};
  // mesh-lifetime device buffers: bump allocations in device chunks that
  // stay allocated when the mesh changes (an upload re-uses them)
  std::vector<std::pair<char *, size_t>> chunks;
  size_t chunk_i = 0, chunk_off = 0;
  // pinned host blocks that stage a mesh upload's copies (reused per upload)
  std::vector<std::pair<char *, size_t>> pins;
  size_t pin_i = 0, pin_off = 0;
  // state and work arrays
  double *U[3] = {nullptr, nullptr, nullptr};
  int cur = 0;
  double *P = nullptr, *P2 = nullptr;
  double *rhs = nullptr, *rhs2 = nullptr, *sl = nullptr, *sl2 = nullptr;
  double *NG = nullptr, *VMM = nullptr, *G = nullptr, *vp = nullptr, *diag = nullptr;
  double *NGP = nullptr;  // [V][9] cached nodal-gradient pseudo-inverses (geometry only)
  double *W = nullptr;    // [C][6] primitives and sound speed of the state a residual works on
  bool ngp_ready = false;
  double *blk = nullptr;
  // viscous terms: per-vertex precomputed blocks (once per mesh with a
  // constant viscosity, every residual with Sutherland's law)
  bool diff_on = false, diff_ready = false;
  DiffParams dparams{};
  long *doff = nullptr;
  double *pre = nullptr, *PD = nullptr, *rdiff = nullptr;
  uint32_t *dssf = nullptr;  // [SE] node-sorted slot -> its sub-element's vertex-local sub-faces, packed
  double *dsvol = nullptr;   // [SE] node-sorted slot -> sub-element volume
  double *mu_cell = nullptr, *mu_blk = nullptr;  // Sutherland: [C] viscosities; block minima + mu0corr
  int diff_fm = 0, diff_em = 0;  // setup LDS extents: 3 * max sub-faces, 3 * max sub-elements of a vertex
  bool diff_large = false;        // a vertex beyond the small size class: large-class kernels
  double *dws = nullptr;          // large class: the setup's workspace, kDiffWsBlocks blocks
  long pre_n = 0;  // doubles in the packed viscous blocks
  int *derr = nullptr;  // [2]: error bits of the path, failed check_mesh cells
  StepState *st = nullptr;
  double host_t = 0.0;
  // optional per-phase kernel timing with HIP events on the context stream
  bool prof = false;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> prof_ev;
};

namespace {

int fail(subfv_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, expr)                                                       \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      return fail((c), SUBFV_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Mesh-lifetime buffer of n elements, 256-byte aligned, taken from the
// context's chunks (a new chunk of at least 16 MiB when none has room).
template <class T>
int dalloc(subfv_ctx *c, T **p, size_t n) {
  const size_t bytes = (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~(size_t)255;
  while (c->chunk_i < c->chunks.size() && c->chunk_off + bytes > c->chunks[c->chunk_i].second) {
    c->chunk_i++;
    c->chunk_off = 0;
  }
  if (c->chunk_i == c->chunks.size()) {
    const size_t sz = std::max<size_t>(bytes, (size_t)16 << 20);
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, sz);
    if (e != hipSuccess) return fail(c, SUBFV_E_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->chunks.push_back({(char *)q, sz});
    c->chunk_off = 0;
  }
  *p = (T *)(c->chunks[c->chunk_i].first + c->chunk_off);
  c->chunk_off += bytes;
  return 0;
}

template <class T>
int upload(subfv_ctx *c, T **p, const std::vector<T> &v) {
  int r = dalloc(c, p, v.size());
  if (r) return r;
  if (!v.empty()) HIPCHK(c, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Host-to-device copy queued on the context's stream from its pinned
// staging blocks (a plain memcpy into pinned memory, then an asynchronous
// DMA copy); the blocks are reused from the start by the next upload, after
// the synchronisation that ends a mesh upload.
int stage_copy(subfv_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!bytes) return 0;
  const size_t need = (bytes + 255) & ~(size_t)255;
  while (c->pin_i < c->pins.size() && c->pin_off + need > c->pins[c->pin_i].second) {
    c->pin_i++;
    c->pin_off = 0;
  }
  if (c->pin_i == c->pins.size()) {
    const size_t sz = std::max(need, (size_t)1 << 20);
    void *q = nullptr;
    HIPCHK(c, hipHostMalloc(&q, sz, hipHostMallocDefault));
    c->pins.push_back({(char *)q, sz});
    c->pin_off = 0;
  }
  char *h = c->pins[c->pin_i].first + c->pin_off;
  std::memcpy(h, src, bytes);
  c->pin_off += need;
  HIPCHK(c, hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, c->stream));
  return 0;
}

template <class T>
int upload_staged(subfv_ctx *c, T **p, const std::vector<T> &v) {
  int r = dalloc(c, p, v.size());
  if (r) return r;
  return stage_copy(c, *p, v.data(), v.size() * sizeof(T));
}

// Forget the mesh; its chunks stay allocated for the next one, so the work
// already queued on the stream must finish first.
void free_mesh(subfv_ctx *c) {
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->chunk_i = 0;
  c->chunk_off = 0;
  c->dm = DevMesh();
  c->has_mesh = false;
  for (auto &u : c->U) u = nullptr;
  c->P = c->P2 = c->rhs = c->rhs2 = c->sl = c->sl2 = nullptr;
  c->NG = c->VMM = c->G = c->vp = c->diag = c->blk = c->NGP = c->W = nullptr;
  c->ngp_ready = false;
  c->doff = nullptr;
  c->pre = c->PD = c->rdiff = c->dsvol = nullptr;
  c->dssf = nullptr;
  c->mu_cell = c->mu_blk = nullptr;
  c->dws = nullptr;
  c->diff_ready = false;
}

inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

// HIP events around one phase when profiling is on (recorded on the stream
// the kernels run on; read back by subfv_profile_read)
struct Prof {
  subfv_ctx *c;
  hipEvent_t a = nullptr, b = nullptr;
  const char *name;
  Prof(subfv_ctx *ctx, const char *nm) : c(ctx), name(nm) {
    if (!c->prof) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      a = b = nullptr;
      return;
    }
    (void)hipEventRecord(a, c->stream);
  }
  ~Prof() {
    if (!a) return;
    (void)hipEventRecord(b, c->stream);
    c->prof_ev.push_back({name, {a, b}});
  }
};

int check_err(subfv_ctx *c) {
  int e = 0;
  HIPCHK(c, hipMemcpyAsync(&e, c->derr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  StepState hs;
  HIPCHK(c, hipMemcpyAsync(&hs, c->st, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int bits = e | hs.err;
  if (bits & kErrNegStarDensity) return fail(c, SUBFV_E_NEG_STAR, "Negative specific volume MPCC !");
  if (bits & kErrBadOmega) return fail(c, SUBFV_E_BAD_OMEGA, "BAD OMEGA " + std::to_string(hs.omega_min));
  if (bits & kErrBadBC) return fail(c, SUBFV_E_BAD_BC, "BC TYPE NOT RECOGNIZED !");
  return 0;
}

// vertex groups of the node sweep (<= kGrpVert vertices, <= kGrpSub sub-faces,
// <= kGrpSlots slots) and, per node-sorted slot, the group-local indices of
// the sub-faces of its vertex that touch its cell, in increasing order (the
// order compute_rhs_around_vert accumulates them, ns_euler_module.F90:506-517)
std::string pack_groups(const subfv::HostMesh &hm, const std::vector<int> &off, const std::vector<int> &ent,
                        const std::vector<int> &secell, std::vector<int> &grp, std::vector<uint32_t> &src) {
  grp.clear();
  src.assign(hm.nse, 0xFFFFFFFFu);
  int v = 0;
  while (v < hm.nv) {
    const int vb = v, sf0 = off[v], se0 = hm.v_seoff[v];
    while (v < hm.nv && v - vb < kGrpVert && off[v + 1] - sf0 <= kGrpSub && hm.v_seoff[v + 1] - se0 <= kGrpSlots)
      v++;
    if (v == vb) return "vertex " + std::to_string(v + 1) + " has more than " + std::to_string(kGrpSub) +
                        " sub-faces or " + std::to_string(kGrpSlots) + " sub-elements";
    const int hdr[kGrpHdr] = {vb, v - vb, sf0, off[v] - sf0, se0, hm.v_seoff[v] - se0, 0, 0};
    grp.insert(grp.end(), hdr, hdr + kGrpHdr);
    for (int u = vb; u < v; u++)
      for (int q = hm.v_seoff[u]; q < hm.v_seoff[u + 1]; q++) {
        const int cell = secell[q];
        int cnt = 0;
        uint32_t code = 0xFFFFFFFFu;
        for (int e = off[u]; e < off[u + 1]; e++) {
          const int s = ent.empty() ? e : ent[e];
          // sums go to the vertex's own sub-elements only, the right one only
          // with the left (compute_rhs_around_vert, ns_euler_module.F90:506-517)
          const int lse = hm.sf_lse[s], rse = hm.sf_rse[s];
          if (lse < 0 || hm.se_vert[lse] != u) continue;
          const bool left = hm.sf_le[s] == cell;
          const bool right = hm.sf_re[s] - 1 == cell && rse >= 0 && hm.se_vert[rse] == u;
          if (left || right) {
            if (cnt == kSlotSrc)
              return "a cell has more than " + std::to_string(kSlotSrc) + " sub-faces at vertex " +
                     std::to_string(u + 1);
            code &= ~(0xFFu << (8 * cnt));
            code |= (uint32_t)(e - sf0) << (8 * cnt);
            cnt++;
          }
        }
        src[q] = code;
      }
  }
  return "";
}

// the device-only tables of a host mesh (see DevTables)
std::shared_ptr<DevTables> make_tables(const subfv::HostMesh &h) {
  auto t = std::make_shared<DevTables>();
  t->sf_re.resize(h.nsf);
  for (int s = 0; s < h.nsf; s++) t->sf_re[s] = h.sf_re[s] - 1;  // device right-neighbour code
  t->se_cell.resize(h.nse);
  for (int q = 0; q < h.nse; q++) t->se_cell[q] = h.se_elem[h.v_se[q]];
  // each cell's sub-elements in increasing vertex id (the scatter order of
  // compute_euler_flux, ns_euler_module.F90:321-336), as node-sorted slots:
  // the gather reads the node sweep's per-slot partials through this table
  std::vector<int> slot_of(h.nse);
  for (int q = 0; q < h.nse; q++) slot_of[h.v_se[q]] = q;
  t->c_slot.resize(h.nse);
  for (int e = 0; e < h.ne; e++) {
    int b = h.e_voff[e], n = h.e_voff[e + 1] - b;
    std::pair<int, int> pv[8];
    for (int j = 0; j < n; j++) pv[j] = {h.se_vert[b + j], slot_of[b + j]};
    std::sort(pv, pv + n);
    for (int j = 0; j < n; j++) t->c_slot[b + j] = pv[j].second;
  }
  if (h.periodic) {
    const int ne_ = h.x_sfoff[h.nv];
    t->prim.resize(h.nsf);
    t->n0.resize(h.nv);
    for (int v = 0; v < h.nv; v++) {
      t->n0[v] = h.sf_off[v + 1] - h.sf_off[v];
      for (int s = h.sf_off[v]; s < h.sf_off[v + 1]; s++) t->prim[s] = h.x_sfoff[v] + (s - h.sf_off[v]);
    }
    t->ent_vert.resize(ne_);
    t->ent_face.resize(ne_);
    t->ent_le.resize(ne_);
    t->sf_re.resize(ne_);
    for (int e = 0; e < ne_; e++) {
      const int s = h.x_sf[e];
      t->ent_vert[e] = h.sf_vert[s];
      t->ent_face[e] = h.sf_face[s];
      t->ent_le[e] = h.sf_le[s];
      t->sf_re[e] = h.sf_re[s] - 1;
    }
    t->ent_fsf.resize(h.f_sf.size());
    for (size_t q = 0; q < h.f_sf.size(); q++) t->ent_fsf[q] = t->prim[h.f_sf[q]];
    t->ent_sesf.resize(h.se_sf.size());
    for (size_t q = 0; q < h.se_sf.size(); q++) t->ent_sesf[q] = t->prim[h.se_sf[q]];
    t->err = pack_groups(h, h.x_sfoff, h.x_sf, t->se_cell, t->grp, t->se_src);
  } else {
    t->err = pack_groups(h, h.sf_off, {}, t->se_cell, t->grp, t->se_src);
  }
  t->ngroups = (int)(t->grp.size() / kGrpHdr);
  return t;
}

int upload_mesh(subfv_ctx *c) {
  const subfv::HostMesh &h = *c->hmp;
  const DevTables &t = *c->tab;
  if (!t.err.empty()) return fail(c, SUBFV_E_INVALID, t.err);
  DevMesh &d = c->dm;
  const bool per = h.periodic;
  c->pin_i = 0;  // the previous upload's copies finished (free_mesh synchronised)
  c->pin_off = 0;
  d.V = h.nv;
  d.C = h.ne;
  d.F = h.nf;
  d.SF = per ? h.x_sfoff[h.nv] : h.nsf;  // list entries on a periodic mesh
  d.SE = h.nse;
  int r;
#define UP(field, vec) \
  if ((r = upload_staged(c, &d.field, vec))) return r
  UP(vcoord, h.vcoord);
  UP(sf_off, per ? h.x_sfoff : h.sf_off);
  if (per) {
    UP(sf_n0, t.n0);
    UP(vx_seoff, h.x_seoff);
    UP(vx_se, h.x_se);
  } else {
    d.sf_n0 = d.vx_seoff = d.vx_se = nullptr;
  }
  UP(se_off, h.v_seoff);
  UP(v_eoff, h.v_eoff);
  UP(v_elem, h.v_elem);
  UP(v_seid, h.v_se);
  UP(f_voff, h.f_voff);
  UP(f_vert, h.f_vert);
  UP(f_left, h.f_left);
  UP(f_right, h.f_right);
  UP(f_sfoff, h.f_sfoff);
  UP(f_sf, per ? t.ent_fsf : h.f_sf);
  UP(sf_vert, per ? t.ent_vert : h.sf_vert);
  UP(sf_face, per ? t.ent_face : h.sf_face);
  UP(sf_le, per ? t.ent_le : h.sf_le);
  UP(sf_re, t.sf_re);
  UP(se_sfoff, h.se_sfoff);
  UP(se_sf, per ? t.ent_sesf : h.se_sf);
  UP(se_elem, h.se_elem);
  UP(se_cell, t.se_cell);
  UP(e_voff, h.e_voff);
  UP(e_vert, h.e_vert);
  UP(e_foff, h.e_foff);
  UP(e_face, h.e_face);
  UP(c_slot, t.c_slot);
  d.ngroups = t.ngroups;
  UP(grp, t.grp);
  UP(se_src, t.se_src);
#undef UP
  if ((r = dalloc(c, &d.vvol, h.nv))) return r;
  if ((r = dalloc(c, &d.vbound, h.nv))) return r;
  if ((r = dalloc(c, &d.Bp, 3L * h.nv))) return r;
  if ((r = dalloc(c, &d.fcoord, 3L * h.nf))) return r;
  if ((r = dalloc(c, &d.fnorm, 3L * h.nf))) return r;
  if ((r = dalloc(c, &d.farea, h.nf))) return r;
  if ((r = dalloc(c, &d.sfn, 4L * d.SF))) return r;
  if ((r = dalloc(c, &d.se_vol, h.nse))) return r;
  if ((r = dalloc(c, &d.ccoord, 3L * h.ne))) return r;
  if ((r = dalloc(c, &d.cvol, h.ne))) return r;
  // state and work arrays
  const long C = h.ne, V = h.nv, SE = h.nse;
  for (int i = 0; i < 3; i++)
    if ((r = dalloc(c, &c->U[i], 5 * C))) return r;
  if ((r = dalloc(c, &c->P, (size_t)kSlotStride * SE))) return r;
  if ((r = dalloc(c, &c->P2, (size_t)kSlotStride * SE))) return r;
  if ((r = dalloc(c, &c->rhs, 5 * C))) return r;
  if ((r = dalloc(c, &c->rhs2, 5 * C))) return r;
  if ((r = dalloc(c, &c->sl, C))) return r;
  if ((r = dalloc(c, &c->sl2, C))) return r;
  if ((r = dalloc(c, &c->NG, 15 * V))) return r;
  if ((r = dalloc(c, &c->VMM, 4 * V))) return r;
  if ((r = dalloc(c, &c->G, 15 * C))) return r;
  if ((r = dalloc(c, &c->vp, 3 * V))) return r;
  if ((r = dalloc(c, &c->blk, nblk(std::max(C, V), kCellThreads) + 64))) return r;
  HIPCHK(c, hipMemsetAsync(c->G, 0, sizeof(double) * 15 * C, c->stream));
  HIPCHK(c, hipMemsetAsync(c->NG, 0, sizeof(double) * 15 * V, c->stream));
  HIPCHK(c, hipMemsetAsync(c->vp, 0, sizeof(double) * 3 * V, c->stream));
  HIPCHK(c, hipMemsetAsync(c->U[0], 0, sizeof(double) * 5 * C, c->stream));
  c->cur = 0;
  return 0;
}

int geometry(subfv_ctx *c) {
  DevMesh &d = c->dm;
  const int T = 256;
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_face_centroid, dim3(nblk(d.F, T)), dim3(T), 0, s, d);
  hipLaunchKernelGGL(k_elem_centroid, dim3(nblk(d.C, T)), dim3(T), 0, s, d);
  hipLaunchKernelGGL(k_subface_norm, dim3(nblk(d.SF, T)), dim3(T), 0, s, d);
  hipLaunchKernelGGL(k_face_norm, dim3(nblk(d.F, T)), dim3(T), 0, s, d);
  hipLaunchKernelGGL(k_subelem_volume, dim3(nblk(d.SE, T)), dim3(T), 0, s, d);
  int *bad = c->derr + 1;
  HIPCHK(c, hipMemsetAsync(bad, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_elem_volume_check, dim3(nblk(d.C, T)), dim3(T), 0, s, d, bad);
  hipLaunchKernelGGL(k_vert_volume_bound, dim3(nblk(d.V, T)), dim3(T), 0, s, d, (int)c->b2d);
  HIPCHK(c, hipGetLastError());
  int hb = 0;
  HIPCHK(c, hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (hb) return fail(c, SUBFV_E_MESH, "MESH GEOMETRY IS NOT OK (" + std::to_string(hb) + " cells)");
  return 0;
}

int apply_bc(subfv_ctx *c) {
  if (!c->has_mesh) return 0;
  DevMesh &d = c->dm;
  hipLaunchKernelGGL(k_wall_normal, dim3(nblk(d.V, 256)), dim3(256), 0, c->stream, d, c->bc);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// upload of a built host connectivity, device geometry and BC data
int finish_mesh(subfv_ctx *c, int32_t boundary_2d) {
  c->gen++;
  c->b2d = boundary_2d != 0;
  int r = upload_mesh(c);
  if (r) return r;
  r = geometry(c);
  if (r) return r;
  c->has_mesh = true;
  c->diff_ready = false;
  r = apply_bc(c);
  if (r) return r;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // the copies are done: staging blocks beyond the first are returned, so
  // pinned host memory held by idle contexts stays at one block each (a
  // server keeps hundreds of contexts)
  while (c->pins.size() > 1) {
    (void)hipHostFree(c->pins.back().first);
    c->pins.pop_back();
  }
  return 0;
}

// --------------------------------------------------------------- launches
// c->W = the primitives and sound speeds of U, for the sweeps and gradients
// of one residual
int launch_prim(subfv_ctx *c, const double *U) {
  if (!c->W) {
    int r = dalloc(c, &c->W, 6 * (size_t)c->dm.C);
    if (r) return r;
  }
  hipLaunchKernelGGL(k_prim, dim3(nblk(c->dm.C, 256)), dim3(256), 0, c->stream, (long)c->dm.C, U, c->W);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// the node sweep on U; W: its primitives from launch_prim, or null
int launch_flux(subfv_ctx *c, const double *U, const double *G, double *P, bool want_vp, const double *W = nullptr) {
  FluxArgs a;
  a.m = c->dm;
  a.bc = c->bc;
  a.U = U;
  a.G = G;
  a.W = W;
  a.P = P;
  a.vp = want_vp ? c->vp : nullptr;
  a.err = c->derr;
  a.t = &c->st->t;
  a.second_order = G != nullptr;
  const int nb = c->dm.ngroups;
  a.per_xcd = (nb + 7) / 8;
  a.nblocks = nb;
  dim3 grid(8 * a.per_xcd), blk(kGrpThreads);
  Prof pf(c, G ? "node_flux_o2" : "node_flux");
  if (c->arith == SUBFV_ARITH_FAST) {
    launch_node_flux_fast(c->bc.scheme, grid.x, c->stream, a);
    HIPCHK(c, hipGetLastError());
    return 0;
  }
  switch (c->bc.scheme) {
    case kMultiPoint: hipLaunchKernelGGL(k_node_flux<kMultiPoint>, grid, blk, 0, c->stream, a); break;
    case kMultiPointIso: hipLaunchKernelGGL(k_node_flux<kMultiPointIso>, grid, blk, 0, c->stream, a); break;
    case kThreeWave: hipLaunchKernelGGL(k_node_flux<kThreeWave>, grid, blk, 0, c->stream, a); break;
    case kTwoWave: hipLaunchKernelGGL(k_node_flux<kTwoWave>, grid, blk, 0, c->stream, a); break;
    case kModifiedThreeWave: hipLaunchKernelGGL(k_node_flux<kModifiedThreeWave>, grid, blk, 0, c->stream, a); break;
    case kMultiPointPressure: hipLaunchKernelGGL(k_node_flux<kMultiPointPressure>, grid, blk, 0, c->stream, a); break;
    case kMultiPointPressurePh: hipLaunchKernelGGL(k_node_flux<kMultiPointPressurePh>, grid, blk, 0, c->stream, a); break;
    default: return fail(c, SUBFV_E_INVALID, "scheme not supported on the device");
  }
  HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_gather(subfv_ctx *c, const double *P, const double *P2, double *rhs, double *rhs2, double *sl,
                  int fixed_dt) {
  GatherArgs g;
  g.m = c->dm;
  g.P = P;
  g.P2 = P2;
  g.rhs = rhs;
  g.rhs2 = rhs2;
  g.sl = sl;
  g.blk = c->blk;
  g.st = c->st;
  g.fixed_dt = fixed_dt;
  Prof pf(c, "gather");
  const int nb = (int)nblk(c->dm.C, kCellThreads);
  g.nblocks = nb;
  g.per_xcd = (nb + 7) / 8;
  hipLaunchKernelGGL(k_gather, dim3(8 * g.per_xcd), dim3(kCellThreads), 0, c->stream, g);
  hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(kCellThreads), 0, c->stream, (const double *)c->blk, nb, c->st,
                     fixed_dt);
  HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_grad(subfv_ctx *c, const double *U, int method, const double *W = nullptr) {
  GradArgs g;
  g.m = c->dm;
  g.U = U;
  g.W = W;
  g.NG = c->NG;
  g.VMM = c->VMM;
  g.G = c->G;
  g.method = method;
  const int nb = c->dm.ngroups;
  g.per_xcd = (nb + 7) / 8;
  g.nblocks = nb;
  // the geometric pseudo-inverses are formed by the first gradient of a
  // mesh and read back by every later one
  if (!c->NGP) {
    int r = dalloc(c, &c->NGP, 9 * (size_t)c->dm.V);
    if (r) return r;
  }
  g.NGP = c->NGP;
  Prof pf(c, "grad");
  if (c->ngp_ready) {
    hipLaunchKernelGGL(k_nodal_grad<kNGCached>, dim3(8 * g.per_xcd), dim3(kGrpThreads), 0, c->stream, g);
  } else {
    hipLaunchKernelGGL(k_nodal_grad<kNGStore>, dim3(8 * g.per_xcd), dim3(kGrpThreads), 0, c->stream, g);
    c->ngp_ready = true;
  }
  g.cblocks = (int)nblk(c->dm.C, kCellThreads);
  g.cper_xcd = (g.cblocks + 7) / 8;
  if (g.method == 4)
    hipLaunchKernelGGL(k_cell_grad<4>, dim3(8 * g.cper_xcd), dim3(kCellThreads), 0, c->stream, g);
  else
    hipLaunchKernelGGL(k_cell_grad<-1>, dim3(8 * g.cper_xcd), dim3(kCellThreads), 0, c->stream, g);
  HIPCHK(c, hipGetLastError());
  return 0;
}

int launch_update(subfv_ctx *c, const double *U, const double *U0, double *Uout, double a1, double a2, double w0,
                  double w1, int flux_limit, int advance_t, bool with_diff = false) {
  UpdateArgs u;
  u.m = c->dm;
  u.U = U;
  u.U0 = U0;
  u.Uout = Uout;
  u.rhs = c->rhs;
  u.rhs2 = c->rhs2;
  u.sl = c->sl;
  u.diag = c->diff_on ? c->diag : nullptr;
  u.rdiff = (c->diff_on && with_diff) ? c->rdiff : nullptr;
  u.blk = c->blk;
  u.st = c->st;
  u.a1 = a1;
  u.a2 = a2;
  u.w0 = w0;
  u.w1 = w1;
  u.flux_limit = flux_limit;
  u.advance_t = advance_t;
  dim3 g(nblk(c->dm.C, kCellThreads)), b(kCellThreads);
  Prof pf(c, "update");
  if (u.diag) {
    hipLaunchKernelGGL(k_update<true>, g, b, 0, c->stream, u);
    hipLaunchKernelGGL(k_finalize_omega, dim3(1), b, 0, c->stream, (const double *)c->blk, (int)g.x, c->st, advance_t);
    hipLaunchKernelGGL(k_update_fixup<true>, g, b, 0, c->stream, u);
  } else {
    hipLaunchKernelGGL(k_update<false>, g, b, 0, c->stream, u);
    hipLaunchKernelGGL(k_finalize_omega, dim3(1), b, 0, c->stream, (const double *)c->blk, (int)g.x, c->st, advance_t);
    hipLaunchKernelGGL(k_update_fixup<false>, g, b, 0, c->stream, u);
  }
  HIPCHK(c, hipGetLastError());
  return 0;
}

// ------------------------------------------------------------- diffusion
// The setup kernel over every vertex: the small class with its matrices in
// dynamic LDS in one launch; the large class with them in the workspace,
// kDiffWsBlocks vertices per launch.
void launch_diff_setup(subfv_ctx *c, DiffArgs a) {
  a.set_fm = c->diff_fm;
  a.set_em = c->diff_em;
  const long V = c->dm.V;
  if (!c->diff_large) {
    a.v0 = 0;
    hipLaunchKernelGGL((k_diff_setup<kDiffMaxSub, false>), dim3((unsigned)V), dim3(64),
                       diff_setup_lds_bytes(a.set_fm, a.set_em), c->stream, a);
    return;
  }
  a.ws = c->dws;
  a.ws_stride = diff_setup_ws_stride(a.set_fm, a.set_em);
  for (long v0 = 0; v0 < V; v0 += kDiffWsBlocks) {
    a.v0 = (int)v0;
    hipLaunchKernelGGL((k_diff_setup<kDiffMaxSubL, true>), dim3((unsigned)std::min<long>(kDiffWsBlocks, V - v0)),
                       dim3(64), 0, c->stream, a);
  }
}

// Once per mesh and parameter set: offsets of the packed per-vertex blocks,
// buffers, and the setup kernel (nodal systems, inverses, products).
int ensure_diffusion(subfv_ctx *c) {
  if (c->diff_ready) return 0;
  const subfv::HostMesh &h = *c->hmp;
  if (h.periodic) return fail(c, SUBFV_E_INVALID, "viscous terms are not available on periodic meshes");
  std::vector<long> off(h.nv + 1, 0);
  int mf = 1, me = 1;
  for (int v = 0; v < h.nv; v++) {
    const int nf = h.sf_off[v + 1] - h.sf_off[v], ns = h.v_seoff[v + 1] - h.v_seoff[v];
    if (nf > kDiffMaxSubL || ns > kDiffMaxSubL)
      return fail(c, SUBFV_E_INVALID, "viscous terms: vertex " + std::to_string(v + 1) + " has more than " +
                                          std::to_string(kDiffMaxSubL) + " sub-faces or sub-elements");
    off[v + 1] = off[v] + diff_block_size(nf, ns);
    mf = std::max(mf, nf);
    me = std::max(me, ns);
  }
  c->diff_fm = 3 * mf;
  c->diff_em = 3 * me;
  c->diff_large = mf > kDiffMaxSub || me > kDiffMaxSub;
  const size_t lds = diff_setup_lds_bytes(c->diff_fm, c->diff_em);
  if (!c->diff_large && lds > (64u << 10))
    HIPCHK(c, hipFuncSetAttribute((const void *)k_diff_setup<kDiffMaxSub, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int r;
  // the buffers' sizes depend on the mesh only: allocated once per mesh and
  // kept when the viscous parameters change (a new mesh resets them)
  if (!c->doff) {
    size_t free_b = 0, total_b = 0;
    HIPCHK(c, hipMemGetInfo(&free_b, &total_b));
    const size_t need = sizeof(double) * ((size_t)off[h.nv] + (size_t)kDiffSlot * h.nse + 29ull * h.ne);
    if (need + ((size_t)1 << 30) > free_b)
      return fail(c, SUBFV_E_DEVICE, "viscous terms: " + std::to_string(need >> 20) + " MiB needed, " +
                                         std::to_string(free_b >> 20) + " MiB free");
    if ((r = upload(c, &c->doff, off))) return r;
    if ((r = dalloc(c, &c->pre, (size_t)off[h.nv]))) return r;
    c->pre_n = off[h.nv];
    if ((r = dalloc(c, &c->PD, (size_t)kDiffSlot * h.nse))) return r;
    if ((r = dalloc(c, &c->rdiff, 4ull * h.ne))) return r;
    // each slot's sub-element sub-faces as vertex-local indices, one byte
    // each in list order (0xFF ends the list), and its volume: the apply
    // reads them at the slot's index instead of through the cell-major lists
    std::vector<uint32_t> ssf(h.nse, 0xFFFFFFFFu);
    for (int v = 0; v < h.nv; v++)
      for (int q = h.v_seoff[v]; q < h.v_seoff[v + 1]; q++) {
        const int se = h.v_se[q], pb = h.se_sfoff[se], pe = h.se_sfoff[se + 1];
        if (pe - pb > kSlotSrc)
          return fail(c, SUBFV_E_INVALID, "viscous terms: a sub-element has more than " +
                                              std::to_string(kSlotSrc) + " sub-faces");
        for (int pp = pb; pp < pe; pp++) {
          const uint32_t loc = (uint32_t)(h.se_sf[pp] - h.sf_off[v]);
          ssf[q] &= ~(0xFFu << (8 * (pp - pb)));
          ssf[q] |= loc << (8 * (pp - pb));
        }
      }
    if ((r = upload(c, &c->dssf, ssf))) return r;
    if (c->diff_large &&
        (r = dalloc(c, &c->dws, (size_t)kDiffWsBlocks * diff_setup_ws_stride(c->diff_fm, c->diff_em))))
      return r;
    if ((r = dalloc(c, &c->dsvol, (size_t)h.nse))) return r;
    hipLaunchKernelGGL(k_slot_vol, dim3(nblk(h.nse, 256)), dim3(256), 0, c->stream, (long)h.nse,
                       (const int *)c->dm.v_seid, (const double *)c->dm.se_vol, c->dsvol);
  }
  if (c->dparams.suth && !c->mu_cell) {
    if ((r = dalloc(c, &c->mu_cell, (size_t)h.ne))) return r;
    if ((r = dalloc(c, &c->mu_blk, (size_t)nblk(h.ne, kCellThreads) + 1))) return r;
  }
  if (!c->diag && (r = dalloc(c, &c->diag, 25ull * h.ne))) return r;
  HIPCHK(c, hipMemsetAsync(c->diag, 0, sizeof(double) * 25 * h.ne, c->stream));
  HIPCHK(c, hipMemsetAsync(c->rdiff, 0, sizeof(double) * 4 * h.ne, c->stream));
  DiffArgs a{};
  a.m = c->dm;
  a.p = c->dparams;
  a.doff = c->doff;
  a.pre = c->pre;
  a.U = nullptr;
  a.PD = c->PD;
  a.rdiff = c->rdiff;
  a.diag = c->diag;
  // with Sutherland's law the blocks depend on the state: formed per residual
  if (!c->dparams.suth) launch_diff_setup(c, a);
  HIPCHK(c, hipGetLastError());
  c->diff_ready = true;
  return 0;
}

// compute_mu with use_sutherland (ns_vectorial_diffusion_module.F90:322-359):
// mu_sutherland of every cell, C0 = 110.4 (ns_global_data_module.F90:20), and
// the block minima of mu0corr
__global__ __launch_bounds__(kCellThreads) void k_compute_mu(long C, const double *U, double cv, double mu0,
                                                             double T0, double *mu, double *blk) {
  const long c = (long)blockIdx.x * kCellThreads + threadIdx.x;
  double v = 1e12;
  if (c < C) {
    double u[5];
    for (int k = 0; k < 5; k++) u[k] = U[5 * c + k];
    const double T = temp_u(u, cv);
    v = mu0 * pow(T / T0, 1.5) * (T0 + 110.4) / (T + 110.4);
    mu[c] = v;
  }
  block_min_store(v, blk);
}

__global__ __launch_bounds__(kCellThreads) void k_mu_min(const double *blk, int nb, double *out) {
  const double v = reduce_min_one_block(blk, nb);
  if (threadIdx.x == 0) *out = v;
}

// rdiff and diag of the state U (build_momentum_diff_mat_into_global_mat_diag
// and build_diff_mat_into_global_mat_diag).  mu_cell / mu_min: given cell viscosities
// and their mu0corr (device), for which the nodal systems are formed anew
// (the constant-viscosity blocks are then invalidated); parts: 1 momentum,
// 2 heat, 3 both.
int launch_diffusion(subfv_ctx *c, const double *U, const double *mu_cell = nullptr,
                     const double *mu_min = nullptr, int parts = 3) {
  int r = ensure_diffusion(c);
  if (r) return r;
  DiffArgs a{};
  a.parts = parts;
  a.m = c->dm;
  a.p = c->dparams;
  a.doff = c->doff;
  a.pre = c->pre;
  a.U = U;
  a.PD = c->PD;
  a.rdiff = c->rdiff;
  a.diag = c->diag;
  a.slot_ssf = c->dssf;
  a.slot_vol = c->dsvol;
  Prof pf(c, "diffusion");
  if (mu_cell) {
    a.mu_cell = mu_cell;
    a.mu_min = mu_min;
    launch_diff_setup(c, a);
    if (!c->dparams.suth) c->diff_ready = false;  // the cached constant-viscosity blocks are overwritten
  } else if (c->dparams.suth) {
    Prof ps(c, "diff_setup");
    const unsigned nb = nblk(c->dm.C, kCellThreads);
    hipLaunchKernelGGL(k_compute_mu, dim3(nb), dim3(kCellThreads), 0, c->stream, (long)c->dm.C, U, c->dparams.cv,
                       c->dparams.suth_mu0, c->dparams.suth_T0, c->mu_cell, c->mu_blk);
    hipLaunchKernelGGL(k_mu_min, dim3(1), dim3(kCellThreads), 0, c->stream, (const double *)c->mu_blk, (int)nb,
                       c->mu_blk + nb);
    a.mu_cell = c->mu_cell;
    a.mu_min = c->mu_blk + nb;
    launch_diff_setup(c, a);
  }
  const int waves = c->diff_large ? 1 : kDiffWaves;
  a.vblocks = (int)nblk(c->dm.V, waves);
  a.vper_xcd = (a.vblocks + 7) / 8;
  a.cblocks = (int)nblk(c->dm.C, kDgCells);
  a.cper_xcd = (a.cblocks + 7) / 8;
  {
    Prof pa(c, "diff_apply");
    if (c->diff_large)
      hipLaunchKernelGGL((k_diff_apply<kDiffMaxSubL, kDiffMaxSubL, kDiffStage, 1, 7>), dim3(8 * a.vper_xcd),
                         dim3(64), 0, c->stream, a);
    else if (c->diff_em <= 3 * kDiffMaxSubC)
      hipLaunchKernelGGL((k_diff_apply<kDiffMaxSub, kDiffMaxSubC, kDiffStageC, kDiffWaves, 8>),
                         dim3(8 * a.vper_xcd), dim3(64 * kDiffWaves), 0, c->stream, a);
    else
      hipLaunchKernelGGL((k_diff_apply<kDiffMaxSub, kDiffMaxSub, kDiffStage, kDiffWaves, 7>), dim3(8 * a.vper_xcd),
                         dim3(64 * kDiffWaves), 0, c->stream, a);
  }
  hipLaunchKernelGGL(k_diff_gather, dim3(8 * a.cper_xcd), dim3(256), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  return 0;
}

__global__ void k_add_rdiff(long C, double *rhs, const double *rdiff) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  for (int k = 1; k < 5; k++) rhs[k * C + c] = rhs[k * C + c] + rdiff[4 * c + k - 1];
}

int sync_state_scalars(subfv_ctx *c, StepState *hs) {
  HIPCHK(c, hipMemcpyAsync(hs, c->st, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int push_state_scalars(subfv_ctx *c, double t) {
  StepState hs{};
  hs.t = t;
  hs.tmax = c->cfg.tmax;
  hs.cfl = c->cfg.cfl;
  hs.lts = c->cfg.local_time_step;
  hs.omega = 1.0;
  HIPCHK(c, hipMemcpyAsync(c->st, &hs, sizeof(StepState), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->derr, 0, sizeof(int), c->stream));
  return 0;
}

int mesh_i32(const subfv::HostMesh &h, const char *name, int32_t *out, int64_t cap) {
  std::vector<int> v;
  std::string n(name);
  auto list = [&](const std::vector<int> &off, const std::vector<int> &ids, int shift) {
    for (size_t i = 0; i < ids.size(); i++) v.push_back(ids[i] + shift);
    (void)off;
  };
  if (n == "vert_nto") list({}, h.vert_nto, 1);
  else if (n == "elem_nto") list({}, h.elem_nto, 1);
  else if (n == "vert_n_sub_faces" && h.periodic)
    for (int i = 0; i < h.nv; i++) v.push_back(h.x_sfoff[i + 1] - h.x_sfoff[i]);
  else if (n == "vert_n_sub_elems" && h.periodic)
    for (int i = 0; i < h.nv; i++)
      v.push_back(h.v_seoff[i + 1] - h.v_seoff[i] + h.x_seoff[i + 1] - h.x_seoff[i]);
  else if (n == "vert_sub_face" && h.periodic) list({}, h.x_sf, 1);
  else if (n == "vert_sub_elem" && h.periodic)
    for (int i = 0; i < h.nv; i++) {
      for (int q = h.v_seoff[i]; q < h.v_seoff[i + 1]; q++) v.push_back(h.v_se[q] + 1);
      for (int q = h.x_seoff[i]; q < h.x_seoff[i + 1]; q++) v.push_back(h.x_se[q] + 1);
    }
  else if (n == "vert_n_sub_faces") for (int i = 0; i < h.nv; i++) v.push_back(h.sf_off[i + 1] - h.sf_off[i]);
  else if (n == "vert_n_sub_elems") for (int i = 0; i < h.nv; i++) v.push_back(h.v_seoff[i + 1] - h.v_seoff[i]);
  else if (n == "vert_n_elems") for (int i = 0; i < h.nv; i++) v.push_back(h.v_eoff[i + 1] - h.v_eoff[i]);
  else if (n == "vert_sub_face") for (int i = 0; i < h.nsf; i++) v.push_back(i + 1);
  else if (n == "vert_sub_elem") list({}, h.v_se, 1);
  else if (n == "vert_elem") list({}, h.v_elem, 1);
  else if (n == "elem_n_vert") for (int i = 0; i < h.ne; i++) v.push_back(h.e_voff[i + 1] - h.e_voff[i]);
  else if (n == "elem_kind") list({}, h.e_kind, 0);
  else if (n == "elem_vert") list({}, h.e_vert, 1);
  else if (n == "elem_face") list({}, h.e_face, 1);
  else if (n == "elem_sub_elem") for (int i = 0; i < h.nse; i++) v.push_back(i + 1);
  else if (n == "elem_tag") list({}, h.e_tag, 0);
  else if (n == "elem_n_neigh_by_vert") for (int i = 0; i < h.ne; i++) v.push_back(h.e_nbvoff[i + 1] - h.e_nbvoff[i]);
  else if (n == "elem_neigh_by_vert") list({}, h.e_nbv, 1);
  else if (n == "face_n_vert") for (int i = 0; i < h.nf; i++) v.push_back(h.f_voff[i + 1] - h.f_voff[i]);
  else if (n == "face_vert") list({}, h.f_vert, 1);
  else if (n == "face_left") list({}, h.f_left, 1);
  else if (n == "face_right") list({}, h.f_right, 0);
  else if (n == "face_sub_face") list({}, h.f_sf, 1);
  else if (n == "sub_face_vert") list({}, h.sf_vert, 1);
  else if (n == "sub_face_face") list({}, h.sf_face, 1);
  else if (n == "sub_face_loc") for (int i = 0; i < h.nv; i++) for (int s = h.sf_off[i]; s < h.sf_off[i + 1]; s++) v.push_back(s - h.sf_off[i] + 1);
  else if (n == "sub_face_left_elem") list({}, h.sf_le, 1);
  else if (n == "sub_face_right_elem") list({}, h.sf_re, 0);
  else if (n == "sub_face_left_sub_elem") list({}, h.sf_lse, 1);
  else if (n == "sub_face_right_sub_elem") list({}, h.sf_rse, 1);
  else if (n == "sub_elem_loc") list({}, h.se_loc, 1);
  else if (n == "sub_elem_vert") list({}, h.se_vert, 1);
  else if (n == "sub_elem_elem") list({}, h.se_elem, 1);
  else if (n == "sub_elem_sub_face") list({}, h.se_sf, 1);
  else return -1;
  if ((int64_t)v.size() > cap) return -1;
  std::copy(v.begin(), v.end(), out);
  return (int64_t)v.size();
}

int dev_to_host_soa(subfv_ctx *c, const double *d, int ncomp, long n, double *out) {
  std::vector<double> tmp((size_t)ncomp * n);
  HIPCHK(c, hipMemcpyAsync(tmp.data(), d, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (long i = 0; i < n; i++)
    for (int k = 0; k < ncomp; k++) out[(size_t)i * ncomp + k] = tmp[(size_t)k * n + i];
  return 0;
}

int host_to_dev_soa(subfv_ctx *c, const double *in, int ncomp, long n, double *d) {
  std::vector<double> tmp((size_t)ncomp * n);
  for (long i = 0; i < n; i++)
    for (int k = 0; k < ncomp; k++) tmp[(size_t)k * n + i] = in[(size_t)i * ncomp + k];
  HIPCHK(c, hipMemcpyAsync(d, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// dt from the device sum_lambda alone (standalone add_timedisc_and_solve)
__global__ __launch_bounds__(kCellThreads) void k_dt_from_sl(DevMesh m, const double *sl, double *blk) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  block_min_store((c < m.C) ? m.cvol[c] / sl[c] : 1e10, blk);
}

// flux_limiting as a standalone pass (ns_euler_recon_module.F90:586-593)
__global__ __launch_bounds__(kCellThreads) void k_flux_limiting(DevMesh m, const double *U, double *rhs,
                                                                const double *rhs2, const double *sl,
                                                                const StepState *st) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.C) return;
  const long C = m.C;
  double cf = m.cvol[c] / st->dt_fl + sl[c];
  double x[5], y[5], r[5], r2[5];
  for (int k = 0; k < 5; k++) {
    r[k] = rhs[k * C + c];
    r2[k] = rhs2[k * C + c];
    x[k] = U[5 * c + k] + r[k] / cf;
    y[k] = (r2[k] - r[k]) / cf;
  }
  double om = max_omega_pos(x, y);
  for (int k = 0; k < 5; k++) rhs[k * C + c] = r[k] + dmin(1.0, om) * (r2[k] - r[k]);
}

}  // namespace

// ======================================================================= C-ABI
extern "C" {

int subfv_create(subfv_ctx **out, int device) {
  if (!out) return SUBFV_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SUBFV_E_DEVICE;
  if (device < 0 || device >= n) return SUBFV_E_INVALID;
  subfv_ctx *c = new subfv_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void **)&c->derr, 2 * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&c->st, sizeof(StepState)) != hipSuccess) {
    delete c;
    return SUBFV_E_DEVICE;
  }
  if (noalloc_setup(device) != hipSuccess) {
    subfv_destroy(c);
    return SUBFV_E_DEVICE;
  }
  (void)hipMemset(c->derr, 0, 2 * sizeof(int));
  (void)hipMemset(c->st, 0, sizeof(StepState));
  // the first device chunk and the upload staging hold a mesh of up to about
  // 10^4 cells, so a context admits such a mesh without allocating; larger
  // meshes add chunks and staging blocks
  void *q = nullptr;
  if (hipMalloc(&q, (size_t)48 << 20) == hipSuccess) c->chunks.push_back({(char *)q, (size_t)48 << 20});
  q = nullptr;
  if (hipHostMalloc(&q, (size_t)8 << 20, hipHostMallocDefault) == hipSuccess) c->pins.push_back({(char *)q, (size_t)8 << 20});
  c->cfg.scheme_id = SUBFV_SCHEME_MULTI_POINT;
  c->cfg.bc_style = 2;
  c->cfg.method = 4;
  c->cfg.cfl = 0.4;
  c->cfg.tmax = 1e30;
  c->bc.scheme = kMultiPoint;
  c->bc.bc_style = 2;
  *out = c;
  return 0;
}

void subfv_destroy(subfv_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_mesh(c);
  for (auto &ch : c->chunks) (void)hipFree(ch.first);
  c->chunks.clear();
  for (auto &ph : c->pins) (void)hipHostFree(ph.first);
  c->pins.clear();
  if (c->derr) (void)hipFree(c->derr);
  if (c->st) (void)hipFree(c->st);
  if (c->ev) (void)hipEventDestroy(c->ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *subfv_error(const subfv_ctx *c) { return c ? c->err.c_str() : "null context"; }

int subfv_host_mesh_build(subfv_host_mesh **out, int32_t n_vert, const double *coord, int32_t n_elems,
                          const int32_t *kind, const int32_t *elem_vert, int32_t n_bfaces,
                          const int32_t *bf_nvert, const int32_t *bf_vert, const int32_t *bf_tag) {
  if (!out || !coord || !kind || !elem_vert) return SUBFV_E_INVALID;
  auto *m = new subfv_host_mesh();
  m->hm = std::make_shared<subfv::HostMesh>();
  std::string e = subfv::build_host_mesh(*m->hm, n_vert, coord, n_elems, kind, elem_vert, n_bfaces, bf_nvert,
                                         bf_vert, bf_tag);
  if (!e.empty()) {
    delete m;
    *out = nullptr;
    return SUBFV_E_INVALID;
  }
  m->tab = make_tables(*m->hm);
  *out = m;
  return 0;
}

int subfv_host_mesh_make_periodic(subfv_host_mesh *m) {
  if (!m || !m->hm) return SUBFV_E_INVALID;
  // a new connectivity: contexts that uploaded the old one keep it
  auto hm = std::make_shared<subfv::HostMesh>(*m->hm);
  if (!subfv::make_periodic(*hm).empty()) return SUBFV_E_INVALID;
  auto tab = make_tables(*hm);
  if (!tab->err.empty()) return SUBFV_E_INVALID;
  m->hm = hm;
  m->tab = tab;
  return 0;
}

int64_t subfv_host_mesh_get_i32(subfv_host_mesh *m, const char *name, int32_t *out, int64_t cap) {
  if (!m || !name || !out) return -1;
  return mesh_i32(*m->hm, name, out, cap);
}

void subfv_host_mesh_free(subfv_host_mesh *m) { delete m; }

int subfv_mesh_build(subfv_ctx *c, int32_t n_vert, const double *coord, int32_t n_elems, const int32_t *kind,
                     const int32_t *elem_vert, int32_t n_bfaces, const int32_t *bf_nvert, const int32_t *bf_vert,
                     const int32_t *bf_tag, int32_t boundary_2d) {
  if (!c || !coord || !kind || !elem_vert) return fail(c, SUBFV_E_INVALID, "null argument");
  if (n_bfaces > 0 && (!bf_nvert || !bf_vert || !bf_tag)) return fail(c, SUBFV_E_INVALID, "null boundary array");
  HIPCHK(c, hipSetDevice(c->device));
  free_mesh(c);
  auto hm = std::make_shared<subfv::HostMesh>();
  // build_mesh's connectivity on the device (mesh_device.hip)
  std::string e = subfv::build_device_mesh(*hm, n_vert, coord, n_elems, kind, elem_vert, n_bfaces, bf_nvert,
                                           bf_vert, bf_tag, (void *)c->stream);
  if (!e.empty()) return fail(c, SUBFV_E_INVALID, e);
  c->tab = make_tables(*hm);
  c->hmp = hm;
  return finish_mesh(c, boundary_2d);
}

int subfv_host_mesh_build_on_device(subfv_host_mesh **out, subfv_ctx *c, int32_t n_vert, const double *coord,
                                    int32_t n_elems, const int32_t *kind, const int32_t *elem_vert, int32_t n_bfaces,
                                    const int32_t *bf_nvert, const int32_t *bf_vert, const int32_t *bf_tag) {
  if (!out || !c || !coord || !kind || !elem_vert) return fail(c, SUBFV_E_INVALID, "null argument");
  if (n_bfaces > 0 && (!bf_nvert || !bf_vert || !bf_tag)) return fail(c, SUBFV_E_INVALID, "null boundary array");
  *out = nullptr;
  HIPCHK(c, hipSetDevice(c->device));
  auto *m = new subfv_host_mesh();
  m->hm = std::make_shared<subfv::HostMesh>();
  std::string e = subfv::build_device_mesh(*m->hm, n_vert, coord, n_elems, kind, elem_vert, n_bfaces, bf_nvert,
                                           bf_vert, bf_tag, (void *)c->stream);
  if (!e.empty()) {
    delete m;
    return fail(c, SUBFV_E_INVALID, e);
  }
  m->tab = make_tables(*m->hm);
  *out = m;
  return 0;
}

int subfv_mesh_upload(subfv_ctx *c, const subfv_host_mesh *hm, int32_t boundary_2d) {
  if (!c || !hm) return fail(c, SUBFV_E_INVALID, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  free_mesh(c);
  c->hmp = hm->hm;
  c->tab = hm->tab;
  return finish_mesh(c, boundary_2d);
}

int64_t subfv_mesh_count(const subfv_ctx *c, const char *what) {
  if (!c || !c->has_mesh || !what) return -1;
  std::string w(what);
  if (w == "n_vert") return c->hmp->nv;
  if (w == "n_elems") return c->hmp->ne;
  if (w == "n_faces") return c->hmp->nf;
  if (w == "n_sub_faces") return c->hmp->nsf;
  if (w == "n_sub_elems") return c->hmp->nse;
  if (w == "n_max_sub_faces") return c->hmp->nmax_sfn;
  if (w == "n_groups") return c->dm.ngroups;
  if (w == "diff_block_doubles") return c->diff_ready ? c->pre_n : 0;
  if (w == "device_bytes_used") {  // taken from the context's device chunks for this mesh
    int64_t u = (int64_t)c->chunk_off;
    for (size_t i = 0; i < c->chunk_i && i < c->chunks.size(); i++) u += (int64_t)c->chunks[i].second;
    return u;
  }
  return -1;
}

int64_t subfv_mesh_get_i32(subfv_ctx *c, const char *name, int32_t *out, int64_t cap) {
  if (!c || !c->has_mesh || !name || !out) return -1;
  if (std::string(name) == "vert_is_bound") {
    std::vector<uint8_t> b(c->hmp->nv);
    if (hipMemcpy(b.data(), c->dm.vbound, b.size(), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (cap < c->hmp->nv) return -1;
    for (int i = 0; i < c->hmp->nv; i++) out[i] = b[i];
    return c->hmp->nv;
  }
  return mesh_i32(*c->hmp, name, out, cap);
}

int64_t subfv_mesh_get_f64(subfv_ctx *c, const char *name, double *out, int64_t cap) {
  if (!c || !c->has_mesh || !name || !out) return -1;
  std::string n(name);
  const DevMesh &d = c->dm;
  auto cp = [&](const double *src, long cnt) -> int64_t {
    if (cnt > cap) return -1;
    if (hipMemcpy(out, src, cnt * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return cnt;
  };
  if (n == "vert_coord") {
    if (3L * d.V > cap) return -1;
    std::copy(c->hmp->vcoord.begin(), c->hmp->vcoord.end(), out);
    return 3L * d.V;
  }
  if (n == "vert_volume") return cp(d.vvol, d.V);
  if (n == "elem_coord") return cp(d.ccoord, 3 * d.C);
  if (n == "elem_volume") return cp(d.cvol, d.C);
  if (n == "face_coord") return cp(d.fcoord, 3 * d.F);
  if (n == "face_norm") return cp(d.fnorm, 3 * d.F);
  if (n == "face_area") return cp(d.farea, d.F);
  if (n == "sub_elem_volume") return cp(d.se_vol, d.SE);
  if (c->hmp->periodic && (n == "sub_face_area" || n == "sub_face_norm")) {
    // by sub-face id: each sub-face's first list entry
    const long nsf = c->hmp->nsf, k = n == "sub_face_area" ? 1 : 3;
    if (k * nsf > cap) return -1;
    std::vector<double> all(4 * d.SF);
    if (hipMemcpy(all.data(), d.sfn, all.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    const std::vector<int> &prim = c->tab->prim;
    for (long q = 0; q < nsf; q++) {
      if (k == 1)
        out[q] = all[3 * d.SF + prim[q]];
      else
        for (int j = 0; j < 3; j++) out[3 * q + j] = all[j * d.SF + prim[q]];
    }
    return k * nsf;
  }
  if (n == "sub_face_area") return cp(d.sfn + 3 * d.SF, d.SF);
  if (n == "sub_face_norm") {
    if (3 * d.SF > cap) return -1;
    return dev_to_host_soa(c, d.sfn, 3, d.SF, out) ? -1 : 3 * d.SF;
  }
  return -1;
}

int subfv_set_arith(subfv_ctx *c, int32_t mode) {
  if (!c) return SUBFV_E_INVALID;
  if (mode != SUBFV_ARITH_EXACT && mode != SUBFV_ARITH_FAST)
    return fail(c, SUBFV_E_INVALID, "arithmetic mode: SUBFV_ARITH_EXACT or SUBFV_ARITH_FAST");
  c->arith = mode;
  c->gen++;
  return 0;
}

int subfv_set_config(subfv_ctx *c, const subfv_config *cfg) {
  if (!c || !cfg) return fail(c, SUBFV_E_INVALID, "null argument");
  switch (cfg->scheme_id) {
    case kMultiPoint: case kMultiPointIso: case kThreeWave: case kTwoWave: case kModifiedThreeWave:
    case kMultiPointPressure: case kMultiPointPressurePh: break;
    default: return fail(c, SUBFV_E_INVALID, "ERROR: scheme unrecognized");
  }
  if (cfg->n_bc < 0 || cfg->n_bc > SUBFV_MNBC) return fail(c, SUBFV_E_INVALID, "n_bc out of range");
  if (cfg->method < 0 || cfg->method > 6) return fail(c, SUBFV_E_INVALID, "Method for second order unknown !");
  for (int i = 0; i < cfg->n_bc; i++)
    if (cfg->bc_euler_id[i] < 1 || cfg->bc_euler_id[i] > 9)
      return fail(c, SUBFV_E_INVALID, "ERROR: bc_type unrecognized");
  c->cfg = *cfg;
  c->gen++;
  BCParams &b = c->bc;
  b = BCParams{};
  b.scheme = cfg->scheme_id;
  b.bc_style = cfg->bc_style;
  b.b2d = cfg->boundary_2d;
  b.excl = cfg->exclude_bound_vert;
  b.n_bc = cfg->n_bc;
  b.lts = cfg->local_time_step;
  for (int i = 0; i < cfg->n_bc; i++) {
    b.bc_id[i + 1] = cfg->bc_euler_id[i];
    for (int k = 0; k < 5; k++) b.bc_w[i + 1][k] = cfg->bc_val[i][k];
  }
  int r = apply_bc(c);
  if (r) return r;
  return push_state_scalars(c, c->host_t);
}

int subfv_set_time(subfv_ctx *c, double t) {
  if (!c) return SUBFV_E_INVALID;
  c->host_t = t;
  c->gen++;
  return push_state_scalars(c, t);
}

int subfv_state_put(subfv_ctx *c, const double *sol) {
  if (!c || !c->has_mesh || !sol) return fail(c, SUBFV_E_INVALID, "no mesh or null state");
  c->gen++;
  return host_to_dev_soa(c, sol, 1, 5 * c->dm.C, c->U[c->cur]);
}

// init_sol with init_1drp (ns_euler_module.F90:36-43) on the device: each
// cell takes the left or right primitive state by its centroid's x
int subfv_init_1drp(subfv_ctx *c, double x1drp, const double *w_left, const double *w_right) {
  if (!c || !c->has_mesh || !w_left || !w_right) return fail(c, SUBFV_E_INVALID, "no mesh or null state");
  c->gen++;
  Init1drp a;
  a.x1drp = x1drp;
  for (int k = 0; k < 5; k++) {
    a.wl[k] = w_left[k];
    a.wr[k] = w_right[k];
  }
  hipLaunchKernelGGL(k_init_1drp, dim3(nblk(c->dm.C, kCellThreads)), dim3(kCellThreads), 0, c->stream, c->dm, a,
                     c->U[c->cur]);
  HIPCHK(c, hipGetLastError());
  return 0;
}

int subfv_state_get(subfv_ctx *c, double *sol) {
  if (!c || !c->has_mesh || !sol) return fail(c, SUBFV_E_INVALID, "no mesh or null state");
  return dev_to_host_soa(c, c->U[c->cur], 1, 5 * c->dm.C, sol);
}

int64_t subfv_get_f64(subfv_ctx *c, const char *name, double *out, int64_t cap) {
  if (!c || !c->has_mesh || !name || !out) return -1;
  std::string n(name);
  if (n == "dt_solve") {  // the dt of the last stage update's solve (before a positivity shrink)
    if (cap < 1) return -1;
    StepState hs;
    if (sync_state_scalars(c, &hs)) return -1;
    *out = hs.dt;
    return 1;
  }
  if (n == "local_time_step") {  // 1.0 when the configuration steps in local time
    if (cap < 1) return -1;
    *out = c->cfg.local_time_step ? 1.0 : 0.0;
    return 1;
  }
  const long C = c->dm.C, V = c->dm.V;
  struct Item { const char *nm; double *p; int nc; long n; };
  Item items[] = {{"sol", c->U[c->cur], 1, 5 * C}, {"rhs", c->rhs, 5, C},      {"rhs_o2", c->rhs2, 5, C},
                  {"sum_lambda", c->sl, 1, C},   {"sum_lambda_o2", c->sl2, 1, C},
                  // gradients are stored in the reference layout already
                  {"cell_grad", c->G, 1, 15 * C}, {"nodal_grad", c->NG, 1, 15 * V}, {"vp", c->vp, 3, V},
                  // viscous block: diag(5,5,C) and the viscous rhs of components 2..5
                  {"diag", c->diag, 1, 25 * C}, {"rhs_diff", c->rdiff, 1, 4 * C}};
  for (auto &it : items)
    if (n == it.nm) {
      if (!it.p) return -1;
      if ((int64_t)it.nc * it.n > cap) return -1;
      return dev_to_host_soa(c, it.p, it.nc, it.n, out) ? -1 : (int64_t)it.nc * it.n;
    }
  return -1;
}

int64_t subfv_put_f64(subfv_ctx *c, const char *name, const double *in, int64_t nval) {
  if (!c || !c->has_mesh || !name || !in) return -1;
  std::string n(name);
  const long C = c->dm.C;
  struct Item { const char *nm; double *p; int nc; long n; };
  Item items[] = {{"sol", c->U[c->cur], 1, 5 * C}, {"rhs", c->rhs, 5, C}, {"rhs_o2", c->rhs2, 5, C},
                  {"sum_lambda", c->sl, 1, C}, {"cell_grad", c->G, 1, 15 * C}};
  c->gen++;
  for (auto &it : items)
    if (n == it.nm) {
      if ((int64_t)it.nc * it.n != nval) return -1;
      return host_to_dev_soa(c, in, it.nc, it.n, it.p) ? -1 : nval;
    }
  return -1;
}

int subfv_compute_euler_flux(subfv_ctx *c, int32_t second_order, int32_t want_vp) {
  if (!c || !c->has_mesh) return fail(c, SUBFV_E_INVALID, "no mesh");
  const double *U = c->U[c->cur];
  double *P = second_order ? c->P2 : c->P;
  int r = launch_flux(c, U, second_order ? c->G : nullptr, P, want_vp != 0);
  if (r) return r;
  r = launch_gather(c, P, nullptr, second_order ? c->rhs2 : c->rhs, nullptr, second_order ? c->sl2 : c->sl, 1);
  if (r) return r;
  return check_err(c);
}

int subfv_compute_cell_grad(subfv_ctx *c, int32_t method) {
  if (!c || !c->has_mesh) return fail(c, SUBFV_E_INVALID, "no mesh");
  if (method < 0 || method > 6) return fail(c, SUBFV_E_INVALID, "Method for second order unknown !");
  int r = launch_grad(c, c->U[c->cur], method);
  if (r) return r;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int subfv_flux_limiting(subfv_ctx *c, double cfl) {
  if (!c || !c->has_mesh) return fail(c, SUBFV_E_INVALID, "no mesh");
  StepState hs;
  int r = sync_state_scalars(c, &hs);
  if (r) return r;
  hs.cfl = cfl;
  HIPCHK(c, hipMemcpyAsync(c->st, &hs, sizeof(StepState), hipMemcpyHostToDevice, c->stream));
  const DevMesh &d = c->dm;
  hipLaunchKernelGGL(k_dt_from_sl, dim3(nblk(d.C, kCellThreads)), dim3(kCellThreads), 0, c->stream, d, c->sl,
                     c->blk);
  hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(kCellThreads), 0, c->stream, (const double *)c->blk,
                     (int)nblk(d.C, kCellThreads), c->st, 0);
  hipLaunchKernelGGL(k_flux_limiting, dim3(nblk(d.C, kCellThreads)), dim3(kCellThreads), 0, c->stream, d,
                     c->U[c->cur], c->rhs, c->rhs2, c->sl, c->st);
  HIPCHK(c, hipGetLastError());
  hs.cfl = c->cfg.cfl;
  HIPCHK(c, hipMemcpyAsync(&c->st->cfl, &hs.cfl, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int subfv_add_timedisc_and_solve(subfv_ctx *c, double a1, double a2, double cfl, double t, double *dt,
                                 double *omega) {
  if (!c || !c->has_mesh) return fail(c, SUBFV_E_INVALID, "no mesh");
  StepState hs{};
  hs.t = t;
  hs.tmax = c->cfg.tmax;
  hs.cfl = cfl;
  hs.lts = c->cfg.local_time_step;
  hs.omega = 1.0;
  HIPCHK(c, hipMemcpyAsync(c->st, &hs, sizeof(StepState), hipMemcpyHostToDevice, c->stream));
  const DevMesh &d = c->dm;
  hipLaunchKernelGGL(k_dt_from_sl, dim3(nblk(d.C, kCellThreads)), dim3(kCellThreads), 0, c->stream, d, c->sl,
                     c->blk);
  hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(kCellThreads), 0, c->stream, (const double *)c->blk,
                     (int)nblk(d.C, kCellThreads), c->st, 0);
  int nxt = (c->cur + 1) % 3;
  int r = launch_update(c, c->U[c->cur], nullptr, c->U[nxt], a1, a2, 0.0, 1.0, 0, 0);
  if (r) return r;
  r = sync_state_scalars(c, &hs);
  if (r) return r;
  if (omega) *omega = hs.omega;
  if (dt) *dt = hs.dt_adv;
  r = check_err(c);
  if (r) return r;
  c->cur = nxt;
  HIPCHK(c, hipMemcpyAsync(c->st, &hs, sizeof(StepState), hipMemcpyHostToDevice, c->stream));
  return 0;
}

int subfv_compute_rhs_ns(subfv_ctx *c) {
  if (!c || !c->has_mesh) return fail(c, SUBFV_E_INVALID, "no mesh");
  int r = subfv_compute_euler_flux(c, 0, 1);
  if (r) return r;
  if (c->cfg.second_order) {
    r = subfv_compute_cell_grad(c, c->cfg.method);
    if (r) return r;
    r = subfv_compute_euler_flux(c, 1, 1);
    if (r) return r;
    r = subfv_flux_limiting(c, c->cfg.cfl);
    if (r) return r;
  }
  if (c->diff_on) {
    // the diffusion block (ns_main.F90:170-175): rhs += viscous terms, diag = their block
    r = launch_diffusion(c, c->U[c->cur]);
    if (r) return r;
    hipLaunchKernelGGL(k_add_rdiff, dim3(nblk(c->dm.C, 256)), dim3(256), 0, c->stream, (long)c->dm.C, c->rhs,
                       (const double *)c->rdiff);
    HIPCHK(c, hipGetLastError());
  }
  return 0;
}

}  // extern "C"

namespace {
// device scratch of one sweep call, freed on every exit path
struct Scratch {
  std::vector<void *> p;
  ~Scratch() {
    for (void *q : p) (void)hipFree(q);
  }
  template <class T>
  hipError_t get(T **out, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) p.push_back(q);
    *out = (T *)q;
    return e;
  }
};
}  // namespace

extern "C" {

// compute_rhs_lagrange (src/lagrange/lagrange_module.F90:14-147)
int subfv_compute_rhs_lagrange(subfv_ctx *c, const double *sol, const double *mass, int32_t b2d, double *rhs,
                               double *vp) {
  if (!c || !c->has_mesh) return fail(c, SUBFV_E_INVALID, "no mesh");
  if (c->hmp->periodic) return fail(c, SUBFV_E_INVALID, "not available on periodic meshes");
  if (!sol || !mass || !rhs || !vp) return fail(c, SUBFV_E_INVALID, "null array");
  HIPCHK(c, hipSetDevice(c->device));
  const DevMesh &m = c->dm;
  Scratch sc;
  double *dsol, *dmass, *T, *dvp, *drhs;
  HIPCHK(c, sc.get(&dsol, 5 * m.C));
  HIPCHK(c, sc.get(&dmass, m.C));
  HIPCHK(c, sc.get(&T, (size_t)kSweepTerm * kSlotSrc * m.SE));
  HIPCHK(c, sc.get(&dvp, 3 * m.V));
  HIPCHK(c, sc.get(&drhs, 5 * m.C));
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(dsol, sol, 5 * m.C * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dmass, mass, m.C * sizeof(double), hipMemcpyHostToDevice, s));
  LagrangeArgs a;
  a.m = m;
  a.sol = dsol;
  a.mass = dmass;
  a.T = T;
  a.vp = dvp;
  a.b2d = b2d ? 1 : 0;
  a.nblocks = m.ngroups;
  a.per_xcd = (m.ngroups + 7) / 8;
  hipLaunchKernelGGL(k_lagrange_group, dim3(8 * a.per_xcd), dim3(kGrpThreads), 0, s, a);
  hipLaunchKernelGGL(k_sweep_gather<5>, dim3(nblk(m.C, kCellThreads)), dim3(kCellThreads), 0, s, m,
                     (const double *)T, drhs);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(rhs, drhs, 5 * m.C * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(vp, dvp, 3 * m.V * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  return 0;
}

// n_steps of the acoustic loop (src/acoustic/acoustic_main.F90:171-290)
int subfv_acoustic_run(subfv_ctx *c, double *sol, int32_t n_steps, double cfl, int32_t n_bc, const int32_t *bc_kind,
                       const double *bc_val, double *t, double *dt_last) {
  if (!c || !c->has_mesh) return fail(c, SUBFV_E_INVALID, "no mesh");
  if (c->hmp->periodic) return fail(c, SUBFV_E_INVALID, "not available on periodic meshes");
  if (!sol || !t || n_steps < 0 || n_bc < 0 || (n_bc > 0 && (!bc_kind || !bc_val)))
    return fail(c, SUBFV_E_INVALID, "bad acoustic arguments");
  for (int i = 0; i < n_bc; i++)
    if (bc_kind[i] < 0 || bc_kind[i] > 2)
      return fail(c, SUBFV_E_BAD_BC, "acoustic boundary kinds are 0 wall, 1 freestream, 2 outflow");
  HIPCHK(c, hipSetDevice(c->device));
  const DevMesh &m = c->dm;
  Scratch sc;
  double *dsol, *T, *blk, *dt_t, *dval;
  int *dkind, *derr;
  const int ng = m.ngroups;
  HIPCHK(c, sc.get(&dsol, 4 * m.C));
  HIPCHK(c, sc.get(&T, (size_t)kSweepTerm * kSlotSrc * m.SE));
  HIPCHK(c, sc.get(&blk, ng));
  HIPCHK(c, sc.get(&dt_t, 2));
  HIPCHK(c, sc.get(&dkind, n_bc + 1));
  HIPCHK(c, sc.get(&dval, 4 * (n_bc + 1)));
  HIPCHK(c, sc.get(&derr, 1));
  std::vector<int> hk(n_bc + 1, 0);
  std::vector<double> hv(4 * (n_bc + 1), 0.0);
  for (int i = 0; i < n_bc; i++) {
    hk[i + 1] = bc_kind[i];
    for (int k = 0; k < 4; k++) hv[4 * (i + 1) + k] = bc_val[4 * i + k];
  }
  const double ht[2] = {0.0, *t};
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(dsol, sol, 4 * m.C * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dkind, hk.data(), hk.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dval, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(dt_t, ht, sizeof(ht), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemsetAsync(derr, 0, sizeof(int), s));
  AcousticArgs a;
  a.m = m;
  a.sol = dsol;
  a.T = T;
  a.blk = blk;
  a.bc_kind = dkind;
  a.bc_val = dval;
  a.n_bc = n_bc;
  a.err = derr;
  a.nblocks = ng;
  a.per_xcd = (ng + 7) / 8;
  for (int it = 0; it < n_steps; it++) {
    hipLaunchKernelGGL(k_acoustic_group, dim3(8 * a.per_xcd), dim3(kGrpThreads), 0, s, a);
    hipLaunchKernelGGL(k_acoustic_dt, dim3(1), dim3(kCellThreads), 0, s, (const double *)blk, ng, dt_t);
    hipLaunchKernelGGL(k_acoustic_update, dim3(nblk(m.C, kCellThreads)), dim3(kCellThreads), 0, s, m,
                       (const double *)T, dsol, cfl, (const double *)dt_t);
  }
  HIPCHK(c, hipGetLastError());
  double hdt[2];
  int herr = 0;
  HIPCHK(c, hipMemcpyAsync(hdt, dt_t, sizeof(hdt), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(sol, dsol, 4 * m.C * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (herr) return fail(c, SUBFV_E_BAD_BC, "acoustic: a boundary tag has no boundary kind");
  *t = hdt[1];
  if (dt_last) *dt_last = hdt[0];
  return 0;
}

// One full residual evaluation on state U into rhs (and rhs2 in second
// order), with dt from the first-order sum_lambda.
static int residual(subfv_ctx *c, const double *U, int fixed_dt) {
  int r;
  if (c->diff_on && (r = launch_diffusion(c, U))) return r;
  if ((r = launch_prim(c, U))) return r;
  if (c->cfg.second_order) {
    if ((r = launch_grad(c, U, c->cfg.method, c->W))) return r;
    if ((r = launch_flux(c, U, nullptr, c->P, false, c->W))) return r;
    if ((r = launch_flux(c, U, c->G, c->P2, false, c->W))) return r;
    return launch_gather(c, c->P, c->P2, c->rhs, c->rhs2, c->sl, fixed_dt);
  }
  if ((r = launch_flux(c, U, nullptr, c->P, false, c->W))) return r;
  return launch_gather(c, c->P, nullptr, c->rhs, nullptr, c->sl, fixed_dt);
}

int subfv_profile_enable(subfv_ctx *c, int32_t on) {
  if (!c) return SUBFV_E_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &e : c->prof_ev) {
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  c->prof_ev.clear();
  c->prof = on != 0;
  return 0;
}

int subfv_profile_read(subfv_ctx *c, const char *phase, double *ms_total, int64_t *count) {
  if (!c || !phase) return SUBFV_E_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double tot = 0.0;
  int64_t n = 0;
  for (auto &e : c->prof_ev)
    if (e.first == phase) {
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, e.second.first, e.second.second));
      tot += ms;
      n++;
    }
  if (ms_total) *ms_total = tot;
  if (count) *count = n;
  return 0;
}

int subfv_set_diffusion(subfv_ctx *c, int32_t activate, double mu_p, double cv_p, double prandtl, int32_t n_bc,
                        const int32_t *bc_v_id, const double *bc_val_v, const int32_t *bc_t_id,
                        const double *bc_val_t) {
  if (!c) return SUBFV_E_INVALID;
  if (n_bc < 0 || n_bc > SUBFV_MNBC) return fail(c, SUBFV_E_INVALID, "n_bc out of range");
  if (activate && !(mu_p > 0.0 && cv_p > 0.0 && prandtl > 0.0))
    return fail(c, SUBFV_E_INVALID, "viscous terms need mu_p, Cv_p and Prandtl > 0");
  DiffParams p{};
  p.mu = mu_p;
  p.mu0 = mu_p;  // compute_mu without Sutherland: mu0corr = mu_p
  p.cv = cv_p;
  p.kap = mu_p * cv_p * kGamma / prandtl;  // kappa (ns_diffusion_module.F90:186-194)
  p.pr = prandtl;
  p.suth = c->dparams.suth;  // set by subfv_set_sutherland
  p.suth_mu0 = c->dparams.suth_mu0;
  p.suth_T0 = c->dparams.suth_T0;
  for (int i = 0; i < 11; i++) {
    p.bc_V_id[i] = 1;
    p.bc_T_id[i] = 1;
  }
  for (int i = 0; i < n_bc; i++) {  // BC index i + 1, as the right-neighbour codes count them
    p.bc_V_id[i + 1] = bc_v_id ? bc_v_id[i] : 1;
    p.bc_T_id[i + 1] = bc_t_id ? bc_t_id[i] : 1;
    for (int d = 0; d < 3; d++) p.bc_val_V[i + 1][d] = bc_val_v ? bc_val_v[3 * i + d] : 0.0;
    p.bc_val_T[i + 1] = bc_val_t ? bc_val_t[i] : 0.0;
  }
  const bool changed = std::memcmp(&p, &c->dparams, sizeof(p)) != 0;
  c->dparams = p;
  c->diff_on = activate != 0;
  if (changed) c->diff_ready = false;
  return 0;
}

int subfv_set_sutherland(subfv_ctx *c, int32_t activate, double mu0, double T0) {
  if (!c) return SUBFV_E_INVALID;
  if (activate && !(mu0 > 0.0 && T0 > 0.0)) return fail(c, SUBFV_E_INVALID, "Sutherland's law needs mu0 and T0 > 0");
  DiffParams p = c->dparams;
  p.suth = activate != 0;
  p.suth_mu0 = activate ? mu0 : 0.0;
  p.suth_T0 = activate ? T0 : 0.0;
  if (std::memcmp(&p, &c->dparams, sizeof(p)) != 0) {
    c->dparams = p;
    c->diff_ready = false;
  }
  return 0;
}

// the cell viscosity buffers (and, past the blocks' minima, mu0corr), once per mesh
int ensure_mu_buffers(subfv_ctx *c) {
  int r;
  if (!c->mu_cell && (r = dalloc(c, &c->mu_cell, (size_t)c->dm.C))) return r;
  if (!c->mu_blk && (r = dalloc(c, &c->mu_blk, (size_t)nblk(c->dm.C, kCellThreads) + 1))) return r;
  return 0;
}

int subfv_compute_mu(subfv_ctx *c, const double *sol, double *mu_arr, double *mu0corr) {
  if (!c || !c->has_mesh || !sol || !mu_arr || !mu0corr) return fail(c, SUBFV_E_INVALID, "no mesh or null array");
  if (!c->diff_on) return fail(c, SUBFV_E_INVALID, "compute_mu: viscous terms not activated (subfv_set_diffusion)");
  HIPCHK(c, hipSetDevice(c->device));
  const long C = c->dm.C;
  int r;
  if (!c->dparams.suth) {  // mu_arr = mu_p, mu0corr = mu_p (:344-345)
    for (long e = 0; e < C; e++) mu_arr[e] = c->dparams.mu;
    *mu0corr = c->dparams.mu;
    return 0;
  }
  if ((r = subfv_state_put(c, sol))) return r;
  if ((r = ensure_mu_buffers(c))) return r;
  const unsigned nb = nblk(C, kCellThreads);
  hipLaunchKernelGGL(k_compute_mu, dim3(nb), dim3(kCellThreads), 0, c->stream, C, (const double *)c->U[c->cur],
                     c->dparams.cv, c->dparams.suth_mu0, c->dparams.suth_T0, c->mu_cell, c->mu_blk);
  hipLaunchKernelGGL(k_mu_min, dim3(1), dim3(kCellThreads), 0, c->stream, (const double *)c->mu_blk, (int)nb,
                     c->mu_blk + nb);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(mu_arr, c->mu_cell, C * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(mu0corr, c->mu_blk + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"

// rhs(2:5, c) += rdiff(1:4, c) and diag(:, :, c) += vdiag(:, :, c) in the
// caller's layouts sol(5, C) and diag(5, 5, C)
__global__ void k_add_diffusion(long C, double *rhs, double *diag, const double *rdiff, const double *vdiag) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  for (int k = 0; k < 4; k++) rhs[5 * c + 1 + k] = rhs[5 * c + 1 + k] + rdiff[4 * c + k];
  for (int q = 0; q < 25; q++) diag[25 * c + q] = diag[25 * c + q] + vdiag[25 * c + q];
}

extern "C" {

int subfv_add_diffusion(subfv_ctx *c, const double *sol, const double *mu_arr, double mu0corr, int32_t parts,
                        double *rhs, double *diag) {
  if (!c || !c->has_mesh || !sol || !mu_arr || !rhs || !diag) return fail(c, SUBFV_E_INVALID, "no mesh or null array");
  if (parts < 1 || parts > 3) return fail(c, SUBFV_E_INVALID, "parts: 1 momentum, 2 heat, 3 both");
  if (!c->diff_on) return fail(c, SUBFV_E_INVALID, "viscous terms not activated (subfv_set_diffusion)");
  HIPCHK(c, hipSetDevice(c->device));
  const long C = c->dm.C;
  int r;
  if ((r = subfv_state_put(c, sol))) return r;
  if ((r = ensure_diffusion(c))) return r;
  if ((r = ensure_mu_buffers(c))) return r;
  const unsigned nb = nblk(C, kCellThreads);
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(c->mu_cell, mu_arr, C * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->mu_blk + nb, &mu0corr, sizeof(double), hipMemcpyHostToDevice, s));
  if ((r = launch_diffusion(c, c->U[c->cur], c->mu_cell, c->mu_blk + nb, parts))) return r;
  Scratch sc;
  double *drhs, *ddiag;
  HIPCHK(c, sc.get(&drhs, 5 * C));
  HIPCHK(c, sc.get(&ddiag, 25 * C));
  HIPCHK(c, hipMemcpyAsync(drhs, rhs, 5 * C * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(ddiag, diag, 25 * C * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_add_diffusion, dim3(nblk(C, 256)), dim3(256), 0, s, C, drhs, ddiag, (const double *)c->rdiff,
                     (const double *)c->diag);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(rhs, drhs, 5 * C * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(diag, ddiag, 25 * C * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  return 0;
}

int subfv_run(subfv_ctx *c, int32_t n_steps, int32_t time_scheme, double *t, double *dt) {
  if (!c || !c->has_mesh) return fail(c, SUBFV_E_INVALID, "no mesh");
  if (time_scheme == 1 && c->cfg.local_time_step)
    return fail(c, SUBFV_E_INVALID, "SSP-RK3 needs local_time_step = .false.");
  int r;
  const int fl = c->cfg.second_order ? 1 : 0;
  for (int it = 0; it < n_steps; it++) {
    int a = c->cur, b = (a + 1) % 3, q = (a + 2) % 3;
    if (time_scheme == 0) {
      if ((r = residual(c, c->U[a], 0))) return r;
      if (c->cfg.second_order) {
        // ns_main.F90:88-95: half step into U*, residual at U*, full step from U^n
        if ((r = launch_update(c, c->U[a], nullptr, c->U[b], 2.0, 1.0, 0.0, 1.0, fl, 0, true))) return r;
        if ((r = residual(c, c->U[b], 0))) return r;
        if ((r = launch_update(c, c->U[a], nullptr, c->U[b], 1.0, 1.0, 0.0, 1.0, fl, 1, true))) return r;
      } else {
        if ((r = launch_update(c, c->U[a], nullptr, c->U[b], 1.0, 1.0, 0.0, 1.0, 0, 1, true))) return r;
      }
      c->cur = b;
    } else {
      // SSP-RK3: U1 = L(Un) -> b; U2 = 3/4 Un + 1/4 L(U1) -> q; Un+1 = 1/3 Un + 2/3 L(U2) -> b
      if ((r = residual(c, c->U[a], 0))) return r;
      if ((r = launch_update(c, c->U[a], nullptr, c->U[b], 1.0, 1.0, 0.0, 1.0, fl, 0, true))) return r;
      if ((r = residual(c, c->U[b], 1))) return r;
      if ((r = launch_update(c, c->U[b], c->U[a], c->U[q], 1.0, 1.0, 0.75, 0.25, fl, 0, true))) return r;
      if ((r = residual(c, c->U[q], 1))) return r;
      if ((r = launch_update(c, c->U[q], c->U[a], c->U[b], 1.0, 1.0, 1.0 / 3.0, 2.0 / 3.0, fl, 2, true))) return r;
      c->cur = b;
    }
  }
  StepState hs;
  r = sync_state_scalars(c, &hs);
  if (r) return r;
  if (t) *t = hs.t;
  if (dt) *dt = (time_scheme == 0) ? hs.dt_adv : hs.dt;
  c->host_t = hs.t;
  return check_err(c);
}

}  // extern "C"

// ===================================================================== BCSR
struct subfv_bcsr {
  subfv_ctx *ctx = nullptr;
  int m = 0, b = 0, nnzb = 0;
  long n = 0;
  int *row = nullptr, *col = nullptr, *diag_idx = nullptr;
  double *v = nullptr, *F64 = nullptr;
  float *v32 = nullptr, *F32 = nullptr;
  // GMRES workspace
  int mr = 0;
  double *V = nullptr, *w = nullptr, *r = nullptr, *x = nullptr, *rhs = nullptr, *tmp = nullptr, *zp = nullptr;
  double *H = nullptr, *scal = nullptr, *z = nullptr, *part = nullptr, *part2 = nullptr;
  double *dcp = nullptr;  // DCGS2 partials ((2 kGsMax + 2) x kGsBlocks)
  // one DCGS2 Arnoldi cycle as a graph, replayed every restart; the restart
  // length, preconditioner, precision and basis it was captured for
  hipGraph_t cyc_graph = nullptr;
  hipGraphExec_t cyc_exec = nullptr;
  int cyc_mr = -1, cyc_kind = -1, cyc_fp32 = -1;
  const void *cyc_v = nullptr;
  double *vr = nullptr;   // A D^-1 (block-Jacobi right preconditioning folded in), FP64
  float *vr32 = nullptr;  // and FP32; formed by the first block-Jacobi GMRES after the values change
  bool vr_ok = false;
  double *dch = nullptr;  // DCGS2 scalars (kDchSize)
  // host copies (0-based) for add_block and the sweep schedules
  std::vector<int> h_row, h_col, h_diag;
  // bcsr_lusgs: rows grouped by dependency level of the lower triangle
  std::vector<int> lvl_off;
  int *lvl_rows = nullptr;
  // bcsr_parallel_lusgs: colour sets
  int nsets = 0;
  std::vector<int> set_off;
  int *set_rows = nullptr, *color = nullptr;
  // bcsr_pcgm
  double *p = nullptr, *pmr = nullptr;
  int *perr = nullptr;
  std::vector<void *> allocs;
};

namespace {

template <class T>
int balloc(subfv_bcsr *A, T **p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return fail(A->ctx, SUBFV_E_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
  A->allocs.push_back(q);
  *p = (T *)q;
  return 0;
}

constexpr int kRedBlocks = 1024;

// y = A x with the values v64 (or v32 in FP32), A's pattern
int spmv_v(subfv_bcsr *A, bool fp32, const double *v64, const float *v32, const double *x, double *y) {
  hipStream_t s = A->ctx->stream;
  long thr = (long)A->m * A->b;
  dim3 g(nblk(thr, 256)), bl(256);
  if (A->b == 4) {
    if (fp32) hipLaunchKernelGGL((k_bcsr_spmv<float, 4>), g, bl, 0, s, A->m, A->row, A->col, v32, x, y);
    else hipLaunchKernelGGL((k_bcsr_spmv<double, 4>), g, bl, 0, s, A->m, A->row, A->col, v64, x, y);
  } else if (A->b == 5) {
    if (fp32) hipLaunchKernelGGL((k_bcsr_spmv<float, 5>), g, bl, 0, s, A->m, A->row, A->col, v32, x, y);
    else hipLaunchKernelGGL((k_bcsr_spmv<double, 5>), g, bl, 0, s, A->m, A->row, A->col, v64, x, y);
  } else if (A->b == 3) {
    if (fp32) hipLaunchKernelGGL((k_bcsr_spmv<float, 3>), g, bl, 0, s, A->m, A->row, A->col, v32, x, y);
    else hipLaunchKernelGGL((k_bcsr_spmv<double, 3>), g, bl, 0, s, A->m, A->row, A->col, v64, x, y);
  } else {
    return fail(A->ctx, SUBFV_E_INVALID, "block size must be 3, 4 or 5");
  }
  HIPCHK(A->ctx, hipGetLastError());
  return 0;
}

int spmv(subfv_bcsr *A, bool fp32, const double *x, double *y) { return spmv_v(A, fp32, A->v, A->v32, x, y); }

// A D^-1 for block-Jacobi GMRES, when the values or factors changed since
int ensure_rprec(subfv_bcsr *A) {
  if (A->vr_ok) return 0;
  dim3 g(nblk(A->nnzb, 256)), bl(256);
  hipStream_t s = A->ctx->stream;
  if (A->b == 3)
    hipLaunchKernelGGL(k_bcsr_rprec<3>, g, bl, 0, s, A->nnzb, A->col, A->v, A->F64, A->vr, A->vr32);
  else if (A->b == 4)
    hipLaunchKernelGGL(k_bcsr_rprec<4>, g, bl, 0, s, A->nnzb, A->col, A->v, A->F64, A->vr, A->vr32);
  else
    hipLaunchKernelGGL(k_bcsr_rprec<5>, g, bl, 0, s, A->nnzb, A->col, A->v, A->F64, A->vr, A->vr32);
  HIPCHK(A->ctx, hipGetLastError());
  A->vr_ok = true;
  return 0;
}

int bjacobi(subfv_bcsr *A, bool fp32, const double *bvec, double *x) {
  hipStream_t s = A->ctx->stream;
  dim3 g(nblk(A->m, 256)), bl(256);
  switch (A->b) {
    case 3:
      if (fp32) hipLaunchKernelGGL((k_bcsr_bjacobi<float, 3>), g, bl, 0, s, A->m, A->F32, bvec, x);
      else hipLaunchKernelGGL((k_bcsr_bjacobi<double, 3>), g, bl, 0, s, A->m, A->F64, bvec, x);
      break;
    case 4:
      if (fp32) hipLaunchKernelGGL((k_bcsr_bjacobi<float, 4>), g, bl, 0, s, A->m, A->F32, bvec, x);
      else hipLaunchKernelGGL((k_bcsr_bjacobi<double, 4>), g, bl, 0, s, A->m, A->F64, bvec, x);
      break;
    default:
      if (fp32) hipLaunchKernelGGL((k_bcsr_bjacobi<float, 5>), g, bl, 0, s, A->m, A->F32, bvec, x);
      else hipLaunchKernelGGL((k_bcsr_bjacobi<double, 5>), g, bl, 0, s, A->m, A->F64, bvec, x);
  }
  HIPCHK(A->ctx, hipGetLastError());
  return 0;
}

// *out = <a, b> (or <w, w>), device scalar
void mgs_step(subfv_bcsr *A, double *w, const double *vprev, const double *h, const double *vnext, double *out) {
  hipLaunchKernelGGL(k_mgs_step, dim3(kRedBlocks), dim3(kVecThreads), 0, A->ctx->stream, A->n, w, vprev, h, vnext,
                     A->part);
  hipLaunchKernelGGL(k_sum_finalize, dim3(1), dim3(kVecThreads), 0, A->ctx->stream, (const double *)A->part,
                     kRedBlocks, out);
}

int norm_host(subfv_bcsr *A, double *w, double *out) {
  mgs_step(A, w, nullptr, nullptr, nullptr, A->scal);
  double ss = 0.0;
  HIPCHK(A->ctx, hipMemcpyAsync(&ss, A->scal, sizeof(double), hipMemcpyDeviceToHost, A->ctx->stream));
  HIPCHK(A->ctx, hipStreamSynchronize(A->ctx->stream));
  *out = sqrt(ss);
  return 0;
}

int ensure_workspace(subfv_bcsr *A, int mr) {
  if (A->mr >= mr) return 0;
  for (double **p : {&A->V, &A->H}) {
    if (*p) {
      (void)hipFree(*p);
      A->allocs.erase(std::find(A->allocs.begin(), A->allocs.end(), (void *)*p));
      *p = nullptr;
    }
  }
  int r;
  if ((r = balloc(A, &A->V, (size_t)(A->n + (A->n & 1)) * (mr + 1)))) return r;  // columns of even length
  if ((r = balloc(A, &A->H, (size_t)(mr + 1) * mr + 8))) return r;
  A->mr = mr;
  return 0;
}


// bcsr_lusgs: out = (L + D)^-1 (L + I)^-1 in, level by level (y in A->w)
template <int B>
void lusgs_b(subfv_bcsr *A, const double *in, double *out) {
  hipStream_t s = A->ctx->stream;
  const int nlev = (int)A->lvl_off.size() - 1;
  for (int l = 0; l < nlev; l++) {
    const int o = A->lvl_off[l], cnt = A->lvl_off[l + 1] - o;
    hipLaunchKernelGGL((k_lusgs_level<B, false>), dim3(nblk(cnt, 256)), dim3(256), 0, s, cnt, A->lvl_rows + o, A->row,
                       A->col, A->v, (const double *)A->F64, in, A->w);
  }
  for (int l = 0; l < nlev; l++) {
    const int o = A->lvl_off[l], cnt = A->lvl_off[l + 1] - o;
    hipLaunchKernelGGL((k_lusgs_level<B, true>), dim3(nblk(cnt, 256)), dim3(256), 0, s, cnt, A->lvl_rows + o, A->row,
                       A->col, A->v, (const double *)A->F64, (const double *)A->w, out);
  }
}

// bcsr_parallel_lusgs: forward over the colour sets, then backward
template <int B>
void plusgs_b(subfv_bcsr *A, const double *in, double *out) {
  hipStream_t s = A->ctx->stream;
  for (int cs = 1; cs <= A->nsets; cs++) {
    const int o = A->set_off[cs - 1], cnt = A->set_off[cs] - o;
    if (cnt)
      hipLaunchKernelGGL((k_plusgs_set<B, false>), dim3(nblk(cnt, 256)), dim3(256), 0, s, cnt, A->set_rows + o, cs,
                         A->color, A->row, A->col, A->v, (const double *)A->F64, in, A->w);
  }
  for (int cs = A->nsets; cs >= 1; cs--) {
    const int o = A->set_off[cs - 1], cnt = A->set_off[cs] - o;
    if (cnt)
      hipLaunchKernelGGL((k_plusgs_set<B, true>), dim3(nblk(cnt, 256)), dim3(256), 0, s, cnt, A->set_rows + o, cs,
                         A->color, A->row, A->col, A->v, (const double *)A->F64, (const double *)A->w, out);
  }
}

int lusgs(subfv_bcsr *A, const double *in, double *out) {
  switch (A->b) {
    case 3: lusgs_b<3>(A, in, out); break;
    case 4: lusgs_b<4>(A, in, out); break;
    default: lusgs_b<5>(A, in, out);
  }
  HIPCHK(A->ctx, hipGetLastError());
  return 0;
}

int plusgs(subfv_bcsr *A, const double *in, double *out) {
  if (A->nsets <= 0) return fail(A->ctx, SUBFV_E_INVALID, "no colour sets: call subfv_bcsr_set_colors first");
  switch (A->b) {
    case 3: plusgs_b<3>(A, in, out); break;
    case 4: plusgs_b<4>(A, in, out); break;
    default: plusgs_b<5>(A, in, out);
  }
  HIPCHK(A->ctx, hipGetLastError());
  return 0;
}

// preconditioners of the GMRES variants: 0 none, 1 block-Jacobi, 2 LU-SGS,
// 3 coloured LU-SGS
int precond(subfv_bcsr *A, int kind, bool fp32, const double *in, double *out) {
  if (kind == 1) return bjacobi(A, fp32, in, out);
  if (kind == 2) return lusgs(A, in, out);
  return plusgs(A, in, out);
}

}  // namespace

extern "C" {

int subfv_bcsr_create(subfv_ctx *c, subfv_bcsr **out, int32_t m, int32_t b, int32_t nnzb, const int32_t *row,
                      const int32_t *col, const double *v) {
  if (!c || !out || !row || !col || !v || m <= 0 || nnzb <= 0) return fail(c, SUBFV_E_INVALID, "bad BCSR arguments");
  if (b < 3 || b > 5) return fail(c, SUBFV_E_INVALID, "block size must be 3, 4 or 5");
  if (row[0] != 1 || row[m] != nnzb + 1) return fail(c, SUBFV_E_INVALID, "row(1) must be 1 and row(m+1) nnzb+1");
  std::vector<int> r0(m + 1), c0(nnzb), di(m, -1);
  for (int i = 0; i <= m; i++) r0[i] = row[i] - 1;
  for (int i = 0; i < m; i++) {
    if (r0[i + 1] < r0[i]) return fail(c, SUBFV_E_INVALID, "row offsets decrease");
    for (int j = r0[i]; j < r0[i + 1]; j++) {
      c0[j] = col[j] - 1;
      if (c0[j] < 0 || c0[j] >= m) return fail(c, SUBFV_E_INVALID, "column id out of range");
      if (c0[j] == i && di[i] < 0) di[i] = j;  // first diagonal block, as bcsr_diag's exit
    }
    if (di[i] < 0) return fail(c, SUBFV_E_INVALID, "block row without a diagonal block");
  }
  // dependency levels of bcsr_lusgs's sweeps: row i waits for the rows
  // col(j) < i met by j = row(i), ... while col(j) < i
  std::vector<int> lvl(m, 0);
  int nlev = 0;
  for (int i = 0; i < m; i++) {
    int l = 0;
    for (int j = r0[i]; j < r0[i + 1] && c0[j] < i; j++) l = std::max(l, lvl[c0[j]] + 1);
    lvl[i] = l;
    nlev = std::max(nlev, l + 1);
  }
  std::vector<int> loff(nlev + 1, 0), lrows(m);
  for (int i = 0; i < m; i++) loff[lvl[i] + 1]++;
  for (int l = 0; l < nlev; l++) loff[l + 1] += loff[l];
  {
    std::vector<int> fill(loff.begin(), loff.end() - 1);
    for (int i = 0; i < m; i++) lrows[fill[lvl[i]]++] = i;
  }
  HIPCHK(c, hipSetDevice(c->device));
  auto *A = new subfv_bcsr();
  A->ctx = c;
  A->m = m;
  A->lvl_off = loff;
  A->b = b;
  A->nnzb = nnzb;
  A->n = (long)m * b;
  size_t nv = (size_t)nnzb * b * b;
  std::vector<float> v32(nv);
  for (size_t q = 0; q < nv; q++) v32[q] = (float)v[q];
  int r = 0;
  auto up = [&](auto **p, const auto *src, size_t cnt) {
    if (r) return;
    r = balloc(A, p, cnt);
    if (!r && hipMemcpy(*p, src, cnt * sizeof(**p), hipMemcpyHostToDevice) != hipSuccess)
      r = fail(c, SUBFV_E_DEVICE, "BCSR upload failed");
  };
  up(&A->row, r0.data(), r0.size());
  up(&A->col, c0.data(), c0.size());
  up(&A->diag_idx, di.data(), di.size());
  up(&A->lvl_rows, lrows.data(), lrows.size());
  up(&A->v, v, nv);
  up(&A->v32, v32.data(), nv);
  if (!r) r = balloc(A, &A->vr, nv);
  if (!r) r = balloc(A, &A->vr32, nv);
  if (!r) r = balloc(A, &A->F64, (size_t)m * b * b);
  if (!r) r = balloc(A, &A->F32, (size_t)m * b * b);
  for (double **p : {&A->w, &A->r, &A->x, &A->rhs, &A->tmp, &A->zp})
    if (!r) r = balloc(A, p, (size_t)A->n);
  if (!r) r = balloc(A, &A->scal, 16);
  if (!r) r = balloc(A, &A->z, 256);
  if (!r) r = balloc(A, &A->part, kRedBlocks);
  if (!r) r = balloc(A, &A->part2, kRedBlocks);
  if (!r) r = balloc(A, &A->dcp, (size_t)(2 * kGsMax + 2) * kGsBlocks);
  if (!r) r = balloc(A, &A->dch, (size_t)kDchSize);
  if (!r) r = balloc(A, &A->perr, 1);
  if (r) {
    subfv_bcsr_destroy(A);
    return r;
  }
  dim3 g(nblk(m, 256)), bl(256);
  hipStream_t s = c->stream;
  switch (b) {
    case 3:
      hipLaunchKernelGGL((k_bcsr_factor<double, 3>), g, bl, 0, s, m, A->diag_idx, A->v, A->F64);
      hipLaunchKernelGGL((k_bcsr_factor<float, 3>), g, bl, 0, s, m, A->diag_idx, A->v, A->F32);
      break;
    case 4:
      hipLaunchKernelGGL((k_bcsr_factor<double, 4>), g, bl, 0, s, m, A->diag_idx, A->v, A->F64);
      hipLaunchKernelGGL((k_bcsr_factor<float, 4>), g, bl, 0, s, m, A->diag_idx, A->v, A->F32);
      break;
    default:
      hipLaunchKernelGGL((k_bcsr_factor<double, 5>), g, bl, 0, s, m, A->diag_idx, A->v, A->F64);
      hipLaunchKernelGGL((k_bcsr_factor<float, 5>), g, bl, 0, s, m, A->diag_idx, A->v, A->F32);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(s));
  A->h_row = std::move(r0);
  A->h_col = std::move(c0);
  A->h_diag = std::move(di);
  *out = A;
  return 0;
}

void subfv_bcsr_destroy(subfv_bcsr *A) {
  if (!A) return;
  if (A->ctx && A->ctx->stream) (void)hipStreamSynchronize(A->ctx->stream);
  if (A->cyc_exec) (void)hipGraphExecDestroy(A->cyc_exec);
  if (A->cyc_graph) (void)hipGraphDestroy(A->cyc_graph);
  for (void *p : A->allocs) (void)hipFree(p);
  delete A;
}

int subfv_bcsr_matmul(subfv_bcsr *A, const double *x, double *y) {
  if (!A || !x || !y) return SUBFV_E_INVALID;
  subfv_ctx *c = A->ctx;
  HIPCHK(c, hipMemcpyAsync(A->x, x, A->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int r = spmv(A, false, A->x, A->tmp);
  if (r) return r;
  HIPCHK(c, hipMemcpyAsync(y, A->tmp, A->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int subfv_bcsr_diag(subfv_bcsr *A, double *x, const double *rhs) {
  if (!A || !x || !rhs) return SUBFV_E_INVALID;
  subfv_ctx *c = A->ctx;
  HIPCHK(c, hipMemcpyAsync(A->rhs, rhs, A->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int r = bjacobi(A, false, A->rhs, A->tmp);
  if (r) return r;
  HIPCHK(c, hipMemcpyAsync(x, A->tmp, A->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"

namespace {

// One DCGS2 Arnoldi step (k_dc_dot, k_dc_coef, k_dc_update) against the st
// columns of V, or the cycle's closing pass (fin: k_dc_dot with z = u and
// k_dc_coef, which corrects the last column of H).
template <int KM, typename VT>
void dc_step_k(subfv_bcsr *A, int st, VT *V, long ldv, double *u, const double *z, double *H, int ldh, bool fin) {
  hipStream_t s = A->ctx->stream;
  dim3 g(kGsBlocks), b(kVecThreads);
  double *hin = A->dch + (st & 1) * (kGsMax + 2), *hout = A->dch + ((st + 1) & 1) * (kGsMax + 2);
  double *coef = A->dch + kDchCoef;
  hipLaunchKernelGGL((k_dc_dot<KM, VT>), g, b, 0, s, A->n, st, (const VT *)V, ldv, (const double *)u,
                     fin ? (const double *)u : z, A->dcp);
  hipLaunchKernelGGL((k_dc_coef<KM>), dim3(1), dim3(kDcCoefThreads), 0, s, st, (const double *)A->dcp, H, ldh,
                     (const double *)hin, hout, coef, st == 0 ? A->dch + kDchBeta0 : nullptr);
  if (!fin)
    hipLaunchKernelGGL((k_dc_update<KM, VT>), g, b, 0, s, A->n, st, V, ldv, u, z, (const double *)coef);
}

template <typename VT>
void dc_step(subfv_bcsr *A, int st, VT *V, long ldv, double *u, const double *z, double *H, int ldh, bool fin) {
  switch ((st + 3) / 4) {
    case 0:
    case 1: dc_step_k<4>(A, st, V, ldv, u, z, H, ldh, fin); break;
    case 2: dc_step_k<8>(A, st, V, ldv, u, z, H, ldh, fin); break;
    case 3: dc_step_k<12>(A, st, V, ldv, u, z, H, ldh, fin); break;
    case 4: dc_step_k<16>(A, st, V, ldv, u, z, H, ldh, fin); break;
    case 5: dc_step_k<20>(A, st, V, ldv, u, z, H, ldh, fin); break;
    case 6: dc_step_k<24>(A, st, V, ldv, u, z, H, ldh, fin); break;
    case 7: dc_step_k<28>(A, st, V, ldv, u, z, H, ldh, fin); break;
    default: dc_step_k<kGsMax>(A, st, V, ldv, u, z, H, ldh, fin);
  }
}

// The restarted GMRES loops of bcsr_gmres (kind 0, :194-290, modifiedArnoldi
// :319-340), bcsr_gmres_diag (1, :1592-1691, modifiedArnoldi_diag
// :1568-1590), bcsr_gmres_lusgs (2, :447-548, modifiedArnoldi_lusgs :642-665)
// and bcsr_gmres_parallel_lusgs (3, :885-1037, :583-611): the Arnoldi process
// on the device, the Givens rotations and back substitution on the host in
// the reference order.
int gmres_impl(subfv_bcsr *A, int kind, double *x_out, const double *rhs, double tol, int32_t mr,
               int32_t precision, int32_t *conv_out, double *stats, double *ws = nullptr) {
  if (!A || !x_out || !rhs || mr < 1 || mr > 200 || kind < 0 || kind > 3) return SUBFV_E_INVALID;
  subfv_ctx *c = A->ctx;
  hipStream_t s = c->stream;
  const bool fp32 = (precision & 1) != 0;
  if (fp32 && kind >= 2) return fail(c, SUBFV_E_INVALID, "the LU-SGS variants run in FP64 only");
  if (kind == 3 && A->nsets <= 0) return fail(c, SUBFV_E_INVALID, "no colour sets: call subfv_bcsr_set_colors first");
  const int maxit = kind == 1 ? 40 : 20;
  const long n = A->n;
  const int ldh = mr + 1;
  // the Krylov basis V (n x (mr + 1)) and H ((mr + 1) x mr): the handle's
  // own workspace, or the caller-given ws that holds both
  int r = 0;
  // V's columns are ldv = n rounded up to even apart (the DCGS2 passes read
  // two consecutive entries at once)
  const long ldv = n + (n & 1);
  double *Vb = ws, *Hb = ws ? ws + (size_t)ldv * (mr + 1) : nullptr;
  if (!ws) {
    if ((r = ensure_workspace(A, mr))) return r;
    Vb = A->V;
    Hb = A->H;
  }
  auto Vc = [&](int j) { return Vb + (long)(j - 1) * ldv; };  // 1-based column
  // Up to kGsMax columns the Arnoldi process is DCGS2 (k_dc_*): the columns
  // orthogonalised twice, the second pass of each delayed into the next
  // step's.  FP32-inner block-Jacobi GMRES then keeps its Krylov basis in
  // FP32 (in the same workspace, half of it): the orthogonalisation reads
  // half the bytes; the products and sums stay FP64, so do the provisional
  // vector and each restart's residual.
  const bool dc = mr <= kGsMax;
  const bool vf32 = dc && fp32 && kind == 1;
  float *Vf = reinterpret_cast<float *>(Vb);
  double *du = A->zp, *dz = A->r;  // the provisional vector and A M^-1 of it (r is free during the cycle)
  auto Hd = [&](int i, int j) { return Hb + (i - 1) + (long)ldh * (j - 1); };
  dim3 vg(kRedBlocks), vb(kVecThreads);
  long nspmv = 0;
  HIPCHK(c, hipMemcpyAsync(A->rhs, rhs, n * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemsetAsync(A->x, 0, n * sizeof(double), s));
  if ((r = spmv(A, false, A->x, A->tmp))) return r;
  nspmv++;
  hipLaunchKernelGGL(k_sub, vg, vb, 0, s, n, A->rhs, A->tmp, A->r);
  double rho = 0.0;
  if ((r = norm_host(A, A->r, &rho))) return r;
  int conv = 0, jrest = 0;
  std::vector<double> H((size_t)ldh * mr), y(mr + 2), cs(mr + 2), sn(mr + 2), z(mr + 2);
  if (rho < tol) {
    conv = 1;
  } else {
    for (jrest = 1; jrest <= maxit; jrest++) {
      std::fill(y.begin(), y.end(), 0.0);
      y[1] = rho;
      HIPCHK(c, hipMemcpyAsync(A->scal, &rho, sizeof(double), hipMemcpyHostToDevice, s));
      if (dc) {
        if (kind == 1 && (r = ensure_rprec(A))) return r;
        // the cycle's launches: u = r / rho (rho in A->scal); step st: z =
        // A M^-1 u, then V(:, st) and the next u; the last column's second pass
        auto cycle = [&]() -> int {
          int rr;
          hipLaunchKernelGGL(k_scale_div, vg, vb, 0, s, n, A->r, du, A->scal, 0, (double *)nullptr);
          for (int st = 0; st < mr; st++) {
            if (kind == 0) {
              if ((rr = spmv(A, fp32, du, dz))) return rr;
            } else if (kind == 1) {
              // A M^-1 u as one product with A D^-1 (A->vr)
              if ((rr = spmv_v(A, fp32, A->vr, A->vr32, du, dz))) return rr;
            } else {
              if ((rr = precond(A, kind, fp32, du, A->tmp))) return rr;
              if ((rr = spmv(A, fp32, A->tmp, dz))) return rr;
            }
            if (vf32)
              dc_step<float>(A, st, Vf, ldv, du, dz, Hb, ldh, false);
            else
              dc_step<double>(A, st, Vb, ldv, du, dz, Hb, ldh, false);
          }
          if (vf32)
            dc_step<float>(A, mr, Vf, ldv, du, dz, Hb, ldh, true);
          else
            dc_step<double>(A, mr, Vb, ldv, du, dz, Hb, ldh, true);
          return 0;
        };
        // captured once per restart length, preconditioner, precision and
        // basis, then replayed every restart (no memory operations inside)
        if (!(A->cyc_exec && A->cyc_mr == mr && A->cyc_kind == kind && A->cyc_fp32 == (int)fp32 &&
              A->cyc_v == (const void *)Vb)) {
          if (A->cyc_exec) (void)hipGraphExecDestroy(A->cyc_exec);
          if (A->cyc_graph) (void)hipGraphDestroy(A->cyc_graph);
          A->cyc_exec = nullptr;
          A->cyc_graph = nullptr;
          A->cyc_mr = -1;
          HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
          const int rc = cycle();
          hipGraph_t g = nullptr;
          const hipError_t ec = hipStreamEndCapture(s, &g);
          if (rc || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : fail(c, SUBFV_E_DEVICE, std::string("GMRES cycle capture: ") + hipGetErrorString(ec));
          }
          A->cyc_graph = g;
          HIPCHK(c, hipGraphInstantiate(&A->cyc_exec, g, nullptr, nullptr, 0));
          A->cyc_mr = mr;
          A->cyc_kind = kind;
          A->cyc_fp32 = (int)fp32;
          A->cyc_v = (const void *)Vb;
        }
        HIPCHK(c, hipGraphLaunch(A->cyc_exec, s));
        nspmv += mr;
      } else {
        hipLaunchKernelGGL(k_scale_div, vg, vb, 0, s, n, A->r, Vc(1), A->scal, 0, (double *)nullptr);
      }
      for (int j = 2; !dc && j <= mr + 1; j++) {
        if (kind == 0) {
          if ((r = spmv(A, fp32, Vc(j - 1), Vc(j)))) return r;
        } else {
          if ((r = precond(A, kind, fp32, Vc(j - 1), A->tmp))) return r;
          if ((r = spmv(A, fp32, A->tmp, Vc(j)))) return r;
        }
        nspmv++;
        // modified Gram-Schmidt against V(1..j-1): each step sums the
        // previous step's dot (h(i-1, j-1)) itself, so partial buffers alternate
        double *w = Vc(j);
        double *pa = A->part, *pb = A->part2;
        hipLaunchKernelGGL(k_mgs_chain, vg, vb, 0, s, n, w, (const double *)nullptr, (const double *)nullptr,
                           (double *)nullptr, (const double *)Vc(1), pa);
        for (int i = 2; i <= j; i++) {
          hipLaunchKernelGGL(k_mgs_chain, vg, vb, 0, s, n, w, (const double *)Vc(i - 1), (const double *)pa,
                             Hd(i - 1, j - 1), (const double *)(i < j ? Vc(i) : nullptr), pb);
          std::swap(pa, pb);
        }
        hipLaunchKernelGGL(k_scale_norm, vg, vb, 0, s, n, w, (const double *)pa, Hd(j, j - 1));
      }
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipMemcpyAsync(H.data(), Hb, H.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      if (dc) {
        // the first column is (r / rho) / beta0: r = rho beta0 V(:, 1)
        double beta0 = 1.0;
        HIPCHK(c, hipMemcpyAsync(&beta0, A->dch + kDchBeta0, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        y[1] = rho * beta0;
      }
      HIPCHK(c, hipStreamSynchronize(s));
      auto h = [&](int i, int j) -> double & { return H[(i - 1) + (size_t)ldh * (j - 1)]; };
      int nr = mr, i, exited = 0;
      for (i = 1; i <= mr; i++) {
        for (int k = 2; k <= i; k++) {
          double t = h(k - 1, i);
          h(k - 1, i) = cs[k - 1] * h(k - 1, i) + sn[k - 1] * h(k, i);
          h(k, i) = -sn[k - 1] * t + cs[k - 1] * h(k, i);
        }
        double delta = sqrt(h(i, i) * h(i, i) + h(i + 1, i) * h(i + 1, i));
        cs[i] = h(i, i) / delta;
        sn[i] = h(i + 1, i) / delta;
        h(i, i) = cs[i] * h(i, i) + sn[i] * h(i + 1, i);
        for (int k = i + 1; k <= mr + 1; k++) h(k, i) = 0.0;
        y[i + 1] = -sn[i] * y[i];
        y[i] = cs[i] * y[i];
        rho = fabs(y[i + 1]);
        if (rho < tol) {
          nr = i + 1;
          exited = 1;
          break;
        }
      }
      if (!exited) i = mr + 1;
      std::fill(z.begin(), z.end(), 0.0);
      if (i >= mr) {
        nr = mr;
        z[nr - 1] = y[nr] / h(nr, nr);
      }
      for (int k = nr - 1; k >= 1; k--) {
        z[k - 1] = y[k];
        for (int l = k + 1; l <= nr; l++) z[k - 1] = z[k - 1] - h(k, l) * z[l - 1];
        z[k - 1] = z[k - 1] / h(k, k);
      }
      HIPCHK(c, hipMemcpyAsync(A->z, z.data(), nr * sizeof(double), hipMemcpyHostToDevice, s));
      if (kind == 0) {
        hipLaunchKernelGGL(k_add_combination, vg, vb, 0, s, n, nr, (const double *)Vb, ldv, (const double *)A->z, A->x);
      } else {
        if (vf32)
          hipLaunchKernelGGL(k_combine<float>, vg, vb, 0, s, n, nr, (const float *)Vf, ldv, (const double *)A->z, A->r);
        else
          hipLaunchKernelGGL(k_combine<double>, vg, vb, 0, s, n, nr, (const double *)Vb, ldv, (const double *)A->z,
                             A->r);
        if ((r = precond(A, kind, fp32, A->r, A->zp))) return r;
        hipLaunchKernelGGL(k_add, vg, vb, 0, s, n, A->x, A->zp);
      }
      if ((r = spmv(A, false, A->x, A->tmp))) return r;
      nspmv++;
      hipLaunchKernelGGL(k_sub, vg, vb, 0, s, n, A->rhs, A->tmp, A->r);
      if ((r = norm_host(A, A->r, &rho))) return r;
      if (rho < tol) {
        conv = 1;
        break;
      }
    }
  }
  HIPCHK(c, hipMemcpyAsync(x_out, A->x, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (conv_out) *conv_out = conv;
  if (stats) {
    stats[0] = (double)std::min(jrest, maxit);
    stats[1] = rho;
    stats[2] = (double)nspmv;
  }
  return conv ? 0 : fail(c, SUBFV_E_NOT_CONVERGED, "GMRES did not converge");
}

}  // namespace

extern "C" {

int subfv_bcsr_gmres_diag(subfv_bcsr *A, double *x, const double *rhs, double tol, int32_t restart,
                          int32_t precision, int32_t *conv, double *stats) {
  return gmres_impl(A, 1, x, rhs, tol, restart, precision, conv, stats);
}

// bcsr_gmres_diag with V and H in the file-scope scratch array: no device
// allocation.  Calls that overlap on one GPU share the array and race.
int subfv_bcsr_gmres_diag_noalloc(subfv_bcsr *A, double *x, const double *rhs, double tol, int32_t restart,
                                  int32_t precision, int32_t *conv, double *stats) {
  if (!A || !x || !rhs || restart < 1 || restart > 200) return SUBFV_E_INVALID;
  const size_t need = (size_t)(A->n + (A->n & 1)) * (restart + 1) + (size_t)(restart + 1) * restart;
  if (need > kNoAllocDoubles)
    return fail(A->ctx, SUBFV_E_INVALID,
                "subfv_bcsr_gmres_diag_noalloc: n' * (restart + 1) + (restart + 1) * restart = " + std::to_string(need) +
                    " doubles exceeds the " + std::to_string(kNoAllocDoubles) + "-double scratch array");
  double *ws = noalloc_get(A->ctx->device);
  if (!ws) return fail(A->ctx, SUBFV_E_DEVICE, "scratch array address not set up for this device");
  return gmres_impl(A, 1, x, rhs, tol, restart, precision, conv, stats, ws);
}

int subfv_bcsr_gmres(subfv_bcsr *A, double *x, const double *rhs, double tol, int32_t restart, int32_t precond_kind,
                     int32_t precision, int32_t *conv, double *stats) {
  return gmres_impl(A, precond_kind, x, rhs, tol, restart, precision, conv, stats);
}

int subfv_bcsr_lusgs(subfv_bcsr *A, double *x, const double *rhs) {
  if (!A || !x || !rhs) return SUBFV_E_INVALID;
  subfv_ctx *c = A->ctx;
  HIPCHK(c, hipMemcpyAsync(A->rhs, rhs, A->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int r = lusgs(A, A->rhs, A->tmp);
  if (r) return r;
  HIPCHK(c, hipMemcpyAsync(x, A->tmp, A->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int subfv_bcsr_set_colors(subfv_bcsr *A, const int32_t *color, int32_t *n_sets) {
  if (!A) return SUBFV_E_INVALID;
  subfv_ctx *c = A->ctx;
  const int m = A->m;
  std::vector<int> col1(m);
  if (color) {
    for (int i = 0; i < m; i++) {
      if (color[i] < 1) return fail(c, SUBFV_E_INVALID, "colours are 1-based");
      col1[i] = color[i];
    }
  } else {
    // greedy colouring of the symmetric closure of the block pattern, rows in order
    std::vector<std::vector<int>> adj(m);
    for (int i = 0; i < m; i++)
      for (int j = A->h_row[i]; j < A->h_row[i + 1]; j++) {
        const int k = A->h_col[j];
        if (k != i) {
          adj[i].push_back(k);
          adj[k].push_back(i);
        }
      }
    std::vector<int> mark(m + 2, -1);
    for (int i = 0; i < m; i++) {
      col1[i] = 0;
    }
    for (int i = 0; i < m; i++) {
      for (int k : adj[i])
        if (col1[k] > 0) mark[col1[k]] = i;
      int cc = 1;
      while (mark[cc] == i) cc++;
      col1[i] = cc;
    }
  }
  int ns = 0;
  for (int i = 0; i < m; i++) ns = std::max(ns, col1[i]);
  std::vector<int> off(ns + 1, 0), rows(m);
  for (int i = 0; i < m; i++) off[col1[i]]++;
  for (int k = 0; k < ns; k++) off[k + 1] += off[k];
  {
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int i = 0; i < m; i++) rows[fill[col1[i] - 1]++] = i;
  }
  for (int **p : {&A->set_rows, &A->color})
    if (*p) {
      (void)hipFree(*p);
      A->allocs.erase(std::find(A->allocs.begin(), A->allocs.end(), (void *)*p));
      *p = nullptr;
    }
  int r;
  if ((r = balloc(A, &A->set_rows, (size_t)m))) return r;
  if ((r = balloc(A, &A->color, (size_t)m))) return r;
  HIPCHK(c, hipMemcpy(A->set_rows, rows.data(), m * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(A->color, col1.data(), m * sizeof(int), hipMemcpyHostToDevice));
  A->set_off = off;
  A->nsets = ns;
  A->cyc_mr = -1;  // a captured GMRES cycle holds the old colour sets
  if (n_sets) *n_sets = ns;
  return 0;
}

int subfv_bcsr_parallel_lusgs(subfv_bcsr *A, double *x, const double *rhs) {
  if (!A || !x || !rhs) return SUBFV_E_INVALID;
  subfv_ctx *c = A->ctx;
  HIPCHK(c, hipMemcpyAsync(A->rhs, rhs, A->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int r = plusgs(A, A->rhs, A->tmp);
  if (r) return r;
  HIPCHK(c, hipMemcpyAsync(x, A->tmp, A->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// bcsr_pcgm (sparse_bcsr_linear_module.F90:1110-1169): CG from the given x
// with lu_solve(3, ...) on each diagonal block as preconditioner; dot
// products by the device reductions, scalars on the host.
int subfv_bcsr_pcgm(subfv_bcsr *A, double *x_io, const double *rhs, double tol, double *stats) {
  if (!A || !x_io || !rhs) return SUBFV_E_INVALID;
  subfv_ctx *c = A->ctx;
  hipStream_t s = c->stream;
  const long n = A->n;
  int r;
  if (!A->p && (r = balloc(A, &A->p, (size_t)n))) return r;
  if (!A->pmr && (r = balloc(A, &A->pmr, (size_t)n))) return r;
  dim3 vg(kRedBlocks), vb(kVecThreads), mg(nblk(A->m, 256)), mb(256);
  auto dot = [&](double *a, const double *b, double *out) -> int {
    mgs_step(A, a, nullptr, nullptr, b, A->scal);
    HIPCHK(c, hipMemcpyAsync(out, A->scal, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return 0;
  };
  auto prec = [&]() {
    hipLaunchKernelGGL(k_pcg_prec, mg, mb, 0, s, A->m, A->b, (const int *)A->diag_idx, (const double *)A->v,
                       (const double *)A->r, A->pmr, A->perr);
  };
  HIPCHK(c, hipMemcpyAsync(A->x, x_io, n * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(A->rhs, rhs, n * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemsetAsync(A->perr, 0, sizeof(int), s));
  HIPCHK(c, hipMemsetAsync(A->pmr, 0, n * sizeof(double), s));
  if ((r = spmv(A, false, A->x, A->tmp))) return r;
  hipLaunchKernelGGL(k_sub, vg, vb, 0, s, n, A->rhs, A->tmp, A->r);
  prec();
  double rdr = 0.0, rpdrp = 0.0;
  if ((r = dot(A->pmr, A->r, &rdr))) return r;
  long k = 0;
  bool conv;
  if (sqrt(rdr) < tol) {
    conv = true;
    rpdrp = rdr;
  } else {
    HIPCHK(c, hipMemcpyAsync(A->p, A->pmr, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    for (k = 1; k <= n; k++) {
      if ((r = spmv(A, false, A->p, A->tmp))) return r;
      double pap = 0.0;
      if ((r = dot(A->p, A->tmp, &pap))) return r;
      const double alpha = rdr / pap;
      hipLaunchKernelGGL(k_axpy, vg, vb, 0, s, n, alpha, (const double *)A->p, A->x);
      hipLaunchKernelGGL(k_axpy, vg, vb, 0, s, n, -alpha, (const double *)A->tmp, A->r);
      prec();
      if ((r = dot(A->pmr, A->r, &rpdrp))) return r;
      if (sqrt(rpdrp) <= tol) {
        double rr = 0.0;
        if ((r = dot(A->r, A->r, &rr))) return r;
        if (sqrt(rr) <= tol) break;
      }
      hipLaunchKernelGGL(k_xpby, vg, vb, 0, s, n, (const double *)A->pmr, rpdrp / rdr, A->p);
      rdr = rpdrp;
    }
    conv = !(sqrt(rpdrp) > tol);
  }
  int herr = 0;
  HIPCHK(c, hipMemcpyAsync(&herr, A->perr, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(x_io, A->x, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (stats) {
    stats[0] = (double)std::min(k, n);
    stats[1] = sqrt(rpdrp);
  }
  if (herr) return fail(c, SUBFV_E_INVALID, "PCG: a diagonal block is singular (lu_solve stop)");
  return conv ? 0 : fail(c, SUBFV_E_NOT_CONVERGED, "PCG did not converge");
}

// New block values v(b, b, nnzb) for the matrix's pattern: device copy, the
// FP32 copy and the block-Jacobi factors refreshed (no allocation).
int subfv_bcsr_set_values(subfv_bcsr *A, const double *v) {
  if (!A || !v) return SUBFV_E_INVALID;
  subfv_ctx *c = A->ctx;
  hipStream_t s = c->stream;
  const long nv = (long)A->nnzb * A->b * A->b;
  HIPCHK(c, hipMemcpyAsync(A->v, v, nv * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_to_f32, dim3(nblk(nv, 256)), dim3(256), 0, s, nv, (const double *)A->v, A->v32);
  A->vr_ok = false;
  dim3 g(nblk(A->m, 256)), bl(256);
  switch (A->b) {
    case 3:
      hipLaunchKernelGGL((k_bcsr_factor<double, 3>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F64);
      hipLaunchKernelGGL((k_bcsr_factor<float, 3>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F32);
      break;
    case 4:
      hipLaunchKernelGGL((k_bcsr_factor<double, 4>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F64);
      hipLaunchKernelGGL((k_bcsr_factor<float, 4>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F32);
      break;
    default:
      hipLaunchKernelGGL((k_bcsr_factor<double, 5>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F64);
      hipLaunchKernelGGL((k_bcsr_factor<float, 5>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F32);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(s));
  return 0;
}

// add_block_to_bcsr (:56-71): v(:,:,k) += M for the first k of row i with
// col(k) == j (1-based i, j); nothing when the block is not stored.  The
// block-Jacobi factors are refreshed when a diagonal block changes.
int subfv_bcsr_add_block(subfv_bcsr *A, int32_t i, int32_t j, const double *M) {
  if (!A || !M || i < 1 || i > A->m || j < 1 || j > A->m) return SUBFV_E_INVALID;
  subfv_ctx *c = A->ctx;
  hipStream_t s = c->stream;
  long k = -1;
  for (int q = A->h_row[i - 1]; q < A->h_row[i]; q++)
    if (A->h_col[q] == j - 1) {
      k = q;
      break;
    }
  if (k < 0) return 0;
  const int bb = A->b * A->b;
  HIPCHK(c, hipMemcpyAsync(A->z, M, bb * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_add_block, dim3(1), dim3(64), 0, s, bb, k, (const double *)A->z, A->v, A->v32);
  A->vr_ok = false;
  if (k == A->h_diag[i - 1]) {
    dim3 g(nblk(A->m, 256)), bl(256);
    switch (A->b) {
      case 3:
        hipLaunchKernelGGL((k_bcsr_factor<double, 3>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F64);
        hipLaunchKernelGGL((k_bcsr_factor<float, 3>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F32);
        break;
      case 4:
        hipLaunchKernelGGL((k_bcsr_factor<double, 4>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F64);
        hipLaunchKernelGGL((k_bcsr_factor<float, 4>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F32);
        break;
      default:
        hipLaunchKernelGGL((k_bcsr_factor<double, 5>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F64);
        hipLaunchKernelGGL((k_bcsr_factor<float, 5>), g, bl, 0, s, A->m, A->diag_idx, A->v, A->F32);
    }
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"

// ----------------------------------------------------------------- batches
namespace {

void batch_drop_graph(subfv_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->graph_ok = false;
}

int bfail(subfv_batch *b, int code, const std::string &m) {
  b->err = m;
  return code;
}

#define BHIP(b, expr)                                                                    \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) return bfail((b), SUBFV_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

template <class T>
int grow(subfv_batch *b, T **p, size_t *cap, size_t n) {
  if (n <= *cap && *p) return 0;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  size_t c = std::max<size_t>(n, 64) * 2;
  BHIP(b, hipMalloc((void **)p, c * sizeof(T)));
  *cap = c;
  b->graph_ok = false;  // captured kernel arguments point at the old buffer
  return 0;
}

// the argument blocks of one instance, as launch_flux / launch_gather /
// launch_update set them for a first-order forward-Euler step
void fill_desc(subfv_ctx *c, InstDesc &d) {
  std::memset(&d, 0, sizeof(d));
  d.fa.m = c->dm;
  d.fa.bc = c->bc;
  d.fa.U = nullptr;
  d.fa.G = nullptr;
  d.fa.P = c->P;
  d.fa.vp = nullptr;
  d.fa.err = c->derr;
  d.fa.t = &c->st->t;
  d.fa.second_order = 0;
  d.fa.nblocks = c->dm.ngroups;
  d.ga.m = c->dm;
  d.ga.P = c->P;
  d.ga.P2 = nullptr;
  d.ga.rhs = c->rhs;
  d.ga.rhs2 = nullptr;
  d.ga.sl = c->sl;
  d.ga.blk = c->blk;
  d.ga.st = c->st;
  d.ga.fixed_dt = 0;
  d.ua.m = c->dm;
  d.ua.U = nullptr;
  d.ua.U0 = nullptr;
  d.ua.Uout = nullptr;
  d.ua.rhs = c->rhs;
  d.ua.rhs2 = nullptr;
  d.ua.sl = c->sl;
  d.ua.diag = nullptr;
  d.ua.rdiff = nullptr;
  d.ua.blk = c->blk;
  d.ua.st = c->st;
  d.ua.a1 = 1.0;
  d.ua.a2 = 1.0;
  d.ua.w0 = 0.0;
  d.ua.w1 = 1.0;
  d.ua.flux_limit = 0;
  d.ua.advance_t = 1;
  for (int i = 0; i < 3; i++) d.U[i] = c->U[i];
  d.cur0 = c->cur;
  d.ngroups = c->dm.ngroups;
  d.ncblk = (int)nblk(c->dm.C, kCellThreads);
}

template <int S>
void launch_phase(subfv_batch *b, int n, int phase) {
  const int per_xcd = (b->total_g + 7) / 8;
  hipLaunchKernelGGL(k_node_flux_multi<S>, dim3(8 * per_xcd), dim3(kGrpThreads), 0, b->stream,
                     (const InstDesc *)b->d_desc, (const int *)b->d_gtab, b->total_g, per_xcd, phase);
  hipLaunchKernelGGL(k_gather_multi, dim3(b->total_c), dim3(kCellThreads), 0, b->stream, (const InstDesc *)b->d_desc,
                     (const int *)b->d_ctab, b->total_c);
  hipLaunchKernelGGL(k_finalize_dt_multi, dim3(n), dim3(kCellThreads), 0, b->stream, (const InstDesc *)b->d_desc);
  hipLaunchKernelGGL(k_update_multi, dim3(b->total_c), dim3(kCellThreads), 0, b->stream, (const InstDesc *)b->d_desc,
                     (const int *)b->d_ctab, b->total_c, phase);
  hipLaunchKernelGGL(k_finalize_omega_multi, dim3(n), dim3(kCellThreads), 0, b->stream, (const InstDesc *)b->d_desc);
  hipLaunchKernelGGL(k_update_fixup_multi, dim3(b->total_c), dim3(kCellThreads), 0, b->stream,
                     (const InstDesc *)b->d_desc, (const int *)b->d_ctab, b->total_c, phase);
}

static_assert(kGrpThreads == kCellThreads, "the persistent kernel runs node groups and cell blocks alike");

// n_steps steps in one cooperative launch of k_batch_persistent
template <int S>
int launch_persistent(subfv_batch *b, int n, int n_steps) {
  auto kern = k_batch_persistent<S>;
  int per_cu = 0, cus = 0;
  BHIP(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kern, kGrpThreads, 0));
  BHIP(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  if (b->wgs_per_cu > 0) per_cu = std::min(per_cu, b->wgs_per_cu);
  // every XCD slice of the queues needs workgroups: a multiple of 8
  const int grid = (per_cu * cus) / 8 * 8;
  if (grid < 8) return bfail(b, SUBFV_E_DEVICE, "persistent batch kernel: no room for a resident grid");
  const size_t nq = 32ull * n_steps + 2ull * n_steps * n;
  int r;
  if ((r = grow(b, &b->d_q, &b->cap_q, nq))) return r;
  BHIP(b, hipMemsetAsync(b->d_q, 0, nq * sizeof(unsigned), b->stream));
  PersistArgs a;
  a.d = b->d_desc;
  a.gtab = b->d_gtab;
  a.ctab = b->d_ctab;
  a.total_g = b->total_g;
  a.total_c = b->total_c;
  a.n_inst = n;
  a.n_steps = n_steps;
  a.qctr = b->d_q;
  a.done = b->d_q + 32ull * n_steps;
  void *args[] = {&a};
  BHIP(b, hipLaunchCooperativeKernel((const void *)kern, dim3(grid), dim3(kGrpThreads), args, 0, b->stream));
  return 0;
}

int launch_persistent_any(subfv_batch *b, int n, int n_steps) {
  switch (b->scheme) {
    case kMultiPoint: return launch_persistent<kMultiPoint>(b, n, n_steps);
    case kMultiPointIso: return launch_persistent<kMultiPointIso>(b, n, n_steps);
    case kThreeWave: return launch_persistent<kThreeWave>(b, n, n_steps);
    case kTwoWave: return launch_persistent<kTwoWave>(b, n, n_steps);
    case kModifiedThreeWave: return launch_persistent<kModifiedThreeWave>(b, n, n_steps);
    case kMultiPointPressure: return launch_persistent<kMultiPointPressure>(b, n, n_steps);
    default: return launch_persistent<kMultiPointPressurePh>(b, n, n_steps);
  }
}

void launch_phase_any(subfv_batch *b, int n, int phase) {
  switch (b->scheme) {
    case kMultiPoint: launch_phase<kMultiPoint>(b, n, phase); break;
    case kMultiPointIso: launch_phase<kMultiPointIso>(b, n, phase); break;
    case kThreeWave: launch_phase<kThreeWave>(b, n, phase); break;
    case kTwoWave: launch_phase<kTwoWave>(b, n, phase); break;
    case kModifiedThreeWave: launch_phase<kModifiedThreeWave>(b, n, phase); break;
    case kMultiPointPressure: launch_phase<kMultiPointPressure>(b, n, phase); break;
    default: launch_phase<kMultiPointPressurePh>(b, n, phase); break;
  }
}

}  // namespace

extern "C" {

int subfv_batch_create(subfv_batch **out, int device) {
  if (!out) return SUBFV_E_INVALID;
  *out = nullptr;
  subfv_batch *b = new subfv_batch();
  b->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return SUBFV_E_DEVICE;
  }
  *out = b;
  return 0;
}

void subfv_batch_destroy(subfv_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_gtab) (void)hipFree(b->d_gtab);
  if (b->d_ctab) (void)hipFree(b->d_ctab);
  if (b->d_q) (void)hipFree(b->d_q);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char *subfv_batch_error(const subfv_batch *b) { return b ? b->err.c_str() : "null batch"; }

int subfv_batch_set_grid(subfv_batch *b, int32_t wgs_per_cu) {
  if (!b || wgs_per_cu < 0) return SUBFV_E_INVALID;
  b->wgs_per_cu = wgs_per_cu;
  return 0;
}

int subfv_batch_run(subfv_batch *b, subfv_ctx **ctxs, int32_t n, int32_t n_steps, int32_t use_graph) {
  if (!b || !ctxs || n < 1 || n_steps < 0 || use_graph < 0 || use_graph > 2) return SUBFV_E_INVALID;
  if (n > (1 << (31 - kTabShift)) - 1) return bfail(b, SUBFV_E_INVALID, "too many instances in one batch");
  BHIP(b, hipSetDevice(b->device));
  const int scheme = ctxs[0] ? ctxs[0]->bc.scheme : 0;
  for (int i = 0; i < n; i++) {
    subfv_ctx *c = ctxs[i];
    if (!c || !c->has_mesh) return bfail(b, SUBFV_E_INVALID, "batch member without a mesh");
    if (c->device != b->device) return bfail(b, SUBFV_E_INVALID, "batch members must share the device");
    if (c->cfg.second_order || c->diff_on)
      return bfail(b, SUBFV_E_INVALID, "batches run the first-order inviscid forward-Euler loop");
    if (c->bc.scheme != scheme) return bfail(b, SUBFV_E_INVALID, "batch members must share the scheme");
    if (c->dm.ngroups > kTabMask || (long)nblk(c->dm.C, kCellThreads) > kTabMask)
      return bfail(b, SUBFV_E_INVALID, "instance too large for a batch");
  }
  std::vector<std::pair<subfv_ctx *, uint64_t>> mem(n);
  for (int i = 0; i < n; i++) mem[i] = {ctxs[i], ctxs[i]->gen};
  // the batch's stream waits for whatever each member has queued on its own
  // stream (uploads, state puts, device inits, time pushes)
  for (int i = 0; i < n; i++) {
    BHIP(b, hipEventRecord(ctxs[i]->ev, ctxs[i]->stream));
    BHIP(b, hipStreamWaitEvent(b->stream, ctxs[i]->ev, 0));
  }
  if (mem != b->members || scheme != b->scheme) {
    // a member's mesh may have changed under the same context pointer (a
    // pooled context re-used for the next request): rebuild the tables and
    // drop the graph whenever a (context, generation) pair differs
    b->members = mem;
    b->scheme = scheme;
    std::vector<int> gt, ct;
    long sum_g = 0, sum_c = 0;
    for (int i = 0; i < n; i++) {
      for (int g = 0; g < ctxs[i]->dm.ngroups; g++) gt.push_back((i << kTabShift) | g);
      const int nc = (int)nblk(ctxs[i]->dm.C, kCellThreads);
      for (int q = 0; q < nc; q++) ct.push_back((i << kTabShift) | q);
      sum_g += ctxs[i]->dm.ngroups;
      sum_c += nc;
    }
    if ((long)gt.size() != sum_g || (long)ct.size() != sum_c)
      return bfail(b, SUBFV_E_INVALID, "batch tables do not match the members' meshes");
    int r;
    if ((r = grow(b, &b->d_gtab, &b->cap_gtab, gt.size()))) return r;
    if ((r = grow(b, &b->d_ctab, &b->cap_ctab, ct.size()))) return r;
    BHIP(b, hipMemcpyAsync(b->d_gtab, gt.data(), gt.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_ctab, ct.data(), ct.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    b->total_g = (int)gt.size();
    b->total_c = (int)ct.size();
    batch_drop_graph(b);
  }
  int r;
  if ((r = grow(b, &b->d_desc, &b->cap_desc, (size_t)n))) return r;
  std::vector<InstDesc> hd(n);
  for (int i = 0; i < n; i++) fill_desc(ctxs[i], hd[i]);
  BHIP(b, hipMemcpyAsync(b->d_desc, hd.data(), n * sizeof(InstDesc), hipMemcpyHostToDevice, b->stream));
  int s = 0;
  if (use_graph == 2 && n_steps > 0) {
    if ((r = launch_persistent_any(b, n, n_steps))) return r;
    s = n_steps;
  } else if (use_graph == 1 && n_steps >= 3) {
    if (!b->graph_ok) {
      batch_drop_graph(b);
      BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      for (int ph = 0; ph < 3; ph++) launch_phase_any(b, n, ph);
      BHIP(b, hipStreamEndCapture(b->stream, &b->graph));
      BHIP(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      b->graph_ok = true;
    }
    for (; s + 3 <= n_steps; s += 3) BHIP(b, hipGraphLaunch(b->gexec, b->stream));
  }
  for (; s < n_steps; s++) launch_phase_any(b, n, s % 3);
  BHIP(b, hipGetLastError());
  for (int i = 0; i < n; i++) ctxs[i]->cur = (ctxs[i]->cur + n_steps) % 3;
  return 0;
}

int subfv_batch_wait(subfv_batch *b) {
  if (!b) return SUBFV_E_INVALID;
  BHIP(b, hipStreamSynchronize(b->stream));
  return 0;
}

}  // extern "C"
