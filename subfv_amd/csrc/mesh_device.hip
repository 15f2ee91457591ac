// The connectivity of build_mesh (src/mesh/mesh_connectivity_module.F90:12-59)
// built on the device, numbered exactly as the reference numbers it (the same
// numbering as the host build of mesh_build.cpp, which the tests hold it to).
// Every step is a data-parallel pass, a stable radix sort or a prefix sum:
//   node -> element lists     stable sort of (vertex, element) pairs
//   vertex renumbering        level-synchronous breadth-first search: each
//                             level's candidates carry their position in the
//                             reference's sequential scan (frontier vertex,
//                             its elements, their vertices); a vertex takes
//                             the smallest, and the level's new vertices are
//                             numbered in that order
//   element renumbering       first appearance over the renumbered vertex
//                             lists: an atomic minimum of list positions, sorted
//   faces                     the element faces sorted by their vertex sets
//                             (stable on file order), so a face's first
//                             occurrence is its owner and its last its right
//                             neighbour; faces numbered by (smallest vertex,
//                             first occurrence), the reference's bucket order
//   sub-faces, sub-elements   vertex-major and cell-major lists by stable sorts
//                             and per-entity passes
// The host copy of the result (HostMesh) keeps the library's host-side
// accessors and tables working unchanged.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <string>
#include <vector>

#include "mesh_build.hpp"

namespace subfv {
namespace {

constexpr int kT = 256;
inline unsigned nb(long n) { return (unsigned)std::max<long>(1, (n + kT - 1) / kT); }
#define GSTRIDE(i, n) for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

__constant__ int8_t c_faces[4][6][4] = {
    {{0, 2, 1, -1}, {0, 1, 3, -1}, {0, 3, 2, -1}, {1, 2, 3, -1}, {-1, -1, -1, -1}, {-1, -1, -1, -1}},  // tet
    {{0, 3, 2, 1}, {2, 3, 7, 6}, {6, 7, 4, 5}, {0, 1, 5, 4}, {0, 4, 7, 3}, {1, 2, 6, 5}},              // hex
    {{0, 2, 1, -1}, {3, 4, 5, -1}, {0, 1, 4, 3}, {1, 2, 5, 4}, {0, 3, 5, 2}, {-1, -1, -1, -1}},        // prism
    {{0, 3, 2, 1}, {0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {0, 4, 3, -1}, {-1, -1, -1, -1}}};     // pyramid

__device__ inline int d_nface(int kind) { return kind == 4 ? 4 : (kind == 5 ? 6 : 5); }

// the stream of the build running on this thread: buffers are allocated and
// freed in its order (hipMallocAsync / hipFreeAsync), so a build never waits
// for work of other streams on the device, e.g. solves running beside it
thread_local hipStream_t t_stream = nullptr;

// device buffer owned for the duration of one build
template <class T>
struct Buf {
  T *p = nullptr;
  size_t n = 0;
  Buf() = default;
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  ~Buf() {
    if (p) (void)hipFreeAsync(p, t_stream);
  }
  hipError_t alloc(size_t count) {
    if (p) (void)hipFreeAsync(p, t_stream);
    p = nullptr;
    n = count;
    return hipMallocAsync((void **)&p, std::max<size_t>(count, 1) * sizeof(T), t_stream);
  }
  operator T *() const { return p; }
};

struct Ctx {
  hipStream_t s;
  Buf<char> tmp;
  hipError_t err = hipSuccess;
  void chk(hipError_t e) {
    if (err == hipSuccess && e != hipSuccess) err = e;
  }
  // stable radix sort of (key, value) pairs on the key's low `bits` bits
  template <class K>
  void sort(const K *kin, K *kout, const int *vin, int *vout, long n, int bits) {
    if (n <= 0 || err != hipSuccess) return;
    size_t need = 0;
    chk(hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, (int)n, 0, bits, s));
    if (need > tmp.n) chk(tmp.alloc(need));
    chk(hipcub::DeviceRadixSort::SortPairs((void *)tmp.p, need, kin, kout, vin, vout, (int)n, 0, bits, s));
  }
  // out[0..n] = exclusive prefix sums of in[0..n-1] (in[n] is read as 0)
  void offsets(int *in, int *out, long n) {
    if (err != hipSuccess) return;
    chk(hipMemsetAsync(in + n, 0, sizeof(int), s));
    size_t need = 0;
    chk(hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, (int)(n + 1), s));
    if (need > tmp.n) chk(tmp.alloc(need));
    chk(hipcub::DeviceScan::ExclusiveSum((void *)tmp.p, need, in, out, (int)(n + 1), s));
  }
};

int bits_for(long maxval) {
  int b = 1;
  while (b < 31 && (1L << b) <= maxval) b++;
  return b;
}

// ---------------------------------------------------------------- kernels
__global__ void k_fill(int *a, long n, int v) { GSTRIDE(i, n) a[i] = v; }
__global__ void k_iota(int *a, long n) { GSTRIDE(i, n) a[i] = (int)i; }

// (vertex, element) pairs of the element lists in element order, and the
// per-vertex counts
__global__ void k_elem_pairs(int ne, const int *e_voff, const int *e_vert, int *key, int *val, int *cnt) {
  GSTRIDE(e, ne) {
    for (int q = e_voff[e]; q < e_voff[e + 1]; q++) {
      key[q] = e_vert[q];
      val[q] = (int)e;
      atomicAdd(&cnt[e_vert[q]], 1);
    }
  }
}

__global__ void k_bfs_count(int nfront, const int *nto, int istart, const int *v_eoff, const int *v_elem,
                            const int *e_voff, int *cnt) {
  GSTRIDE(i, nfront) {
    const int v = nto[istart + i];
    int c = 0;
    for (int q = v_eoff[v]; q < v_eoff[v + 1]; q++) {
      const int e = v_elem[q];
      c += e_voff[e + 1] - e_voff[e];
    }
    cnt[i] = c;
  }
}

// every unvisited vertex met from the frontier keeps the smallest scan
// position that meets it; the first to meet it lists it
__global__ void k_bfs_cand(int nfront, const int *nto, int istart, const int *v_eoff, const int *v_elem,
                           const int *e_voff, const int *e_vert, const int *base, const int *otn, int *prio,
                           int *list, int *nlist) {
  GSTRIDE(i, nfront) {
    const int v = nto[istart + i];
    int p = base[i];
    for (int q = v_eoff[v]; q < v_eoff[v + 1]; q++) {
      const int e = v_elem[q];
      for (int r = e_voff[e]; r < e_voff[e + 1]; r++, p++) {
        const int w = e_vert[r];
        if (otn[w] < 0) {
          const int old = atomicMin(&prio[w], p);
          if (old == INT_MAX) list[atomicAdd(nlist, 1)] = w;
        }
      }
    }
  }
}

__global__ void k_gather_keys(int n, const int *list, const int *prio, unsigned *key) {
  GSTRIDE(i, n) key[i] = (unsigned)prio[list[i]];
}

__global__ void k_bfs_assign(int n, const int *sorted, int iloc, int *otn, int *nto, int *prio) {
  GSTRIDE(i, n) {
    const int w = sorted[i];
    otn[w] = iloc + (int)i;
    nto[iloc + i] = w;
    prio[w] = INT_MAX;
  }
}

// the lowest index of a negative entry (res starts at INT_MAX)
__global__ void k_first_negative(int n, const int *a, int *res) {
  GSTRIDE(i, n) if (a[i] < 0) atomicMin(res, (int)i);
}

__global__ void k_set_one(int *otn, int *nto, const int *res, int iloc) {
  const int w = *res;
  otn[w] = iloc;
  nto[iloc] = w;
}

// renumbered coordinates and the vertex -> element list lengths
__global__ void k_vert_new(int nv, const int *nto, const double *coord, double *vc, const int *v_eoff, int *cnt) {
  GSTRIDE(n, nv) {
    const int o = nto[n];
    for (int d = 0; d < 3; d++) vc[3 * n + d] = coord[3L * o + d];
    cnt[n] = v_eoff[o + 1] - v_eoff[o];
  }
}

__global__ void k_copy_lists(int n, const int *perm, const int *off_old, const int *list_old, const int *off_new,
                             int *list_new, const int *map) {
  GSTRIDE(i, n) {
    const int o = perm[i];
    int w = off_new[i];
    for (int q = off_old[o]; q < off_old[o + 1]; q++) list_new[w++] = map ? map[list_old[q]] : list_old[q];
  }
}

__global__ void k_map(long n, int *a, const int *map) { GSTRIDE(i, n) a[i] = map[a[i]]; }

__global__ void k_first_pos(int nv, const int *v_eoff, const int *v_elem, int *first) {
  GSTRIDE(n, nv) {
    for (int q = v_eoff[n]; q < v_eoff[n + 1]; q++) atomicMin(&first[v_elem[q]], q);
  }
}

__global__ void k_inverse(int n, const int *perm, int *inv) { GSTRIDE(i, n) inv[perm[i]] = (int)i; }

// minimal faces in file order: vertices (new ids), owner, local face, and
// their sorted vertex sets as four sort keys (triangles padded with INT_MAX)
__global__ void k_minimal_faces(int ne, const int *kind, const int *e_voff, const int *e_vert, const int *mf_off,
                                const int *eotn, int *mf_v, int *mf_nv, int *mf_left, int *mf_leftf, unsigned *k0,
                                unsigned *k1, unsigned *k2, unsigned *k3) {
  GSTRIDE(e, ne) {
    const int kd = kind[e] - 4;
    const int nf = d_nface(kind[e]);
    for (int j = 0; j < nf; j++) {
      const int f = mf_off[e] + j;
      const int8_t *lf = c_faces[kd][j];
      const int n = lf[3] < 0 ? 3 : 4;
      int k[4];
      for (int q = 0; q < 4; q++) {
        const int v = q < n ? e_vert[e_voff[e] + lf[q]] : -1;
        mf_v[4L * f + q] = v;
        k[q] = q < n ? v : INT_MAX;
      }
      for (int a = 1; a < 4; a++)  // insertion sort of four keys
        for (int b = a; b > 0 && k[b - 1] > k[b]; b--) {
          const int t = k[b];
          k[b] = k[b - 1];
          k[b - 1] = t;
        }
      mf_nv[f] = n;
      mf_left[f] = eotn[e];
      mf_leftf[f] = j;
      k0[f] = (unsigned)k[0];
      k1[f] = (unsigned)k[1];
      k2[f] = (unsigned)k[2];
      k3[f] = (unsigned)k[3];
    }
  }
}

__global__ void k_gather_u(long n, const int *idx, const unsigned *src, unsigned *dst) {
  GSTRIDE(i, n) dst[i] = src[idx[i]];
}

__device__ inline bool same4(const unsigned *k0, const unsigned *k1, const unsigned *k2, const unsigned *k3, int a,
                             int b) {
  return k0[a] == k0[b] && k1[a] == k1[b] && k2[a] == k2[b] && k3[a] == k3[b];
}

// group starts of the sorted minimal faces (flag 1 at a new vertex set)
__global__ void k_group_flags(long n, const int *s, const unsigned *k0, const unsigned *k1, const unsigned *k2,
                              const unsigned *k3, int *flag) {
  GSTRIDE(i, n) flag[i] = (i == 0 || !same4(k0, k1, k2, k3, s[i], s[i - 1])) ? 1 : 0;
}

__global__ void k_group_starts(long n, const int *flag, const int *gid_excl, int *gstart) {
  GSTRIDE(i, n) if (flag[i]) gstart[gid_excl[i]] = (int)i;
}

// per face (group): its owner (first occurrence), right neighbour (last
// occurrence when there are two or more) and the bucket order key
__global__ void k_group_info(int ng, int nmf, const int *gstart, const int *s, const unsigned *k0, const int *mf_left,
                             const int *mf_leftf, int *src, int *right, int *rightf, unsigned long long *okey) {
  GSTRIDE(g, ng) {
    const int a = gstart[g], b = g + 1 < ng ? gstart[g + 1] : nmf;
    const int f = s[a];
    src[g] = f;
    if (b - a >= 2) {
      const int l = s[b - 1];
      right[g] = mf_left[l] + 1;
      rightf[g] = mf_leftf[l];
    } else {
      right[g] = 0;
      rightf[g] = 0;
    }
    okey[g] = ((unsigned long long)k0[f] << 32) | (unsigned)f;
  }
}

__device__ inline int cmp4(const unsigned *a, unsigned b0, unsigned b1, unsigned b2, unsigned b3) {
  if (a[0] != b0) return a[0] < b0 ? -1 : 1;
  if (a[1] != b1) return a[1] < b1 ? -1 : 1;
  if (a[2] != b2) return a[2] < b2 ? -1 : 1;
  if (a[3] != b3) return a[3] < b3 ? -1 : 1;
  return 0;
}

// boundary faces: the group of the same vertex set (binary search over the
// groups in key order); the last boundary face of a set in file order wins
__global__ void k_boundary_match(int nbf, const int *bf_off, const int *bf_vert, const int *otn, int ng,
                                 const int *gstart, const int *s, const unsigned *k0, const unsigned *k1,
                                 const unsigned *k2, const unsigned *k3, int *bmax) {
  GSTRIDE(j, nbf) {
    const int n = bf_off[j + 1] - bf_off[j];
    int k[4];
    for (int q = 0; q < 4; q++) k[q] = q < n ? otn[bf_vert[bf_off[j] + q]] : INT_MAX;
    for (int a = 1; a < 4; a++)
      for (int b = a; b > 0 && k[b - 1] > k[b]; b--) {
        const int t = k[b];
        k[b] = k[b - 1];
        k[b - 1] = t;
      }
    int lo = 0, hi = ng - 1, found = -1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int f = s[gstart[mid]];
      const unsigned key[4] = {k0[f], k1[f], k2[f], k3[f]};
      const int c = cmp4(key, (unsigned)k[0], (unsigned)k[1], (unsigned)k[2], (unsigned)k[3]);
      if (c == 0) {
        found = mid;
        break;
      }
      if (c < 0)
        lo = mid + 1;
      else
        hi = mid - 1;
    }
    if (found >= 0) atomicMax(&bmax[found], (int)j);
  }
}

// faces in their final order
__global__ void k_faces_out(int ng, const int *order, const int *src, const int *right, const int *rightf,
                            const int *bmax, const int *bf_tag, const int *mf_nv, const int *mf_left,
                            const int *mf_leftf, int *f_nv, int *f_left, int *f_leftf, int *f_right, int *f_rightf,
                            int *f_src) {
  GSTRIDE(fid, ng) {
    const int g = order[fid];
    const int f = src[g];
    int r = right[g], rf = rightf[g];
    if (r <= 0 && bmax[g] >= 0) {
      r = -bf_tag[bmax[g]];
      rf = 0;
    }
    f_nv[fid] = mf_nv[f];
    f_left[fid] = mf_left[f];
    f_leftf[fid] = mf_leftf[f];
    f_right[fid] = r;
    f_rightf[fid] = rf;
    f_src[fid] = f;
  }
}

__global__ void k_face_verts(int nf, const int *f_src, const int *f_voff, const int *mf_v, int *f_vert, int *key,
                             int *val, int *vcnt) {
  GSTRIDE(fid, nf) {
    const int f = f_src[fid];
    for (int q = f_voff[fid]; q < f_voff[fid + 1]; q++) {
      const int v = mf_v[4L * f + (q - f_voff[fid])];
      f_vert[q] = v;
      key[q] = v;
      val[q] = (int)fid;
      atomicAdd(&vcnt[v], 1);
    }
  }
}

__global__ void k_elem_face(int nf, const int *f_left, const int *f_leftf, const int *f_right, const int *f_rightf,
                            const int *e_foff, int *e_face) {
  GSTRIDE(fid, nf) {
    e_face[e_foff[f_left[fid]] + f_leftf[fid]] = (int)fid;
    if (f_right[fid] > 0) e_face[e_foff[f_right[fid] - 1] + f_rightf[fid]] = (int)fid;
  }
}

__global__ void k_kind_counts(int ne, const int *kind, const int *ento, int *kn, int *vcnt, int *fcnt) {
  GSTRIDE(n, ne) {
    const int k = kind[ento[n]];
    kn[n] = k;
    vcnt[n] = k == 4 ? 4 : (k == 5 ? 8 : (k == 6 ? 6 : 5));
    fcnt[n] = d_nface(k);
  }
}

__global__ void k_sub_faces(long nsf, const int *sf_vert, const int *v_face, const int *f_left, const int *f_right,
                            int *sf_face, int *sf_le, int *sf_re, int *fkey, int *fval, int *ekey, int *eval) {
  GSTRIDE(q, nsf) {
    const int f = v_face[q];
    sf_face[q] = f;
    sf_le[q] = f_left[f];
    sf_re[q] = f_right[f];
    fkey[q] = f;
    fval[q] = (int)q;
    ekey[2 * q] = f_left[f];
    eval[2 * q] = (int)q;
    ekey[2 * q + 1] = f_right[f] > 0 ? f_right[f] - 1 : INT_MAX;
    eval[2 * q + 1] = (int)q;
  }
}

__global__ void k_elem_sf_counts(int ne, const int *e_foff, const int *e_face, const int *f_voff, int *cnt) {
  GSTRIDE(e, ne) {
    int c = 0;
    for (int q = e_foff[e]; q < e_foff[e + 1]; q++) c += f_voff[e_face[q] + 1] - f_voff[e_face[q]];
    cnt[e] = c;
  }
}

__global__ void k_se_counts(int ne, const int *e_voff, const int *e_vert, const int *e_sfoff, const int *e_sf,
                            const int *sf_vert, int *se_elem, int *se_vert, int *cnt) {
  GSTRIDE(e, ne) {
    for (int j = e_voff[e]; j < e_voff[e + 1]; j++) {
      int n = 0;
      for (int q = e_sfoff[e]; q < e_sfoff[e + 1]; q++)
        if (sf_vert[e_sf[q]] == e_vert[j]) n++;
      cnt[j] = n;
      se_elem[j] = (int)e;
      se_vert[j] = e_vert[j];
    }
  }
}

__global__ void k_se_fill(int ne, const int *e_voff, const int *e_vert, const int *e_sfoff, const int *e_sf,
                          const int *sf_vert, const int *se_sfoff, int *se_sf) {
  GSTRIDE(e, ne) {
    for (int j = e_voff[e]; j < e_voff[e + 1]; j++) {
      int loc = se_sfoff[j];
      for (int q = e_sfoff[e]; q < e_sfoff[e + 1]; q++)
        if (sf_vert[e_sf[q]] == e_vert[j]) se_sf[loc++] = e_sf[q];
    }
  }
}

__global__ void k_sf_se(int nse, const int *se_elem, const int *se_sfoff, const int *se_sf, const int *sf_le,
                        const int *sf_re, int *sf_lse, int *sf_rse, int *key, int *val, const int *se_vert, int *vcnt) {
  GSTRIDE(j, nse) {
    const int e = se_elem[j];
    for (int q = se_sfoff[j]; q < se_sfoff[j + 1]; q++) {
      const int sf = se_sf[q];
      if (sf_le[sf] == e) sf_lse[sf] = (int)j;
      if (sf_re[sf] == e + 1) sf_rse[sf] = (int)j;
    }
    key[j] = se_vert[j];
    val[j] = (int)j;
    atomicAdd(&vcnt[se_vert[j]], 1);
  }
}

__global__ void k_se_loc(int nv, const int *v_seoff, const int *v_se, int *se_loc) {
  GSTRIDE(v, nv) {
    for (int p = v_seoff[v]; p < v_seoff[v + 1]; p++) se_loc[v_se[p]] = p - v_seoff[v];
  }
}

__global__ void k_elem_tag(int nf, const int *f_left, const int *f_right, int *tag) {
  GSTRIDE(f, nf) if (f_right[f] < 0) atomicMin(&tag[f_left[f]], -f_right[f]);
}

__global__ void k_tag_fix(int ne, int *tag) { GSTRIDE(e, ne) if (tag[e] == INT_MAX) tag[e] = -1; }

// neighbours by vertex: the cells around the cell's vertices, vertex by
// vertex, first occurrence kept (written to a bound-sized scratch, then packed)
__global__ void k_nbv_bound(int ne, const int *e_voff, const int *e_vert, const int *v_eoff, int *ub) {
  GSTRIDE(e, ne) {
    int c = 0;
    for (int j = e_voff[e]; j < e_voff[e + 1]; j++) c += v_eoff[e_vert[j] + 1] - v_eoff[e_vert[j]];
    ub[e] = c;
  }
}

__global__ void k_nbv_scratch(int ne, const int *e_voff, const int *e_vert, const int *v_eoff, const int *v_elem,
                              const int *ub_off, int *scratch, int *cnt) {
  GSTRIDE(e, ne) {
    int *out = scratch + ub_off[e];
    int c = 0;
    for (int j = e_voff[e]; j < e_voff[e + 1]; j++) {
      const int v = e_vert[j];
      for (int k = v_eoff[v]; k < v_eoff[v + 1]; k++) {
        const int o = v_elem[k];
        if (o == (int)e) continue;
        bool seen = false;
        for (int t = 0; t < c && !seen; t++) seen = out[t] == o;
        if (!seen) out[c++] = o;
      }
    }
    cnt[e] = c;
  }
}

__global__ void k_nbv_pack(int ne, const int *ub_off, const int *scratch, const int *off, int *nbv) {
  GSTRIDE(e, ne) {
    for (int t = 0; t < off[e + 1] - off[e]; t++) nbv[off[e] + t] = scratch[ub_off[e] + t];
  }
}

template <class T>
hipError_t down(std::vector<T> &v, const T *d, size_t n, hipStream_t s) {
  v.resize(n);
  if (!n) return hipSuccess;
  return hipMemcpyAsync(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s);
}

int read1(Ctx &c, const int *d) {
  int h = 0;
  c.chk(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, c.s));
  c.chk(hipStreamSynchronize(c.s));
  return h;
}

}  // namespace

std::string build_device_mesh(HostMesh &m, int nv, const double *coord, int ne, const int32_t *kind,
                              const int32_t *elem_vert, int nbf, const int32_t *bf_nvert, const int32_t *bf_vert,
                              const int32_t *bf_tag, void *stream) {
  if (nv <= 0 || ne <= 0) return "mesh has no vertices or no 3-D elements";
  // ---- the msh data: element kinds and vertex lists, boundary faces
  std::vector<int> h_kind(ne), h_evoff(ne + 1, 0), h_mfoff(ne + 1, 0);
  for (int i = 0; i < ne; i++) {
    const int nk = gmsh_n_vert(kind[i]);
    if (!nk) return "unsupported element kind " + std::to_string(kind[i]);
    h_kind[i] = kind[i];
    h_evoff[i + 1] = h_evoff[i] + nk;
    h_mfoff[i + 1] = h_mfoff[i] + gmsh_n_face(kind[i]);
  }
  std::vector<int> h_evert(h_evoff[ne]);
  for (int q = 0; q < h_evoff[ne]; q++) {
    const int v = elem_vert[q] - 1;
    if (v < 0 || v >= nv) return "element vertex id out of range";
    h_evert[q] = v;
  }
  std::vector<int> h_bfoff(nbf + 1, 0), h_bfvert, h_bftag(std::max(nbf, 1), 0);
  for (int j = 0; j < nbf; j++) {
    const int n = bf_nvert[j];
    if (n < 3 || n > 4) return "boundary face with unsupported vertex count";
    if (bf_tag[j] < 0 || bf_tag[j] > 10) return "boundary tag outside 0..10 (mnbc)";
    h_bfoff[j + 1] = h_bfoff[j] + n;
    h_bftag[j] = bf_tag[j];
  }
  h_bfvert.resize(std::max(h_bfoff[nbf], 1));
  for (int q = 0; q < h_bfoff[nbf]; q++) {
    const int v = bf_vert[q] - 1;
    if (v < 0 || v >= nv) return "boundary face vertex id out of range";
    h_bfvert[q] = v;
  }
  const int nq = h_evoff[ne], nmf = h_mfoff[ne];

  t_stream = (hipStream_t)stream;
  Ctx c;
  c.s = (hipStream_t)stream;
  auto fail = [&](const char *where) { return std::string("device mesh build (") + where + "): " + hipGetErrorString(c.err); };
#define UPV(buf, vec)                                                                                       \
  do {                                                                                                      \
    c.chk(buf.alloc((vec).size()));                                                                         \
    if (!(vec).empty())                                                                                     \
      c.chk(hipMemcpyAsync(buf.p, (vec).data(), (vec).size() * sizeof((vec)[0]), hipMemcpyHostToDevice, c.s)); \
  } while (0)
  Buf<double> d_coord;
  Buf<int> d_kind, d_evoff, d_evert, d_mfoff, d_bfoff, d_bfvert, d_bftag;
  c.chk(d_coord.alloc(3L * nv));
  c.chk(hipMemcpyAsync(d_coord.p, coord, 3L * nv * sizeof(double), hipMemcpyHostToDevice, c.s));
  UPV(d_kind, h_kind);
  UPV(d_evoff, h_evoff);
  UPV(d_evert, h_evert);
  UPV(d_mfoff, h_mfoff);
  UPV(d_bfoff, h_bfoff);
  UPV(d_bfvert, h_bfvert);
  UPV(d_bftag, h_bftag);
#undef UPV
  if (c.err) return fail("upload");
  const int vbits = bits_for(nv), ebits = bits_for(ne);

  // ---- build_node_elem_neigh: vertex -> element lists, element order
  Buf<int> k1, v1, k2, v2, cnt, v_eoff, v_elem;
  c.chk(k1.alloc(nq));
  c.chk(v1.alloc(nq));
  c.chk(k2.alloc(nq));
  c.chk(cnt.alloc(nv + 1));
  c.chk(v_eoff.alloc(nv + 1));
  c.chk(v_elem.alloc(nq));
  c.chk(hipMemsetAsync(cnt.p, 0, (nv + 1) * sizeof(int), c.s));
  hipLaunchKernelGGL(k_elem_pairs, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)d_evoff, (const int *)d_evert,
                     k1.p, v1.p, cnt.p);
  c.sort<int>(k1, k2, v1, v_elem, nq, vbits);
  c.offsets(cnt, v_eoff, nv);
  if (c.err) return fail("vertex lists");

  // ---- renumber_vert_mesh: level-synchronous breadth-first search
  Buf<int> otn, nto, prio, list, slist, fcnt, fbase, nlist, res;
  Buf<unsigned> pkey, pkey2;
  c.chk(otn.alloc(nv));
  c.chk(nto.alloc(nv));
  c.chk(prio.alloc(nv));
  c.chk(list.alloc(nv));
  c.chk(slist.alloc(nv));
  c.chk(pkey.alloc(nv));
  c.chk(pkey2.alloc(nv));
  c.chk(fcnt.alloc(nv + 1));
  c.chk(fbase.alloc(nv + 1));
  c.chk(nlist.alloc(1));
  c.chk(res.alloc(1));
  hipLaunchKernelGGL(k_fill, dim3(nb(nv)), dim3(kT), 0, c.s, otn.p, (long)nv, -1);
  hipLaunchKernelGGL(k_fill, dim3(nb(nv)), dim3(kT), 0, c.s, prio.p, (long)nv, INT_MAX);
  {
    const int zero = 0;
    c.chk(hipMemcpyAsync(otn.p, &zero, sizeof(int), hipMemcpyHostToDevice, c.s));
    c.chk(hipMemcpyAsync(nto.p, &zero, sizeof(int), hipMemcpyHostToDevice, c.s));
    c.chk(hipStreamSynchronize(c.s));
  }
  int iloc = 1, istart = 0, iend = 0;
  while (iloc < nv && !c.err) {
    const int nfront = iend - istart + 1;
    int found = 0;
    if (nfront > 0) {
      hipLaunchKernelGGL(k_bfs_count, dim3(nb(nfront)), dim3(kT), 0, c.s, nfront, (const int *)nto, istart,
                         (const int *)v_eoff, (const int *)v_elem, (const int *)d_evoff, fcnt.p);
      c.offsets(fcnt, fbase, nfront);
      c.chk(hipMemsetAsync(nlist.p, 0, sizeof(int), c.s));
      hipLaunchKernelGGL(k_bfs_cand, dim3(nb(nfront)), dim3(kT), 0, c.s, nfront, (const int *)nto, istart,
                         (const int *)v_eoff, (const int *)v_elem, (const int *)d_evoff, (const int *)d_evert,
                         (const int *)fbase, (const int *)otn, prio.p, list.p, nlist.p);
      found = read1(c, nlist);
      if (found > 0) {
        hipLaunchKernelGGL(k_gather_keys, dim3(nb(found)), dim3(kT), 0, c.s, found, (const int *)list,
                           (const int *)prio, pkey.p);
        c.sort<unsigned>(pkey, pkey2, list, slist, found, 31);
        hipLaunchKernelGGL(k_bfs_assign, dim3(nb(found)), dim3(kT), 0, c.s, found, (const int *)slist, iloc, otn.p,
                           nto.p, prio.p);
        iloc += found;
      }
    }
    if (found == 0) {
      // nothing reached: restart at the lowest unvisited vertex
      const int big = INT_MAX;
      c.chk(hipMemcpyAsync(res.p, &big, sizeof(int), hipMemcpyHostToDevice, c.s));
      hipLaunchKernelGGL(k_first_negative, dim3(nb(nv)), dim3(kT), 0, c.s, nv, (const int *)otn, res.p);
      hipLaunchKernelGGL(k_set_one, dim3(1), dim3(1), 0, c.s, otn.p, nto.p, (const int *)res, iloc);
      iloc++;
    }
    istart = iend + 1;
    iend = iloc - 1;
  }
  if (c.err) return fail("vertex numbering");

  // ---- renumbered vertices and their element lists
  Buf<double> vc;
  Buf<int> v_eoff2, v_elem2;
  c.chk(vc.alloc(3L * nv));
  c.chk(v_eoff2.alloc(nv + 1));
  c.chk(v_elem2.alloc(nq));
  hipLaunchKernelGGL(k_vert_new, dim3(nb(nv)), dim3(kT), 0, c.s, nv, (const int *)nto, (const double *)d_coord, vc.p,
                     (const int *)v_eoff, cnt.p);
  c.offsets(cnt, v_eoff2, nv);
  hipLaunchKernelGGL(k_copy_lists, dim3(nb(nv)), dim3(kT), 0, c.s, nv, (const int *)nto, (const int *)v_eoff,
                     (const int *)v_elem, (const int *)v_eoff2, v_elem2.p, (const int *)nullptr);
  hipLaunchKernelGGL(k_map, dim3(nb(nq)), dim3(kT), 0, c.s, (long)nq, d_evert.p, (const int *)otn);

  // ---- renumber_elem_mesh: first appearance over the vertex lists
  Buf<int> first, eidx, ento, eotn;
  Buf<unsigned> fkey, fkey2;
  c.chk(first.alloc(ne));
  c.chk(eidx.alloc(ne));
  c.chk(ento.alloc(ne));
  c.chk(eotn.alloc(ne));
  c.chk(fkey.alloc(ne));
  c.chk(fkey2.alloc(ne));
  hipLaunchKernelGGL(k_fill, dim3(nb(ne)), dim3(kT), 0, c.s, first.p, (long)ne, INT_MAX);
  hipLaunchKernelGGL(k_first_pos, dim3(nb(nv)), dim3(kT), 0, c.s, nv, (const int *)v_eoff2, (const int *)v_elem2,
                     first.p);
  hipLaunchKernelGGL(k_iota, dim3(nb(ne)), dim3(kT), 0, c.s, eidx.p, (long)ne);
  hipLaunchKernelGGL(k_gather_u, dim3(nb(ne)), dim3(kT), 0, c.s, (long)ne, (const int *)eidx, (const unsigned *)first.p,
                     fkey.p);
  c.sort<unsigned>(fkey, fkey2, eidx, ento, ne, 31);
  hipLaunchKernelGGL(k_inverse, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)ento, eotn.p);
  if (c.err) return fail("element numbering");

  // ---- build_faces: minimal faces sorted by vertex set, stable on file order
  Buf<int> mf_v, mf_nv, mf_left, mf_leftf, sidx, sidx2;
  Buf<unsigned> kk[4], kt, kt2;
  c.chk(mf_v.alloc(4L * nmf));
  c.chk(mf_nv.alloc(nmf));
  c.chk(mf_left.alloc(nmf));
  c.chk(mf_leftf.alloc(nmf));
  c.chk(sidx.alloc(nmf));
  c.chk(sidx2.alloc(nmf));
  for (auto &b : kk) c.chk(b.alloc(nmf));
  c.chk(kt.alloc(nmf));
  c.chk(kt2.alloc(nmf));
  hipLaunchKernelGGL(k_minimal_faces, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)d_kind, (const int *)d_evoff,
                     (const int *)d_evert, (const int *)d_mfoff, (const int *)eotn, mf_v.p, mf_nv.p, mf_left.p,
                     mf_leftf.p, kk[0].p, kk[1].p, kk[2].p, kk[3].p);
  hipLaunchKernelGGL(k_iota, dim3(nb(nmf)), dim3(kT), 0, c.s, sidx.p, (long)nmf);
  for (int pass = 3; pass >= 0; pass--) {  // least significant key first
    hipLaunchKernelGGL(k_gather_u, dim3(nb(nmf)), dim3(kT), 0, c.s, (long)nmf, (const int *)sidx, (const unsigned *)kk[pass],
                       kt.p);
    c.sort<unsigned>(kt, kt2, sidx, sidx2, nmf, 31);
    std::swap(sidx.p, sidx2.p);
  }
  Buf<int> flag, gid, gstart;
  c.chk(flag.alloc(nmf + 1));
  c.chk(gid.alloc(nmf + 1));
  hipLaunchKernelGGL(k_group_flags, dim3(nb(nmf)), dim3(kT), 0, c.s, (long)nmf, (const int *)sidx,
                     (const unsigned *)kk[0], (const unsigned *)kk[1], (const unsigned *)kk[2], (const unsigned *)kk[3],
                     flag.p);
  c.offsets(flag, gid, nmf);
  const int ngr = read1(c, gid.p + nmf);
  c.chk(gstart.alloc(ngr));
  hipLaunchKernelGGL(k_group_starts, dim3(nb(nmf)), dim3(kT), 0, c.s, (long)nmf, (const int *)flag, (const int *)gid,
                     gstart.p);
  Buf<int> gsrc, gright, grightf, gidx, order, bmax;
  Buf<unsigned long long> okey, okey2;
  c.chk(gsrc.alloc(ngr));
  c.chk(gright.alloc(ngr));
  c.chk(grightf.alloc(ngr));
  c.chk(gidx.alloc(ngr));
  c.chk(order.alloc(ngr));
  c.chk(bmax.alloc(ngr));
  c.chk(okey.alloc(ngr));
  c.chk(okey2.alloc(ngr));
  hipLaunchKernelGGL(k_group_info, dim3(nb(ngr)), dim3(kT), 0, c.s, ngr, nmf, (const int *)gstart, (const int *)sidx,
                     (const unsigned *)kk[0], (const int *)mf_left, (const int *)mf_leftf, gsrc.p, gright.p, grightf.p,
                     okey.p);
  hipLaunchKernelGGL(k_iota, dim3(nb(ngr)), dim3(kT), 0, c.s, gidx.p, (long)ngr);
  c.sort<unsigned long long>(okey, okey2, gidx, order, ngr, 32 + vbits);
  hipLaunchKernelGGL(k_fill, dim3(nb(ngr)), dim3(kT), 0, c.s, bmax.p, (long)ngr, -1);
  if (nbf > 0)
    hipLaunchKernelGGL(k_boundary_match, dim3(nb(nbf)), dim3(kT), 0, c.s, nbf, (const int *)d_bfoff,
                       (const int *)d_bfvert, (const int *)otn, ngr, (const int *)gstart, (const int *)sidx,
                       (const unsigned *)kk[0], (const unsigned *)kk[1], (const unsigned *)kk[2],
                       (const unsigned *)kk[3], bmax.p);
  const int nfc = ngr;
  Buf<int> f_nv, f_voff, f_left, f_leftf, f_right, f_rightf, f_src, f_vert;
  c.chk(f_nv.alloc(nfc + 1));
  c.chk(f_voff.alloc(nfc + 1));
  c.chk(f_left.alloc(nfc));
  c.chk(f_leftf.alloc(nfc));
  c.chk(f_right.alloc(nfc));
  c.chk(f_rightf.alloc(nfc));
  c.chk(f_src.alloc(nfc));
  hipLaunchKernelGGL(k_faces_out, dim3(nb(nfc)), dim3(kT), 0, c.s, nfc, (const int *)order, (const int *)gsrc,
                     (const int *)gright, (const int *)grightf, (const int *)bmax, (const int *)d_bftag,
                     (const int *)mf_nv, (const int *)mf_left, (const int *)mf_leftf, f_nv.p, f_left.p, f_leftf.p,
                     f_right.p, f_rightf.p, f_src.p);
  c.offsets(f_nv, f_voff, nfc);
  const int nfv = read1(c, f_voff.p + nfc);
  c.chk(f_vert.alloc(nfv));
  // ---- new element storage, faces of each element
  Buf<int> kn, e_vcnt, e_fcnt, e_voff, e_vert, e_foff, e_face;
  c.chk(kn.alloc(ne));
  c.chk(e_vcnt.alloc(ne + 1));
  c.chk(e_fcnt.alloc(ne + 1));
  c.chk(e_voff.alloc(ne + 1));
  c.chk(e_vert.alloc(nq));
  c.chk(e_foff.alloc(ne + 1));
  hipLaunchKernelGGL(k_kind_counts, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)d_kind, (const int *)ento, kn.p,
                     e_vcnt.p, e_fcnt.p);
  c.offsets(e_vcnt, e_voff, ne);
  c.offsets(e_fcnt, e_foff, ne);
  hipLaunchKernelGGL(k_copy_lists, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)ento, (const int *)d_evoff,
                     (const int *)d_evert, (const int *)e_voff, e_vert.p, (const int *)nullptr);
  c.chk(e_face.alloc(nmf));
  hipLaunchKernelGGL(k_fill, dim3(nb(nmf)), dim3(kT), 0, c.s, e_face.p, (long)nmf, -1);
  hipLaunchKernelGGL(k_elem_face, dim3(nb(nfc)), dim3(kT), 0, c.s, nfc, (const int *)f_left, (const int *)f_leftf,
                     (const int *)f_right, (const int *)f_rightf, (const int *)e_foff, e_face.p);
  // v_elem in the new element numbering
  hipLaunchKernelGGL(k_map, dim3(nb(nq)), dim3(kT), 0, c.s, (long)nq, v_elem2.p, (const int *)eotn);
  {
    // a face of three or more elements leaves an element face unset
    const int big = INT_MAX;
    c.chk(hipMemcpyAsync(res.p, &big, sizeof(int), hipMemcpyHostToDevice, c.s));
    hipLaunchKernelGGL(k_first_negative, dim3(nb(nmf)), dim3(kT), 0, c.s, nmf, (const int *)e_face, res.p);
    if (read1(c, res.p) != INT_MAX && !c.err) return "face shared by more than two elements";
  }
  if (c.err) return fail("faces");

  // ---- build_node_face_connectivity: vertex -> faces, face order
  Buf<int> fk, fv, vf_key, v_face, vfcnt, v_foff;
  c.chk(fk.alloc(nfv));
  c.chk(fv.alloc(nfv));
  c.chk(vf_key.alloc(nfv));
  c.chk(v_face.alloc(nfv));
  c.chk(vfcnt.alloc(nv + 1));
  c.chk(v_foff.alloc(nv + 1));
  c.chk(hipMemsetAsync(vfcnt.p, 0, (nv + 1) * sizeof(int), c.s));
  hipLaunchKernelGGL(k_face_verts, dim3(nb(nfc)), dim3(kT), 0, c.s, nfc, (const int *)f_src, (const int *)f_voff,
                     (const int *)mf_v, f_vert.p, fk.p, fv.p, vfcnt.p);
  c.sort<int>(fk, vf_key, fv, v_face, nfv, vbits);
  c.offsets(vfcnt, v_foff, nv);

  // ---- build_sub_faces: vertex-major ids; face and element sub-face lists
  const long nsf = nfv;
  Buf<int> sf_face, sf_le, sf_re, f_sfk, f_sfv, f_sfk2, f_sf, ek, ev, ek2, e_sf_all, e_sfcnt, e_sfoff;
  c.chk(sf_face.alloc(nsf));
  c.chk(sf_le.alloc(nsf));
  c.chk(sf_re.alloc(nsf));
  c.chk(f_sfk.alloc(nsf));
  c.chk(f_sfv.alloc(nsf));
  c.chk(f_sfk2.alloc(nsf));
  c.chk(f_sf.alloc(nsf));
  c.chk(ek.alloc(2 * nsf));
  c.chk(ev.alloc(2 * nsf));
  c.chk(ek2.alloc(2 * nsf));
  c.chk(e_sf_all.alloc(2 * nsf));
  c.chk(e_sfcnt.alloc(ne + 1));
  c.chk(e_sfoff.alloc(ne + 1));
  hipLaunchKernelGGL(k_sub_faces, dim3(nb(nsf)), dim3(kT), 0, c.s, nsf, (const int *)vf_key, (const int *)v_face,
                     (const int *)f_left, (const int *)f_right, sf_face.p, sf_le.p, sf_re.p, f_sfk.p, f_sfv.p, ek.p,
                     ev.p);
  c.sort<int>(f_sfk, f_sfk2, f_sfv, f_sf, nsf, bits_for(nfc));
  c.sort<int>(ek, ek2, ev, e_sf_all, 2 * nsf, 31);
  hipLaunchKernelGGL(k_elem_sf_counts, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)e_foff, (const int *)e_face,
                     (const int *)f_voff, e_sfcnt.p);
  c.offsets(e_sfcnt, e_sfoff, ne);
  if (c.err) return fail("sub-faces");

  // ---- build_sub_elems: cell-major ids and their sub-face lists
  const int nse = nq;
  Buf<int> se_elem, se_vert, se_cnt, se_sfoff, se_sf, sf_lse, sf_rse, sek, sev, sek2, v_se, vscnt, v_seoff, se_loc;
  c.chk(se_elem.alloc(nse));
  c.chk(se_vert.alloc(nse));
  c.chk(se_cnt.alloc(nse + 1));
  c.chk(se_sfoff.alloc(nse + 1));
  hipLaunchKernelGGL(k_se_counts, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)e_voff, (const int *)e_vert,
                     (const int *)e_sfoff, (const int *)e_sf_all, (const int *)vf_key, se_elem.p, se_vert.p, se_cnt.p);
  c.offsets(se_cnt, se_sfoff, nse);
  const int nsesf = read1(c, se_sfoff.p + nse);
  c.chk(se_sf.alloc(nsesf));
  hipLaunchKernelGGL(k_se_fill, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)e_voff, (const int *)e_vert,
                     (const int *)e_sfoff, (const int *)e_sf_all, (const int *)vf_key, (const int *)se_sfoff, se_sf.p);
  c.chk(sf_lse.alloc(nsf));
  c.chk(sf_rse.alloc(nsf));
  c.chk(sek.alloc(nse));
  c.chk(sev.alloc(nse));
  c.chk(sek2.alloc(nse));
  c.chk(v_se.alloc(nse));
  c.chk(vscnt.alloc(nv + 1));
  c.chk(v_seoff.alloc(nv + 1));
  c.chk(se_loc.alloc(nse));
  hipLaunchKernelGGL(k_fill, dim3(nb(nsf)), dim3(kT), 0, c.s, sf_lse.p, nsf, -1);
  hipLaunchKernelGGL(k_fill, dim3(nb(nsf)), dim3(kT), 0, c.s, sf_rse.p, nsf, -1);
  c.chk(hipMemsetAsync(vscnt.p, 0, (nv + 1) * sizeof(int), c.s));
  hipLaunchKernelGGL(k_sf_se, dim3(nb(nse)), dim3(kT), 0, c.s, nse, (const int *)se_elem, (const int *)se_sfoff,
                     (const int *)se_sf, (const int *)sf_le, (const int *)sf_re, sf_lse.p, sf_rse.p, sek.p, sev.p,
                     (const int *)se_vert, vscnt.p);
  c.sort<int>(sek, sek2, sev, v_se, nse, vbits);
  c.offsets(vscnt, v_seoff, nv);
  hipLaunchKernelGGL(k_se_loc, dim3(nb(nv)), dim3(kT), 0, c.s, nv, (const int *)v_seoff, (const int *)v_se, se_loc.p);

  // ---- compute_elem_tag and compute_neigh_by_vert
  Buf<int> e_tag, ub, ub_off, nbv_scratch, nbv_cnt, nbv_off, nbv;
  c.chk(e_tag.alloc(ne));
  hipLaunchKernelGGL(k_fill, dim3(nb(ne)), dim3(kT), 0, c.s, e_tag.p, (long)ne, INT_MAX);
  hipLaunchKernelGGL(k_elem_tag, dim3(nb(nfc)), dim3(kT), 0, c.s, nfc, (const int *)f_left, (const int *)f_right,
                     e_tag.p);
  hipLaunchKernelGGL(k_tag_fix, dim3(nb(ne)), dim3(kT), 0, c.s, ne, e_tag.p);
  c.chk(ub.alloc(ne + 1));
  c.chk(ub_off.alloc(ne + 1));
  hipLaunchKernelGGL(k_nbv_bound, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)e_voff, (const int *)e_vert,
                     (const int *)v_eoff2, ub.p);
  c.offsets(ub, ub_off, ne);
  const int nub = read1(c, ub_off.p + ne);
  c.chk(nbv_scratch.alloc(nub));
  c.chk(nbv_cnt.alloc(ne + 1));
  c.chk(nbv_off.alloc(ne + 1));
  hipLaunchKernelGGL(k_nbv_scratch, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)e_voff, (const int *)e_vert,
                     (const int *)v_eoff2, (const int *)v_elem2, (const int *)ub_off, nbv_scratch.p, nbv_cnt.p);
  c.offsets(nbv_cnt, nbv_off, ne);
  const int nnbv = read1(c, nbv_off.p + ne);
  c.chk(nbv.alloc(nnbv));
  hipLaunchKernelGGL(k_nbv_pack, dim3(nb(ne)), dim3(kT), 0, c.s, ne, (const int *)ub_off, (const int *)nbv_scratch,
                     (const int *)nbv_off, nbv.p);
  c.chk(hipGetLastError());
  if (c.err) return fail("sub-elements");

  // ---- the host copy
  m = HostMesh();
  m.nv = nv;
  m.ne = ne;
  m.nf = nfc;
  m.nsf = (int)nsf;
  m.nse = nse;
  hipStream_t s = c.s;
  c.chk(down(m.vcoord, vc.p, 3L * nv, s));
  c.chk(down(m.vert_nto, nto.p, nv, s));
  c.chk(down(m.elem_nto, ento.p, ne, s));
  c.chk(down(m.e_kind, kn.p, ne, s));
  c.chk(down(m.e_voff, e_voff.p, ne + 1, s));
  c.chk(down(m.e_vert, e_vert.p, nq, s));
  c.chk(down(m.e_foff, e_foff.p, ne + 1, s));
  c.chk(down(m.e_face, e_face.p, nmf, s));
  c.chk(down(m.e_sfoff, e_sfoff.p, ne + 1, s));
  c.chk(down(m.v_eoff, v_eoff2.p, nv + 1, s));
  c.chk(down(m.v_elem, v_elem2.p, nq, s));
  c.chk(down(m.f_voff, f_voff.p, nfc + 1, s));
  c.chk(down(m.f_vert, f_vert.p, nfv, s));
  c.chk(down(m.f_left, f_left.p, nfc, s));
  c.chk(down(m.f_leftf, f_leftf.p, nfc, s));
  c.chk(down(m.f_right, f_right.p, nfc, s));
  c.chk(down(m.f_rightf, f_rightf.p, nfc, s));
  c.chk(down(m.f_sfoff, f_voff.p, nfc + 1, s));
  c.chk(down(m.f_sf, f_sf.p, nsf, s));
  c.chk(down(m.v_foff, v_foff.p, nv + 1, s));
  c.chk(down(m.v_face, v_face.p, nsf, s));
  c.chk(down(m.sf_off, v_foff.p, nv + 1, s));
  c.chk(down(m.sf_vert, vf_key.p, nsf, s));
  c.chk(down(m.sf_face, sf_face.p, nsf, s));
  c.chk(down(m.sf_le, sf_le.p, nsf, s));
  c.chk(down(m.sf_re, sf_re.p, nsf, s));
  c.chk(down(m.sf_lse, sf_lse.p, nsf, s));
  c.chk(down(m.sf_rse, sf_rse.p, nsf, s));
  c.chk(down(m.se_elem, se_elem.p, nse, s));
  c.chk(down(m.se_vert, se_vert.p, nse, s));
  c.chk(down(m.se_loc, se_loc.p, nse, s));
  c.chk(down(m.se_sfoff, se_sfoff.p, nse + 1, s));
  c.chk(down(m.se_sf, se_sf.p, nsesf, s));
  c.chk(down(m.v_seoff, v_seoff.p, nv + 1, s));
  c.chk(down(m.v_se, v_se.p, nse, s));
  c.chk(down(m.e_tag, e_tag.p, ne, s));
  c.chk(down(m.e_nbvoff, nbv_off.p, ne + 1, s));
  c.chk(down(m.e_nbv, nbv.p, nnbv, s));
  std::vector<int> esf_all;
  c.chk(down(esf_all, e_sf_all.p, 2 * nsf, s));
  c.chk(hipStreamSynchronize(s));
  if (c.err) return fail("download");
  m.e_sf.assign(esf_all.begin(), esf_all.begin() + m.e_sfoff[ne]);
  for (int v = 0; v < nv; v++) {
    m.nmax_sfn = std::max(m.nmax_sfn, m.sf_off[v + 1] - m.sf_off[v]);
    m.nmax_sen = std::max(m.nmax_sen, m.v_seoff[v + 1] - m.v_seoff[v]);
  }
  (void)ebits;
  return "";
}

}  // namespace subfv
