// Connectivity build numbered like subfv's build_mesh.  Each step cites the
// reference routine whose numbering it reproduces; the data structures are
// flat CSR arrays instead of the reference's per-entity allocatables.
#include "mesh_build.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace subfv {

int gmsh_n_vert(int k) {
  switch (k) {
    case 4: return 4;
    case 5: return 8;
    case 6: return 6;
    case 7: return 5;
    default: return 0;
  }
}
int gmsh_n_face(int k) {
  switch (k) {
    case 4: return 4;
    case 5: return 6;
    case 6: return 5;
    case 7: return 5;
    default: return 0;
  }
}

// Element-local faces in gmsh order, as the msh 2.2 reader lists them
// (src/mesh/mesh_reading_module.F90:1163-1316); -1 pads triangles.
static const int8_t kTet[4][4] = {{0, 2, 1, -1}, {0, 1, 3, -1}, {0, 3, 2, -1}, {1, 2, 3, -1}};
static const int8_t kHex[6][4] = {{0, 3, 2, 1}, {2, 3, 7, 6}, {6, 7, 4, 5},
                                  {0, 1, 5, 4}, {0, 4, 7, 3}, {1, 2, 6, 5}};
static const int8_t kPri[5][4] = {{0, 2, 1, -1}, {3, 4, 5, -1}, {0, 1, 4, 3}, {1, 2, 5, 4}, {0, 3, 5, 2}};
static const int8_t kPyr[5][4] = {{0, 3, 2, 1}, {0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {0, 4, 3, -1}};

static const int8_t *local_face(int kind, int j) {
  switch (kind) {
    case 4: return kTet[j];
    case 5: return kHex[j];
    case 6: return kPri[j];
    default: return kPyr[j];
  }
}

struct Key4 {
  int v[4];
  int n;
};
static inline Key4 make_key(const int *vv, int n) {
  Key4 k;
  k.n = n;
  for (int i = 0; i < 4; i++) k.v[i] = (i < n) ? vv[i] : 0x7fffffff;
  std::sort(k.v, k.v + 4);
  return k;
}
static inline bool same_key(const Key4 &a, const Key4 &b) {
  return a.n == b.n && a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2] && a.v[3] == b.v[3];
}

static void csr_from_counts(std::vector<int> &off, const std::vector<int> &cnt) {
  off.assign(cnt.size() + 1, 0);
  for (size_t i = 0; i < cnt.size(); i++) off[i + 1] = off[i] + cnt[i];
}

std::string build_host_mesh(HostMesh &m, int nv, const double *coord, int ne, const int32_t *kind,
                            const int32_t *elem_vert, int nbf, const int32_t *bf_nvert,
                            const int32_t *bf_vert, const int32_t *bf_tag) {
  if (nv <= 0 || ne <= 0) return "mesh has no vertices or no 3-D elements";
  m = HostMesh();
  m.nv = nv;
  m.ne = ne;
  // ---- reader: elements and their faces (file order)
  std::vector<int> e_kind(ne), e_voff(ne + 1, 0);
  for (int i = 0; i < ne; i++) {
    int nk = gmsh_n_vert(kind[i]);
    if (!nk) return "unsupported element kind " + std::to_string(kind[i]);
    e_kind[i] = kind[i];
    e_voff[i + 1] = e_voff[i] + nk;
  }
  std::vector<int> e_vert(e_voff[ne]);
  for (int q = 0; q < e_voff[ne]; q++) {
    int v = elem_vert[q] - 1;
    if (v < 0 || v >= nv) return "element vertex id out of range";
    e_vert[q] = v;
  }
  // ---- build_node_elem_neigh (mesh_connectivity_module.F90:475-509)
  std::vector<int> cnt(nv, 0);
  for (int q = 0; q < e_voff[ne]; q++) cnt[e_vert[q]]++;
  std::vector<int> v_eoff;
  csr_from_counts(v_eoff, cnt);
  std::vector<int> v_elem(v_eoff[nv]);
  {
    std::vector<int> fill(nv, 0);
    for (int e = 0; e < ne; e++)
      for (int q = e_voff[e]; q < e_voff[e + 1]; q++) {
        int v = e_vert[q];
        v_elem[v_eoff[v] + fill[v]++] = e;
      }
  }
  // ---- renumber_vert_mesh (:851-926): breadth first from vertex 1, restart
  //      at the lowest unvisited vertex when a level discovers nothing
  std::vector<int> otn(nv, -1), nto(nv, -1);
  otn[0] = 0;
  nto[0] = 0;
  int iloc = 1, istart = 0, iend = 0;
  while (iloc < nv) {
    int before = iloc;
    for (int i = istart; i <= iend; i++) {
      int v = nto[i];
      for (int q = v_eoff[v]; q < v_eoff[v + 1]; q++) {
        int e = v_elem[q];
        for (int r = e_voff[e]; r < e_voff[e + 1]; r++) {
          int w = e_vert[r];
          if (otn[w] < 0) {
            otn[w] = iloc;
            nto[iloc] = w;
            iloc++;
          }
        }
      }
    }
    if (iloc == before) {
      for (int i = 0; i < nv; i++)
        if (otn[i] < 0) {
          otn[i] = iloc;
          nto[iloc] = i;
          iloc++;
          break;
        }
    }
    istart = iend + 1;
    iend = iloc - 1;
  }
  m.vert_nto = nto;
  m.vcoord.resize(3 * (size_t)nv);
  for (int n = 0; n < nv; n++)
    for (int d = 0; d < 3; d++) m.vcoord[3 * (size_t)n + d] = coord[3 * (size_t)nto[n] + d];
  {  // vertex element lists move with their vertex
    std::vector<int> c2(nv);
    for (int n = 0; n < nv; n++) c2[n] = v_eoff[nto[n] + 1] - v_eoff[nto[n]];
    std::vector<int> off2;
    csr_from_counts(off2, c2);
    std::vector<int> el2(off2[nv]);
    for (int n = 0; n < nv; n++)
      std::copy(v_elem.begin() + v_eoff[nto[n]], v_elem.begin() + v_eoff[nto[n] + 1], el2.begin() + off2[n]);
    v_eoff.swap(off2);
    v_elem.swap(el2);
  }
  for (auto &v : e_vert) v = otn[v];
  // ---- renumber_elem_mesh (:752-849): first appearance over vertex lists
  std::vector<int> eotn(ne, -1), ento(ne);
  {
    int il = 0;
    for (int v = 0; v < nv; v++)
      for (int q = v_eoff[v]; q < v_eoff[v + 1]; q++) {
        int le = v_elem[q];
        if (eotn[le] < 0) eotn[le] = il++;
      }
    if (il != ne) return "mesh has elements with no vertex";
    for (int e = 0; e < ne; e++) ento[eotn[e]] = e;
  }
  m.elem_nto = ento;
  // minimal faces in file order (element order of the file, local face order)
  std::vector<int> mf_off(ne + 1, 0);
  for (int e = 0; e < ne; e++) mf_off[e + 1] = mf_off[e] + gmsh_n_face(e_kind[e]);
  const int nmf = mf_off[ne];
  std::vector<int> mf_nv(nmf), mf_left(nmf), mf_leftf(nmf), mf_v(4 * (size_t)nmf);
  for (int e = 0; e < ne; e++) {
    int nf = gmsh_n_face(e_kind[e]);
    for (int j = 0; j < nf; j++) {
      const int8_t *lf = local_face(e_kind[e], j);
      int f = mf_off[e] + j;
      int n = (lf[3] < 0) ? 3 : 4;
      mf_nv[f] = n;
      mf_left[f] = eotn[e];
      mf_leftf[f] = j;
      for (int q = 0; q < 4; q++) mf_v[4 * (size_t)f + q] = (q < n) ? e_vert[e_voff[e] + lf[q]] : -1;
    }
  }
  // permute element storage into the new numbering
  m.e_kind.resize(ne);
  m.e_voff.assign(ne + 1, 0);
  for (int n = 0; n < ne; n++) {
    m.e_kind[n] = e_kind[ento[n]];
    m.e_voff[n + 1] = m.e_voff[n] + gmsh_n_vert(m.e_kind[n]);
  }
  m.e_vert.resize(m.e_voff[ne]);
  for (int n = 0; n < ne; n++)
    std::copy(e_vert.begin() + e_voff[ento[n]], e_vert.begin() + e_voff[ento[n] + 1], m.e_vert.begin() + m.e_voff[n]);
  for (auto &e : v_elem) e = eotn[e];
  m.v_eoff = v_eoff;
  m.v_elem = v_elem;
  // ---- build_faces (:61-211): buckets keyed by the smallest vertex id,
  //      insertion order inside a bucket
  std::vector<Key4> mkey(nmf);
  int id_glob_max = -1;
  std::vector<int> ncol(nv, 0);
  for (int f = 0; f < nmf; f++) {
    mkey[f] = make_key(&mf_v[4 * (size_t)f], mf_nv[f]);
    ncol[mkey[f].v[0]]++;
    id_glob_max = std::max(id_glob_max, mkey[f].v[0]);
  }
  std::vector<int> bstart;
  csr_from_counts(bstart, ncol);
  std::vector<int> bfill(nv, 0);
  std::vector<int> s_src(nmf, -1), s_right(nmf, 0), s_rightf(nmf, 0);
  int n_faces = 0;
  for (int f = 0; f < nmf; f++) {
    int b = mkey[f].v[0];
    int found = -1;
    for (int s = bstart[b]; s < bstart[b] + bfill[b]; s++)
      if (same_key(mkey[s_src[s]], mkey[f])) {
        found = s;
        break;
      }
    if (found >= 0) {
      s_right[found] = mf_left[f] + 1;
      s_rightf[found] = mf_leftf[f];
    } else {
      int s = bstart[b] + bfill[b]++;
      s_src[s] = f;
      n_faces++;
    }
  }
  {
    size_t pos = 0;
    for (int j = 0; j < nbf; j++) {
      int n = bf_nvert[j];
      if (n < 3 || n > 4) return "boundary face with unsupported vertex count";
      if (bf_tag[j] < 0 || bf_tag[j] > 10) return "boundary tag outside 0..10 (mnbc)";
      int vv[4];
      for (int q = 0; q < n; q++) {
        int fv = bf_vert[pos + q] - 1;
        if (fv < 0 || fv >= nv) return "boundary face vertex id out of range";
        vv[q] = otn[fv];
      }
      pos += n;
      Key4 k = make_key(vv, n);
      int b = k.v[0];
      if (b > id_glob_max) continue;
      for (int s = bstart[b]; s < bstart[b] + bfill[b]; s++)
        if (same_key(mkey[s_src[s]], k)) {
          if (s_right[s] <= 0) {
            s_right[s] = -bf_tag[j];
            s_rightf[s] = 0;
          }
          break;
        }
    }
  }
  m.nf = n_faces;
  m.f_voff.assign(n_faces + 1, 0);
  m.f_left.resize(n_faces);
  m.f_leftf.resize(n_faces);
  m.f_right.resize(n_faces);
  m.f_rightf.resize(n_faces);
  m.e_foff.assign(ne + 1, 0);
  for (int n = 0; n < ne; n++) m.e_foff[n + 1] = m.e_foff[n] + gmsh_n_face(m.e_kind[n]);
  m.e_face.assign(m.e_foff[ne], -1);
  {
    int fid = 0;
    std::vector<int> fv;
    fv.reserve(4 * (size_t)n_faces);
    for (int b = 0; b < nv; b++)
      for (int s = bstart[b]; s < bstart[b] + bfill[b]; s++) {
        int src = s_src[s];
        for (int q = 0; q < mf_nv[src]; q++) fv.push_back(mf_v[4 * (size_t)src + q]);
        m.f_voff[fid + 1] = m.f_voff[fid] + mf_nv[src];
        m.f_left[fid] = mf_left[src];
        m.f_leftf[fid] = mf_leftf[src];
        m.f_right[fid] = s_right[s];
        m.f_rightf[fid] = s_rightf[s];
        m.e_face[m.e_foff[mf_left[src]] + mf_leftf[src]] = fid;
        if (s_right[s] > 0) m.e_face[m.e_foff[s_right[s] - 1] + s_rightf[s]] = fid;
        fid++;
      }
    m.f_vert.swap(fv);
  }
  for (int q = 0; q < m.e_foff[ne]; q++)
    if (m.e_face[q] < 0) return "face shared by more than two elements";
  // ---- build_node_face_connectivity (:299-333)
  cnt.assign(nv, 0);
  for (int v : m.f_vert) cnt[v]++;
  csr_from_counts(m.v_foff, cnt);
  m.v_face.resize(m.v_foff[nv]);
  {
    std::vector<int> fill(nv, 0);
    for (int f = 0; f < m.nf; f++)
      for (int q = m.f_voff[f]; q < m.f_voff[f + 1]; q++) {
        int v = m.f_vert[q];
        m.v_face[m.v_foff[v] + fill[v]++] = f;
      }
  }
  // ---- renumber_faces_mesh (:928-974): first appearance over the vertex
  //      face lists (the bucket order already is that order; checked here)
  {
    std::vector<int> fotn(m.nf, -1);
    int il = 0;
    for (int v = 0; v < nv; v++)
      for (int q = m.v_foff[v]; q < m.v_foff[v + 1]; q++) {
        int f = m.v_face[q];
        if (fotn[f] < 0) fotn[f] = il++;
      }
    for (int f = 0; f < m.nf; f++)
      if (fotn[f] != f) return "internal error: face renumbering is not the identity";
  }
  // ---- build_sub_faces (:335-402): vertex-major sub-face ids
  m.nsf = (int)m.v_face.size();
  m.sf_off = m.v_foff;
  m.sf_vert.resize(m.nsf);
  m.sf_face.resize(m.nsf);
  m.sf_le.resize(m.nsf);
  m.sf_re.resize(m.nsf);
  m.f_sfoff = m.f_voff;
  m.f_sf.assign(m.f_voff[m.nf], -1);
  cnt.assign(ne, 0);
  for (int e = 0; e < ne; e++)
    for (int q = m.e_foff[e]; q < m.e_foff[e + 1]; q++) cnt[e] += m.f_voff[m.e_face[q] + 1] - m.f_voff[m.e_face[q]];
  csr_from_counts(m.e_sfoff, cnt);
  m.e_sf.assign(m.e_sfoff[ne], -1);
  {
    std::vector<int> ffill(m.nf, 0), efill(ne, 0);
    int id = 0;
    for (int v = 0; v < nv; v++)
      for (int q = m.v_foff[v]; q < m.v_foff[v + 1]; q++, id++) {
        int f = m.v_face[q];
        m.f_sf[m.f_sfoff[f] + ffill[f]++] = id;
        m.sf_face[id] = f;
        m.sf_vert[id] = v;
        m.sf_le[id] = m.f_left[f];
        m.sf_re[id] = m.f_right[f];
        int l = m.f_left[f];
        m.e_sf[m.e_sfoff[l] + efill[l]++] = id;
        int r = m.f_right[f];
        if (r > 0) m.e_sf[m.e_sfoff[r - 1] + efill[r - 1]++] = id;
      }
  }
  // ---- build_sub_elems (:404-452): cell-major sub-element ids
  m.nse = m.e_voff[ne];
  m.se_elem.resize(m.nse);
  m.se_vert.resize(m.nse);
  m.se_loc.resize(m.nse);
  m.se_sfoff.assign(m.nse + 1, 0);
  for (int e = 0; e < ne; e++)
    for (int j = m.e_voff[e]; j < m.e_voff[e + 1]; j++) {
      int n = 0;
      for (int q = m.e_sfoff[e]; q < m.e_sfoff[e + 1]; q++)
        if (m.sf_vert[m.e_sf[q]] == m.e_vert[j]) n++;
      m.se_sfoff[j + 1] = n;
      m.se_elem[j] = e;
      m.se_vert[j] = m.e_vert[j];
    }
  for (int j = 0; j < m.nse; j++) m.se_sfoff[j + 1] += m.se_sfoff[j];
  m.se_sf.resize(m.se_sfoff[m.nse]);
  for (int e = 0; e < ne; e++)
    for (int j = m.e_voff[e]; j < m.e_voff[e + 1]; j++) {
      int loc = m.se_sfoff[j];
      for (int q = m.e_sfoff[e]; q < m.e_sfoff[e + 1]; q++)
        if (m.sf_vert[m.e_sf[q]] == m.e_vert[j]) m.se_sf[loc++] = m.e_sf[q];
    }
  // ---- find_sub_elems_around_sub_face (:454-473)
  m.sf_lse.assign(m.nsf, -1);
  m.sf_rse.assign(m.nsf, -1);
  for (int j = 0; j < m.nse; j++) {
    int e = m.se_elem[j];
    for (int q = m.se_sfoff[j]; q < m.se_sfoff[j + 1]; q++) {
      int sf = m.se_sf[q];
      if (m.sf_le[sf] == e) m.sf_lse[sf] = j;
      if (m.sf_re[sf] == e + 1) m.sf_rse[sf] = j;
    }
  }
  // ---- build_node_sub_face_connectivity (:511-556)
  cnt.assign(nv, 0);
  for (int j = 0; j < m.nse; j++) cnt[m.se_vert[j]]++;
  csr_from_counts(m.v_seoff, cnt);
  m.v_se.resize(m.nse);
  {
    std::vector<int> fill(nv, 0);
    for (int j = 0; j < m.nse; j++) {
      int v = m.se_vert[j];
      m.se_loc[j] = fill[v];
      m.v_se[m.v_seoff[v] + fill[v]++] = j;
    }
  }
  // ---- compute_elem_tag (:558-579)
  m.e_tag.assign(ne, -1);
  for (int f = 0; f < m.nf; f++)
    if (m.f_right[f] < 0) {
      int tag = -m.f_right[f];
      int e = m.f_left[f];
      if (m.e_tag[e] < 0 || m.e_tag[e] > tag) m.e_tag[e] = tag;
    }
  // ---- compute_neigh_by_vert (:595-654): the cells around the cell's
  // vertices, vertex by vertex in the cell's order, first occurrence kept
  m.e_nbvoff.assign(ne + 1, 0);
  m.e_nbv.clear();
  {
    std::vector<int> seen(ne, -1);
    for (int e = 0; e < ne; e++) {
      for (int j = m.e_voff[e]; j < m.e_voff[e + 1]; j++) {
        const int v = m.e_vert[j];
        for (int k = m.v_eoff[v]; k < m.v_eoff[v + 1]; k++) {
          const int o = m.v_elem[k];
          if (o != e && seen[o] != e) {
            seen[o] = e;
            m.e_nbv.push_back(o);
          }
        }
      }
      m.e_nbvoff[e + 1] = (int)m.e_nbv.size();
    }
  }
  m.nmax_sfn = 0;
  m.nmax_sen = 0;
  for (int v = 0; v < nv; v++) {
    m.nmax_sfn = std::max(m.nmax_sfn, m.sf_off[v + 1] - m.sf_off[v]);
    m.nmax_sen = std::max(m.nmax_sen, m.v_seoff[v + 1] - m.v_seoff[v]);
  }
  return "";
}

// ---- periodic_mesh: stitch_periodic_faces (:976-1073) and
//      build_periodic_stencil (:1075-1328)
namespace {

constexpr double kPerTol = 1e-12;
inline bool on_plane(double x, double c) { return x < c + kPerTol && x > c - kPerTol; }

// Vertex lists that grow as partners are processed; a vertex not yet touched
// reads its built CSR range.
struct PeriodicLists {
  const HostMesh &m;
  std::vector<std::vector<int>> sf, se;
  std::vector<char> own;
  explicit PeriodicLists(const HostMesh &mesh) : m(mesh), sf(mesh.nv), se(mesh.nv), own(mesh.nv, 0) {}
  void touch(int v) {
    if (own[v]) return;
    own[v] = 1;
    sf[v].clear();
    for (int s = m.sf_off[v]; s < m.sf_off[v + 1]; s++) sf[v].push_back(s);
    se[v].assign(m.v_se.begin() + m.v_seoff[v], m.v_se.begin() + m.v_seoff[v + 1]);
  }
};

}  // namespace

std::string make_periodic(HostMesh &m) {
  if (m.periodic) return "the mesh is periodic already";
  // face centres as vertex averages, as build_mesh takes them (:44-52)
  std::vector<double> fc(3L * m.nf);
  for (int f = 0; f < m.nf; f++) {
    double c[3] = {0.0, 0.0, 0.0};
    for (int q = m.f_voff[f]; q < m.f_voff[f + 1]; q++)
      for (int d = 0; d < 3; d++) c[d] = c[d] + m.vcoord[3L * m.f_vert[q] + d];
    const double n = (double)(m.f_voff[f + 1] - m.f_voff[f]);
    for (int d = 0; d < 3; d++) fc[3L * f + d] = c[d] / n;
  }
  // the two other axes of direction d, in the reference's order
  auto axes = [](int d, int &a, int &b) {
    a = d == 0 ? 1 : 0;
    b = d == 2 ? 1 : 2;
  };
  std::vector<int> lo[3], hi[3];
  for (int f = 0; f < m.nf; f++)
    for (int d = 0; d < 3; d++) {
      if (on_plane(fc[3L * f + d], -0.5)) lo[d].push_back(f);
      if (on_plane(fc[3L * f + d], 0.5)) hi[d].push_back(f);
    }
  for (int d = 0; d < 3; d++)
    if (lo[d].size() != hi[d].size()) return "Error periodic boundaries ! (unmatched face counts)";
  for (int d = 0; d < 3; d++) {
    int a, b;
    axes(d, a, b);
    for (int ii : lo[d])
      for (int jj : hi[d]) {
        const double da = fc[3L * ii + a] - fc[3L * jj + a], db = fc[3L * ii + b] - fc[3L * jj + b];
        if (std::sqrt(da * da + db * db) < kPerTol) {
          m.f_right[ii] = m.f_left[jj] + 1;
          m.f_right[jj] = m.f_left[ii] + 1;
        }
      }
  }
  for (int s = 0; s < m.nsf; s++) {
    m.sf_le[s] = m.f_left[m.sf_face[s]];
    m.sf_re[s] = m.f_right[m.sf_face[s]];
  }
  PeriodicLists L(m);
  // sub-faces joining the same two cells, in either orientation (left cells
  // compared one-based, as the reference stores them)
  auto same_cells = [&](int p, int q) {
    const int lp = m.sf_le[p] + 1, rp = m.sf_re[p], lq = m.sf_le[q] + 1, rq = m.sf_re[q];
    return (lp == rq && rp == lq) || (lp == lq && rp == rq);
  };
  std::string err;
  auto process_pair = [&](int v1, int v2) {
    L.touch(v1);
    L.touch(v2);
    std::vector<int> &s1 = L.sf[v1], &s2 = L.sf[v2];
    // step 1: right sub-element of v1's stitched sub-faces, from v2's list
    for (int sfki : s1) {
      const int re = m.sf_re[sfki];
      if (re > 0 && m.sf_rse[sfki] < 0)
        for (int id_se : L.se[v2])
          if (m.se_elem[id_se] == re - 1) {
            m.sf_rse[sfki] = id_se;
            break;
          }
    }
    // step 2: v2's sub-faces appended, without repeats
    const int n1 = (int)s1.size(), n2 = (int)s2.size();
    int nm = 0;
    for (int p : s1)
      for (int q : s2)
        if (same_cells(p, q)) nm++;
    const int st = n1 + n2 - nm;
    if (st < n1) {
      err = "periodic stencil of vertex " + std::to_string(v1 + 1) + " cannot be formed";
      return;
    }
    std::vector<int> nl(s1);
    nl.resize(st, -1);
    for (int ki = n1; ki < st; ki++)
      for (int q : s2) {
        bool found = false;
        for (int p = 0; p < ki; p++)
          if (nl[p] >= 0 && same_cells(nl[p], q)) found = true;
        if (!found) {
          nl[ki] = q;
          break;
        }
      }
    for (int x : nl)
      if (x < 0) {
        err = "periodic stencil of vertex " + std::to_string(v1 + 1) + " cannot be filled";
        return;
      }
    s1.swap(nl);
    // step 3: v2's own sub-elements appended
    std::vector<int> add;
    for (int id_se : L.se[v2])
      if (m.se_vert[id_se] == v2) add.push_back(id_se);
    L.se[v1].insert(L.se[v1].end(), add.begin(), add.end());
  };
  auto pair = [&](int a, int b) {
    if (err.empty()) process_pair(a, b);
    if (err.empty()) process_pair(b, a);
  };
  std::vector<int> vlo[3], vhi[3];
  for (int v = 0; v < m.nv; v++)
    for (int d = 0; d < 3; d++) {
      if (on_plane(m.vcoord[3L * v + d], -0.5)) vlo[d].push_back(v);
      if (on_plane(m.vcoord[3L * v + d], 0.5)) vhi[d].push_back(v);
    }
  const double *X = m.vcoord.data();
  for (int d = 0; d < 3; d++) {
    int a, b;
    axes(d, a, b);
    for (int ii : vlo[d])
      for (int jj : vhi[d]) {
        const double da = X[3L * ii + a] - X[3L * jj + a], db = X[3L * ii + b] - X[3L * jj + b];
        if (std::sqrt(da * da + db * db) < kPerTol) pair(ii, jj);
      }
  }
  // edges: lower on d1 and d2 <-> upper on both, matched along the third axis
  const int edge[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
  for (const auto &e : edge)
    for (int ii : vlo[e[0]]) {
      if (!on_plane(X[3L * ii + e[1]], -0.5)) continue;
      for (int jj : vhi[e[0]])
        if (on_plane(X[3L * jj + e[1]], 0.5) && std::fabs(X[3L * ii + e[2]] - X[3L * jj + e[2]]) < kPerTol)
          pair(ii, jj);
    }
  for (int ii : vlo[0]) {
    if (!(on_plane(X[3L * ii + 1], -0.5) && on_plane(X[3L * ii + 2], -0.5))) continue;
    for (int jj : vhi[0])
      if (on_plane(X[3L * jj + 1], 0.5) && on_plane(X[3L * jj + 2], 0.5)) pair(ii, jj);
  }
  if (!err.empty()) return err;
  // flatten: the sweep's lists and the partners' sub-elements
  m.x_sfoff.assign(m.nv + 1, 0);
  m.x_seoff.assign(m.nv + 1, 0);
  m.x_sf.clear();
  m.x_se.clear();
  for (int v = 0; v < m.nv; v++) {
    if (L.own[v]) {
      m.x_sf.insert(m.x_sf.end(), L.sf[v].begin(), L.sf[v].end());
      const int n0 = m.v_seoff[v + 1] - m.v_seoff[v];
      m.x_se.insert(m.x_se.end(), L.se[v].begin() + n0, L.se[v].end());
    } else {
      for (int s = m.sf_off[v]; s < m.sf_off[v + 1]; s++) m.x_sf.push_back(s);
    }
    m.x_sfoff[v + 1] = (int)m.x_sf.size();
    m.x_seoff[v + 1] = (int)m.x_se.size();
    m.nmax_sfn = std::max(m.nmax_sfn, m.x_sfoff[v + 1] - m.x_sfoff[v]);
    m.nmax_sen = std::max(m.nmax_sen, m.v_seoff[v + 1] - m.v_seoff[v] + m.x_seoff[v + 1] - m.x_seoff[v]);
  }
  m.periodic = true;
  return "";
}

}  // namespace subfv
