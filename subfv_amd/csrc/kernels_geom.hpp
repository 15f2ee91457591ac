// Geometry of the subface mesh on the device, one thread per entity, in the
// operation order of compute_geometry_mesh
// (src/mesh/mesh_geometry_module.F90:16-57, 59-288, 290-411, 714-738).
#pragma once
#include "device_mesh.hpp"
#include "physics.hpp"

namespace sfv {

SFV_DEV void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = -(a[0] * b[2] - a[2] * b[0]);
  c[2] = a[0] * b[1] - a[1] * b[0];
}
SFV_DEV double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
SFV_DEV double nrm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// compute_face_centroid (:210-235)
__global__ void k_face_centroid(DevMesh m) {
  long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  int b = m.f_voff[f], n = m.f_voff[f + 1] - b;
  double cg[3] = {0, 0, 0};
  for (int j = 0; j < n; j++)
    for (int d = 0; d < 3; d++) cg[d] = cg[d] + m.vcoord[3L * m.f_vert[b + j] + d];
  for (int d = 0; d < 3; d++) cg[d] = cg[d] / (double)n;
  double c[3] = {0, 0, 0}, sa = 0.0;
  for (int j = 0; j < n; j++) {
    const double *cv = &m.vcoord[3L * m.f_vert[b + j]];
    const double *cvp = &m.vcoord[3L * m.f_vert[b + (j + 1) % n]];
    double a[3], bb[3], x[3];
    for (int d = 0; d < 3; d++) {
      a[d] = cv[d] - cg[d];
      bb[d] = cvp[d] - cg[d];
    }
    cross3(a, bb, x);
    double area = 0.5 * nrm3(x);
    for (int d = 0; d < 3; d++) c[d] = c[d] + area * (cv[d] + cvp[d] + cg[d]) / 3.0;
    sa = sa + area;
  }
  for (int d = 0; d < 3; d++) m.fcoord[3 * f + d] = c[d] / sa;
}

// compute_elem_centroid (:75-127)
__global__ void k_elem_centroid(DevMesh m) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.C) return;
  int vb = m.e_voff[e], nv = m.e_voff[e + 1] - vb;
  double ce[3] = {0, 0, 0};
  for (int j = 0; j < nv; j++)
    for (int d = 0; d < 3; d++) ce[d] = ce[d] + m.vcoord[3L * m.e_vert[vb + j] + d];
  for (int d = 0; d < 3; d++) ce[d] = ce[d] / (double)nv;
  double c[3] = {0, 0, 0}, sv = 0.0;
  for (int q = m.e_foff[e]; q < m.e_foff[e + 1]; q++) {
    int f = m.e_face[q];
    const double *cf = &m.fcoord[3L * f];
    int fb = m.f_voff[f], n = m.f_voff[f + 1] - fb;
    bool left = (m.f_left[f] == e);
    for (int k = 0; k < n; k++) {
      const double *cp = &m.vcoord[3L * m.f_vert[fb + k]];
      const double *cpp = &m.vcoord[3L * m.f_vert[fb + (k + 1) % n]];
      double a[3], bb[3], x[3], fe[3];
      for (int d = 0; d < 3; d++) {
        a[d] = cp[d] - cf[d];
        bb[d] = cpp[d] - cf[d];
        fe[d] = cf[d] - ce[d];
      }
      if (left)
        cross3(a, bb, x);
      else
        cross3(bb, a, x);
      double vol = dot3(x, fe) / 6.0;
      for (int d = 0; d < 3; d++) c[d] = c[d] + vol * (cp[d] + cpp[d] + cf[d] + ce[d]) / 4.0;
      sv = sv + vol;
    }
  }
  for (int d = 0; d < 3; d++) m.ccoord[3 * e + d] = c[d] / sv;
}

// position of vertex v in face f and the midpoints towards its two face
// neighbours (shared by compute_subface_norm and compute_subelem_volume)
SFV_DEV void subface_pts(const DevMesh &m, int f, int v, double *cp, double *cpm, double *cpp) {
  int b = m.f_voff[f], n = m.f_voff[f + 1] - b, jj = 0;
  for (int j = 0; j < n; j++)
    if (m.f_vert[b + j] == v) {
      jj = j;
      break;
    }
  const double *c = &m.vcoord[3L * v];
  const double *cm = &m.vcoord[3L * m.f_vert[b + (jj - 1 + n) % n]];
  const double *cq = &m.vcoord[3L * m.f_vert[b + (jj + 1) % n]];
  for (int d = 0; d < 3; d++) {
    cp[d] = c[d];
    cpm[d] = 0.5 * (cm[d] + c[d]);
    cpp[d] = 0.5 * (cq[d] + c[d]);
  }
}

// compute_subface_norm (:237-271), unnormalised
SFV_DEV void subface_raw_norm(const DevMesh &m, int sf, double *nr) {
  int v = m.sf_vert[sf], f = m.sf_face[sf];
  double cp[3], cpm[3], cpp[3];
  subface_pts(m, f, v, cp, cpm, cpp);
  const double *cf = &m.fcoord[3L * f];
  double a[3], b[3], c[3], x1[3], x2[3];
  for (int d = 0; d < 3; d++) {
    a[d] = cpm[d] - cf[d];
    b[d] = cp[d] - cf[d];
    c[d] = cpp[d] - cf[d];
  }
  cross3(a, b, x1);
  cross3(b, c, x2);
  for (int d = 0; d < 3; d++) nr[d] = 0.5 * (x1[d] + x2[d]);
}

__global__ void k_subface_norm(DevMesh m) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m.SF) return;
  double n[3];
  subface_raw_norm(m, (int)s, n);
  double a = nrm3(n);
  m.sfn[s] = n[0] / a;
  m.sfn[m.SF + s] = n[1] / a;
  m.sfn[2 * m.SF + s] = n[2] / a;
  m.sfn[3 * m.SF + s] = a;
}

// compute_face_norm (:177-191) then normalisation (:39-43)
__global__ void k_face_norm(DevMesh m) {
  long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.F) return;
  double acc[3] = {0, 0, 0};
  for (int q = m.f_sfoff[f]; q < m.f_sfoff[f + 1]; q++) {
    double n[3];
    subface_raw_norm(m, m.f_sf[q], n);
    for (int d = 0; d < 3; d++) acc[d] = acc[d] + n[d];
  }
  double a = nrm3(acc);
  m.farea[f] = a;
  for (int d = 0; d < 3; d++) m.fnorm[3 * f + d] = acc[d] / a;
}

// compute_subelem_volume (:129-175)
__global__ void k_subelem_volume(DevMesh m) {
  long se = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (se >= m.SE) return;
  int e = m.se_elem[se];
  const double *ce = &m.ccoord[3L * e];
  double vol = 0.0;
  for (int q = m.se_sfoff[se]; q < m.se_sfoff[se + 1]; q++) {
    int sf = m.se_sf[q];
    int v = m.sf_vert[sf], f = m.sf_face[sf];
    double cp[3], cpm[3], cpp[3];
    subface_pts(m, f, v, cp, cpm, cpp);
    const double *cf = &m.fcoord[3L * f];
    double a[3], b[3], c[3], fe[3], x1[3], x2[3];
    for (int d = 0; d < 3; d++) {
      a[d] = cpm[d] - cf[d];
      b[d] = cp[d] - cf[d];
      c[d] = cpp[d] - cf[d];
      fe[d] = cf[d] - ce[d];
    }
    if (m.f_left[f] == e) {
      cross3(a, b, x1);
      cross3(b, c, x2);
    } else {
      cross3(c, b, x1);
      cross3(b, a, x2);
    }
    vol = vol + dot3(x1, fe) / 6.0 + dot3(x2, fe) / 6.0;
  }
  m.se_vol[se] = vol;
}

// compute_elem_volume (:59-73) and check_mesh (:290-411)
__global__ void k_elem_volume_check(DevMesh m, int *bad) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.C) return;
  double v = 0.0;
  for (int j = m.e_voff[e]; j < m.e_voff[e + 1]; j++) v = v + m.se_vol[j];
  m.cvol[e] = v;
  bool ok = true;
  double vol = 0.0, nt[3] = {0, 0, 0}, ns[3] = {0, 0, 0};
  for (int q = m.e_foff[e]; q < m.e_foff[e + 1]; q++) {
    int f = m.e_face[q];
    double sg = (m.f_left[f] == e) ? 1.0 : -1.0;
    double an[3];
    for (int d = 0; d < 3; d++) an[d] = m.farea[f] * (sg > 0 ? m.fnorm[3L * f + d] : -m.fnorm[3L * f + d]);
    vol = vol + dot3(&m.fcoord[3L * f], an);
  }
  vol = vol / 3.0;
  ok = ok && (fabs(v - vol) < 1e-14);
  for (int q = m.e_foff[e]; q < m.e_foff[e + 1]; q++) {
    int f = m.e_face[q];
    bool lf = (m.f_left[f] == e);
    for (int d = 0; d < 3; d++) nt[d] = nt[d] + m.farea[f] * (lf ? m.fnorm[3L * f + d] : -m.fnorm[3L * f + d]);
  }
  ok = ok && (nrm3(nt) < 1e-14);
  for (int q = m.e_foff[e]; q < m.e_foff[e + 1]; q++) {
    int f = m.e_face[q];
    bool lf = (m.f_left[f] == e);
    for (int k = m.f_sfoff[f]; k < m.f_sfoff[f + 1]; k++) {
      int sf = m.f_sf[k];
      double a = m.sfn[3 * m.SF + sf];
      for (int d = 0; d < 3; d++) {
        double nr = m.sfn[d * m.SF + sf];
        ns[d] = ns[d] + a * (lf ? nr : -nr);
      }
    }
  }
  ok = ok && (nrm3(ns) < 1e-14);
  for (int j = m.e_voff[e]; j < m.e_voff[e + 1]; j++) ok = ok && (m.se_vol[j] > 0.0);
  if (!ok) atomicAdd(bad, 1);
}

// compute_vert_volume (:273-288) and find_if_vert_is_bound (:714-738)
__global__ void k_vert_volume_bound(DevMesh m, int b2d) {
  long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= m.V) return;
  double vol = 0.0;
  for (int q = m.se_off[v]; q < m.se_off[v + 1]; q++) vol = vol + m.se_vol[m.v_seid[q]];
  if (m.vx_seoff)  // periodic: the partners' sub-elements follow (build_periodic_stencil step 3)
    for (int q = m.vx_seoff[v]; q < m.vx_seoff[v + 1]; q++) vol = vol + m.se_vol[m.vx_se[q]];
  m.vvol[v] = vol;
  uint8_t bnd = 0;
  // faces around v are exactly the faces of its own sub-faces
  const int s_end = m.sf_n0 ? m.sf_off[v] + m.sf_n0[v] : m.sf_off[v + 1];
  for (int s = m.sf_off[v]; s < s_end; s++) {
    int f = m.sf_face[s];
    int re = m.f_right[f];
    if (b2d) {
      if (re <= 0 && fabs(m.fnorm[3L * f + 2]) < 1e-12) bnd = 1;
    } else {
      if (re <= 0) bnd = 1;
    }
  }
  m.vbound[v] = bnd;
}

// wall_normal (ns_euler_primitives_module.F90:162-190), per vertex
__global__ void k_wall_normal(DevMesh m, BCParams bc) {
  long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= m.V) return;
  double B[3] = {0, 0, 0};
  for (int s = m.sf_off[v]; s < m.sf_off[v + 1]; s++) {
    int re = m.sf_re[s];
    if (re < 0) {
      double nz = m.sfn[2 * m.SF + s];
      if (!(bc.b2d && fabs(nz) > 1e-12)) {
        if (re == -1 || bc.bc_id[-re - 1] == kBCWall) {
          double a = m.sfn[3 * m.SF + s];
          for (int d = 0; d < 3; d++) B[d] = B[d] + a * m.sfn[d * m.SF + s];
        }
      }
    }
  }
  for (int d = 0; d < 3; d++) m.Bp[d * m.V + v] = B[d];
}

}  // namespace sfv
