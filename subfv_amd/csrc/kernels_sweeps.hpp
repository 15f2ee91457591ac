// The node sweeps of src/lagrange and src/acoustic on the vertex-group
// structure of the Euler sweep (one 256-thread workgroup per group of
// consecutive vertices, thread = sub-face, per-vertex sums in sub-face order
// from LDS, the 3x3 solves in the first wavefront).
//
//   compute_rhs_lagrange   src/lagrange/lagrange_module.F90:14-147
//   acoustic time step     src/acoustic/acoustic_main.F90:171-290,
//                          multi_point_pressure acoustic_module.F90:11-35,
//                          rotate :232-262
//
// Both reference loops scatter into the cells with one running sum per cell
// (vertex order, then sub-face order).  Each term is therefore stored on its
// own: T[q][e][k] is the term of the e-th contributing sub-face of
// node-sorted slot q, and the gather adds a cell's terms in slot order
// (increasing vertex id), then contributor order, from zero.
#pragma once
#include "kernels_cell.hpp"
#include "kernels_flux.hpp"

namespace sfv {

constexpr int kSweepTerm = 5;  // doubles per stored term (acoustic uses 4)

struct SweepLds {
  double red[12][kRedStride];
  double vs[kGrpVert][12];
  double vn[kGrpVert][3];
  int vtab[kGrpVert + 1];
  int lc[kGrpThreads];
};

struct LagrangeArgs {
  DevMesh m;
  const double *sol;   // [C][5] (rho, u, v, w, E)
  const double *mass;  // [C]
  double *T;           // [SE][kSlotSrc][5] signed terms
  double *vp;          // [V][3]
  int b2d;
  int per_xcd, nblocks;
};

constexpr double kGammaL = 7.0 / 5.0;

// pressure (lagrange_module.F90:429-436): norm2(u(2:4)) squared
SFV_DEV double pressure_l(const double *u) {
  const double nrm = sqrt(u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
  return (kGammaL - 1.0) / u[0] * (u[4] - 0.5 * (nrm * nrm));
}

SFV_DEV double det2(double a, double b, double c, double d) { return a * d - c * b; }

// inverse_3_by_3 (linear_solver_module.F90:134-172), column-major in and out
SFV_DEV void inverse_3_by_3(const double *M, double *inv) {
#define MM(i, j) M[((i)-1) + 3 * ((j)-1)]
#define TT(i, j) MM(j, i)
  double adj[9];  // adj[(i-1) + 3 (j-1)]
  adj[0] = det2(TT(2, 2), TT(2, 3), TT(3, 2), TT(3, 3));
  adj[3] = -det2(TT(2, 1), TT(2, 3), TT(3, 1), TT(3, 3));
  adj[6] = det2(TT(2, 1), TT(2, 2), TT(3, 1), TT(3, 2));
  adj[1] = -det2(TT(1, 2), TT(1, 3), TT(3, 2), TT(3, 3));
  adj[4] = det2(TT(1, 1), TT(1, 3), TT(3, 1), TT(3, 3));
  adj[7] = -det2(TT(1, 1), TT(1, 2), TT(3, 1), TT(3, 2));
  adj[2] = det2(TT(1, 2), TT(1, 3), TT(2, 2), TT(2, 3));
  adj[5] = -det2(TT(1, 1), TT(1, 3), TT(2, 1), TT(2, 3));
  adj[8] = det2(TT(1, 1), TT(1, 2), TT(2, 1), TT(2, 2));
  const double det = MM(1, 1) * adj[0] + MM(1, 2) * adj[1] + MM(1, 3) * adj[2];
  const double r = 1.0 / det;
#pragma unroll
  for (int q = 0; q < 9; q++) inv[q] = r * adj[q];
#undef TT
#undef MM
}

SFV_DEV void matvec3(const double *A, const double *x, double *y) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) acc = acc + A[i + 3 * j] * x[j];
    y[i] = acc;
  }
}

// store each slot's contributor terms: thread q owns slot q of the group;
// rows 0..NV-1 of red hold the left terms, NV..2NV-1 the right terms
template <int NV>
SFV_DEV void store_slot_terms(const DevMesh &m, SweepLds &L, int se0, int nse, double *T) {
  const int tid = threadIdx.x;
  if (tid >= nse) return;
  const uint32_t src = m.se_src[se0 + tid];
  const int cell = m.se_cell[se0 + tid];
  double *o = T + (long)kSweepTerm * kSlotSrc * (se0 + tid);
#pragma unroll
  for (int e = 0; e < kSlotSrc; e++) {
    const int j = (src >> (8 * e)) & 0xFFu;
    if (j == 0xFF) break;
    const int r0 = (L.lc[j] == cell) ? 0 : NV;
#pragma unroll
    for (int k = 0; k < NV; k++) o[kSweepTerm * e + k] = L.red[r0 + k][j];
  }
}

__global__ __launch_bounds__(kGrpThreads) void k_lagrange_group(LagrangeArgs a) {
  __shared__ SweepLds L;
  const DevMesh &m = a.m;
  const int lb = xcd_block(a.per_xcd);
  if (lb >= a.nblocks) return;
  const int tid = threadIdx.x;
  const int *gh = m.grp + kGrpHdr * (long)lb;
  const int vb = gh[0], nv = gh[1], sf0 = gh[2], nsf = gh[3], se0 = gh[4], nse = gh[5];
  if (tid <= nv) L.vtab[tid] = m.sf_off[vb + tid] - sf0;
  const bool active = tid < nsf;
  const long s = sf0 + (active ? tid : 0);
  const int f = m.sf_face[s];
  const double n[3] = {m.fnorm[3L * f], m.fnorm[3L * f + 1], m.fnorm[3L * f + 2]};
  const double area = m.sfn[3 * m.SF + s];
  const int idl = m.sf_le[s], idr = m.sf_re[s];
  __syncthreads();
  const int vl = seg_of(L.vtab, nv, tid);
  double ul[5], ur[5] = {1.0, 0.0, 0.0, 0.0, 1.0};
  for (int k = 0; k < 5; k++) ul[k] = a.sol[5L * idl + k];
  const double pl = pressure_l(ul), vl_n = ul[1] * n[0] + ul[2] * n[1] + ul[3] * n[2];
  double pr = 0.0, vr_n = 0.0, l1, l2 = 0.0;
  double cm = 0.0, cr = 0.0;  // Mp += cm n n^T, Rp += cr n
  bool addB = false;
  if (idr >= 0) {
    for (int k = 0; k < 5; k++) ur[k] = a.sol[5L * idr + k];
    pr = pressure_l(ur);
    vr_n = ur[1] * n[0] + ur[2] * n[1] + ur[3] * n[2];
    l1 = dmax(dmax(sqrt(kGammaL * pl * ul[0]) / ul[0], sqrt(dmax(0.0, pr - pl) / ul[0])), -(vr_n - vl_n) / ul[0]);
    l2 = dmax(dmax(sqrt(kGammaL * pr * ur[0]) / ur[0], sqrt(dmax(0.0, pl - pr) / ur[0])), -(vr_n - vl_n) / ur[0]);
    const double vbar = (l1 * vl_n + l2 * vr_n - (pr - pl)) / (l2 + l1);
    cm = area * (l1 + l2);
    cr = cm;  // Rp += ((area (l1 + l2)) n) vbar
    for (int d = 0; d < 3; d++) L.red[6 + d][tid] = active ? cr * n[d] * vbar : 0.0;
  } else if (a.b2d && fabs(n[2]) > 1e-8) {
    l1 = sqrt(kGammaL * pl * ul[0]) / ul[0];
    cm = area * 2 * l1;
    for (int d = 0; d < 3; d++) L.red[6 + d][tid] = 0.0;
  } else {
    l1 = sqrt(kGammaL * pl * ul[0]) / ul[0];
    addB = true;
    cm = area * l1;
    const double sr = area * (pl + l1 * vl_n);
    for (int d = 0; d < 3; d++) L.red[6 + d][tid] = active ? sr * n[d] : 0.0;
  }
  L.red[0][tid] = active ? cm * (n[0] * n[0]) : 0.0;
  L.red[1][tid] = active ? cm * (n[0] * n[1]) : 0.0;
  L.red[2][tid] = active ? cm * (n[0] * n[2]) : 0.0;
  L.red[3][tid] = active ? cm * (n[1] * n[1]) : 0.0;
  L.red[4][tid] = active ? cm * (n[1] * n[2]) : 0.0;
  L.red[5][tid] = active ? cm * (n[2] * n[2]) : 0.0;
  for (int d = 0; d < 3; d++) L.red[9 + d][tid] = (active && addB) ? area * n[d] : 0.0;
  __syncthreads();
  for (int p = tid; p < nv * 12; p += kGrpThreads) {
    const int vv = p / 12, c = p - vv * 12;
    L.vs[vv][c] = seg_sum(L.red[c], L.vtab[vv], L.vtab[vv + 1]);
  }
  __syncthreads();
  if (tid < nv) {  // first wavefront only
    const double *S = L.vs[tid];
    const double Mp[9] = {S[0], S[1], S[2], S[1], S[3], S[4], S[2], S[4], S[5]};
    const double Rp[3] = {S[6], S[7], S[8]}, Bp[3] = {S[9], S[10], S[11]};
    double Mi[9], v3[3];
    inverse_3_by_3(Mp, Mi);
    if (dmax(dmax(fabs(Bp[0]), fabs(Bp[1])), fabs(Bp[2])) > 1e-8) {
      double mb[3], w[3];
      matvec3(Mi, Bp, mb);
      const double Pp = (Rp[0] * mb[0] + Rp[1] * mb[1] + Rp[2] * mb[2]) / (Bp[0] * mb[0] + Bp[1] * mb[1] + Bp[2] * mb[2]);
      for (int d = 0; d < 3; d++) w[d] = Rp[d] - Pp * Bp[d];
      matvec3(Mi, w, v3);
    } else {
      matvec3(Mi, Rp, v3);
    }
    for (int d = 0; d < 3; d++) {
      L.vn[tid][d] = v3[d];
      a.vp[3L * (vb + tid) + d] = v3[d];
    }
  }
  __syncthreads();
  const double v_et = L.vn[vl][0] * n[0] + L.vn[vl][1] * n[1] + L.vn[vl][2] * n[2];
  const double pl_et = pl - l1 * (v_et - vl_n);
  const double fl[5] = {-v_et, pl_et * n[0], pl_et * n[1], pl_et * n[2], pl_et * v_et};
  const double al = area / a.mass[idl];
  for (int k = 0; k < 5; k++) L.red[k][tid] = -(al * fl[k]);
  if (idr >= 0) {
    const double pr_et = pr + l2 * (v_et - vr_n);
    const double fr[5] = {-v_et, pr_et * n[0], pr_et * n[1], pr_et * n[2], pr_et * v_et};
    const double ar = area / a.mass[idr];
    for (int k = 0; k < 5; k++) L.red[5 + k][tid] = ar * fr[k];
  }
  L.lc[tid] = idl;
  __syncthreads();
  store_slot_terms<5>(m, L, se0, nse, a.T);
}

// rhs of each cell: its slots' terms in increasing vertex id, then
// contributor order, added from zero
template <int NV>
__global__ __launch_bounds__(kCellThreads) void k_sweep_gather(DevMesh m, const double *T, double *out) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.C) return;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0.0;
  for (int j = m.e_voff[c]; j < m.e_voff[c + 1]; j++) {
    const int q = m.c_slot[j];
    const uint32_t src = m.se_src[q];
    const double *t = T + (long)kSweepTerm * kSlotSrc * q;
    for (int e = 0; e < kSlotSrc; e++) {
      if (((src >> (8 * e)) & 0xFFu) == 0xFF) break;
#pragma unroll
      for (int k = 0; k < NV; k++) acc[k] = acc[k] + t[kSweepTerm * e + k];
    }
  }
#pragma unroll
  for (int k = 0; k < NV; k++) out[(long)NV * c + k] = acc[k];
}

// ------------------------------------------------------------------ acoustic
struct AcousticArgs {
  DevMesh m;
  const double *sol;    // [C][4] (p, u, v, w)
  double *T;            // [SE][kSlotSrc][5] terms (4 used)
  double *blk;          // per-group minima of the dt candidates
  const int *bc_kind;   // by tag 1..n_bc: 0 wall, 1 freestream, 2 outflow
  const double *bc_val; // [tag][4]
  int n_bc;
  int *err;
  int per_xcd, nblocks;
};

constexpr double kRhoA = 2.0, kAA = 1.0;

// rotate (acoustic_module.F90:232-262)
SFV_DEV void rotate3(double *v, const double *n1, bool forward) {
  double n2[3] = {n1[2], 0.0, -n1[0]}, n3[3], vt[3];
  const double nn2 = sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
  if (nn2 > 1e-10) {
    for (int d = 0; d < 3; d++) n2[d] = n2[d] / nn2;
  } else {
    n2[0] = 0.0;
    n2[1] = -n1[2];
    n2[2] = n1[1];
    const double q = sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
    for (int d = 0; d < 3; d++) n2[d] = n2[d] / q;
  }
  n3[0] = n1[1] * n2[2] - n1[2] * n2[1];
  n3[1] = -(n1[0] * n2[2] - n1[2] * n2[0]);
  n3[2] = n1[0] * n2[1] - n1[1] * n2[0];
  const double q3 = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
  for (int d = 0; d < 3; d++) n3[d] = n3[d] / q3;
  if (forward) {
    vt[0] = v[0] * n1[0] + v[1] * n1[1] + v[2] * n1[2];
    vt[1] = v[0] * n2[0] + v[1] * n2[1] + v[2] * n2[2];
    vt[2] = v[0] * n3[0] + v[1] * n3[1] + v[2] * n3[2];
  } else {
    for (int d = 0; d < 3; d++) vt[d] = v[0] * n1[d] + v[1] * n2[d] + v[2] * n3[d];
  }
  for (int d = 0; d < 3; d++) v[d] = vt[d];
}

__global__ __launch_bounds__(kGrpThreads) void k_acoustic_group(AcousticArgs a) {
  __shared__ SweepLds L;
  const DevMesh &m = a.m;
  const int lb = xcd_block(a.per_xcd);
  if (lb >= a.nblocks) return;
  const int tid = threadIdx.x;
  const int *gh = m.grp + kGrpHdr * (long)lb;
  const int vb = gh[0], nv = gh[1], sf0 = gh[2], nsf = gh[3], se0 = gh[4], nse = gh[5];
  if (tid <= nv) L.vtab[tid] = m.sf_off[vb + tid] - sf0;
  const bool active = tid < nsf;
  const long s = sf0 + (active ? tid : 0);
  const int f = m.sf_face[s];
  const double n[3] = {m.fnorm[3L * f], m.fnorm[3L * f + 1], m.fnorm[3L * f + 2]};
  const double area = m.sfn[3 * m.SF + s];
  const int il = m.sf_le[s], ir = m.sf_re[s];
  __syncthreads();
  const int vl = seg_of(L.vtab, nv, tid);
  double sl[4], sr[4];
  for (int k = 0; k < 4; k++) sl[k] = a.sol[4L * il + k];
  int kind = -1;
  if (ir == -1) {
    kind = 0;
  } else if (ir <= -2) {
    const int tag = -ir - 1;
    kind = (tag <= a.n_bc) ? a.bc_kind[tag] : 3;
  }
  if (ir >= 0) {
    for (int k = 0; k < 4; k++) sr[k] = a.sol[4L * ir + k];
  } else if (kind == 0) {
    const double un = sl[1] * n[0] + sl[2] * n[1] + sl[3] * n[2];
    sr[0] = sl[0];
    for (int d = 0; d < 3; d++) sr[1 + d] = sl[1 + d] - 2.0 * un * n[d];
  } else if (kind == 1) {
    for (int k = 0; k < 4; k++) sr[k] = a.bc_val[4 * (-ir - 1) + k];
  } else {
    for (int k = 0; k < 4; k++) sr[k] = sl[k];
    if (kind != 2 && active) atomicOr(a.err, (int)kErrBadBC);
  }
  rotate3(sl + 1, n, true);
  rotate3(sr + 1, n, true);
  const double pt = 0.5 * (sl[0] + sr[0]) - 0.5 * kRhoA * kAA * (sr[1] - sl[1]);
  L.red[0][tid] = active ? area * pt : 0.0;
  L.red[1][tid] = active ? area : 0.0;
  __syncthreads();
  for (int p = tid; p < nv * 2; p += kGrpThreads) {
    const int vv = p >> 1, c = p & 1;
    L.vs[vv][c] = seg_sum(L.red[c], L.vtab[vv], L.vtab[vv + 1]);
  }
  __syncthreads();
  const double p_nodal = L.vs[vl][0] / L.vs[vl][1];
  const double smax = kAA;
  const double ul_et = sl[1] - (p_nodal - sl[0]) / (kRhoA * kAA);
  const double ur_et = sr[1] + (p_nodal - sr[0]) / (kRhoA * kAA);
  double fl[4] = {-(kRhoA * (kAA * kAA) * ul_et), -(p_nodal / kRhoA), -0.0, -0.0};
  double fr[4] = {kRhoA * (kAA * kAA) * ur_et, p_nodal / kRhoA, 0.0, 0.0};
  rotate3(fl + 1, n, false);
  rotate3(fr + 1, n, false);
  for (int k = 0; k < 4; k++) L.red[k][tid] = area * fl[k];
  for (int k = 0; k < 4; k++) L.red[4 + k][tid] = area * fr[k];
  L.lc[tid] = il;
  double cand = 1e30;
  if (active) {
    cand = dmin(cand, m.cvol[il] / (m.farea[f] * smax));
    if (ir >= 0) cand = dmin(cand, m.cvol[ir] / (m.farea[f] * smax));
  }
  __syncthreads();
  store_slot_terms<4>(m, L, se0, nse, a.T);
  block_min_store(cand, a.blk, lb);
}

// dt = min over the groups (exact in any order); t = t + dt
__global__ __launch_bounds__(kCellThreads) void k_acoustic_dt(const double *blk, int nblk, double *dt_t) {
  const double mn = reduce_min_one_block(blk, nblk);
  if (threadIdx.x == 0) {
    const double dt = dmin(1e30, mn);
    dt_t[0] = dt;
    dt_t[1] = dt_t[1] + dt;
  }
}

// flux of each cell (its terms from zero), then sol = sol + (cfl dt / vol) flux
__global__ __launch_bounds__(kCellThreads) void k_acoustic_update(DevMesh m, const double *T, double *sol, double cfl,
                                                                  const double *dt_t) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.C) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = m.e_voff[c]; j < m.e_voff[c + 1]; j++) {
    const int q = m.c_slot[j];
    const uint32_t src = m.se_src[q];
    const double *t = T + (long)kSweepTerm * kSlotSrc * q;
    for (int e = 0; e < kSlotSrc; e++) {
      if (((src >> (8 * e)) & 0xFFu) == 0xFF) break;
#pragma unroll
      for (int k = 0; k < 4; k++) acc[k] = acc[k] + t[kSweepTerm * e + k];
    }
  }
  const double w = cfl * dt_t[0] / m.cvol[c];
#pragma unroll
  for (int k = 0; k < 4; k++) sol[4L * c + k] = sol[4L * c + k] + w * acc[k];
}

}  // namespace sfv
