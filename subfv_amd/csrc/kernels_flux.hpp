// Node-centred subface sweep: compute_euler_flux / compute_rhs_around_vert
// (src/ns/ns_euler_module.F90:151-339, 341-519) with the nodal solvers of
// src/ns/ns_euler_rs_module.F90:288-553, 681-758.
//
// Mapping: one 256-thread workgroup owns a group of consecutive vertices
// whose sub-faces (contiguous, vertex-major) fit in its lanes; thread =
// sub-face.  Every per-vertex sum (nodal matrix and right-hand side, flux per
// sub-element) is formed in reference order by one thread reading the
// vertex's sub-face terms from LDS (loads batched, additions in order), so
// results are deterministic and atomic-free.  The 3x3 nodal solves of the
// group run in the first wavefront only (one lane per vertex), so the other
// wavefronts issue no solve instructions.
//
// Output: one row of six partial sums per node-sorted sub-element slot q
// (vertex-major, so a group's rows are one contiguous run): P[6*q + k] = sum
// over the vertex's sub-faces of area*flux_k (k < 5) and P[6*q + 5] = sum of
// area*max(0, -s).  Thread q of the group owns slot q; its contributors (the
// group-local sub-faces touching the slot's cell, in increasing order) come
// from the static table se_src.
#pragma once
#include "device_mesh.hpp"
#include "physics.hpp"

namespace sfv {

struct FluxArgs {
  DevMesh m;
  BCParams bc;
  const double *U;      // [C][5] conservative state, sol(5, n_elems)
  const double *G;      // [C][15] cell gradients, 3*k + d fastest (second order)
  const double *W;      // [C][6] primitives and sound speed of U (k_prim), or null
  double *P;            // [SE][6] node-sorted slots
  double *vp;           // [3][V] nodal velocity output or null
  int *err;
  const double *t;      // physical time (time-dependent BCs)
  int second_order;
  int per_xcd;          // logical blocks per XCD slice
  int nblocks;          // logical blocks (= vertex groups)
};

SFV_DEV void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kFluxWaves = 4;  // wavefronts per workgroup of the wavefront-segment kernels

// blockIdx -> logical block so that each XCD gets a contiguous slice of the
// (spatially ordered) work and its L2 sees neighbouring cells.
SFV_DEV int xcd_block(int per_xcd) {
  int b = blockIdx.x;
  return (b & 7) * per_xcd + (b >> 3);
}

// In-order sum row[b] + ... + row[e-1] starting from 0.0; loads are issued
// eight at a time, the additions stay sequential (adding the +0.0 padding
// leaves every partial sum unchanged).
SFV_DEV double seg_sum(const double *row, int b, int e) {
  double acc = 0.0;
  for (int j0 = b; j0 < e; j0 += 8) {
    double t[8];
#pragma unroll
    for (int q = 0; q < 8; q++) t[q] = (j0 + q < e) ? row[j0 + q] : 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) acc = acc + t[q];
  }
  return acc;
}

// largest i in [0, n) with tab[i] <= x (tab increasing, tab[0] = 0 <= x)
SFV_DEV int seg_of(const int *tab, int n, int x) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

struct FluxLds {
  double red[12][kRedStride];
  double sum[kGrpVert][9];
  double vn[kGrpVert][3];
  double bp[kGrpVert][3];
  int vtab[kGrpVert + 1];
  int lc[kGrpThreads];
  uint8_t vbound[kGrpVert];
};

// Nodal velocity of one vertex from its summed matrix S (6 distinct entries)
// and right-hand side R (S[6..8]): ns_euler_rs_module.F90:406-421.
SFV_DEV void nodal_velocity(const BCParams &bc, const double *S, const double *B, double *vn3) {
  double M[9] = {S[0], S[1], S[2], S[1], S[3], S[4], S[2], S[4], S[5]};
  const double R[3] = {S[6], S[7], S[8]};
  if (bc.bc_style == 2) {
    const double mb = dmax(dmax(fabs(B[0]), fabs(B[1])), fabs(B[2]));
    if (mb > 1e-8) {
      pinv3_sym(S, M);
      double mR[3], mB[3];
      for (int i = 0; i < 3; i++) {
        mR[i] = M[i] * R[0] + M[i + 3] * R[1] + M[i + 6] * R[2];
        mB[i] = M[i] * B[0] + M[i + 3] * B[1] + M[i + 6] * B[2];
      }
      const double pb = (mR[0] * B[0] + mR[1] * B[1] + mR[2] * B[2]) / (mB[0] * B[0] + mB[1] * B[1] + mB[2] * B[2]);
      double rr[3];
      for (int d = 0; d < 3; d++) rr[d] = R[d] - pb * B[d];
      for (int i = 0; i < 3; i++) vn3[i] = M[i] * rr[0] + M[i + 3] * rr[1] + M[i + 6] * rr[2];
      const double sc = (vn3[0] * B[0] + vn3[1] * B[1] + vn3[2] * B[2]) / (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
      for (int d = 0; d < 3; d++) vn3[d] = vn3[d] - sc * B[d];
    } else {
      lu_solve<3>(M, R, vn3);
    }
  } else {
    pinv3_sym(S, M);
    for (int i = 0; i < 3; i++) vn3[i] = M[i] * R[0] + M[i + 3] * R[1] + M[i + 6] * R[2];
  }
}

// The sweep of one vertex group lb of the mesh in a, state U.
template <int SCHEME>
SFV_DEV void node_flux_group(const FluxArgs &a, const double *__restrict__ U, int lb, FluxLds &L) {
  const DevMesh &m = a.m;
  const BCParams &bc = a.bc;
  const int tid = threadIdx.x;
  const int *gh = m.grp + kGrpHdr * (long)lb;
  const int vb = gh[0], nv = gh[1], sf0 = gh[2], nsf = gh[3], se0 = gh[4], nse = gh[5];

  // ---- stage per-vertex data; issue the sub-face and slot loads
  if (tid <= nv) L.vtab[tid] = m.sf_off[vb + tid] - sf0;
  if (tid < nv) {
    L.vbound[tid] = m.vbound[vb + tid];
    if (SCHEME == kMultiPoint || SCHEME == kMultiPointIso) {
      L.bp[tid][0] = m.Bp[vb + tid];
      L.bp[tid][1] = m.Bp[m.V + vb + tid];
      L.bp[tid][2] = m.Bp[2 * m.V + vb + tid];
    }
  }
  const bool active = tid < nsf;
  // inactive threads mirror the group's first sub-face so every load is in
  // bounds; their contributions are masked below
  const long s = sf0 + (active ? tid : 0);
  double n[3], area;
  n[0] = m.sfn[s];
  n[1] = m.sfn[m.SF + s];
  n[2] = m.sfn[2 * m.SF + s];
  area = m.sfn[3 * m.SF + s];
  const int le = m.sf_le[s];
  const int re = m.sf_re[s];
  const bool has_slot = tid < nse;
  uint32_t src = 0xFFFFFFFFu;
  int slot_cell = -1;
  if (has_slot) {
    src = m.se_src[se0 + tid];
    slot_cell = m.se_cell[se0 + tid];
  }
  __syncthreads();
  const int vl = seg_of(L.vtab, nv, tid);
  const int v = vb + vl;

  // ---- left/right primitive states (reconstruct_lr_w)
  const long C = m.C;
  double wl[5], wr[5];
  double al_c = 0.0, ar_c = 0.0;  // the cells' sound speeds from W (first order)
  const double *W = a.W;
  {
    double u[5];
    if (W) {
      for (int k = 0; k < 5; k++) wl[k] = W[6L * le + k];
      al_c = W[6L * le + 5];
    } else {
      for (int k = 0; k < 5; k++) u[k] = U[5L * le + k];
      to_prim(u, wl);
    }
    if (a.second_order) {
      double dx[3];
      const double *g = a.G + 15L * le;
      for (int d = 0; d < 3; d++) dx[d] = m.vcoord[3L * v + d] - m.ccoord[3L * le + d];
      for (int k = 0; k < 5; k++) wl[k] = wl[k] + (g[3 * k] * dx[0] + g[3 * k + 1] * dx[1] + g[3 * k + 2] * dx[2]);
    }
    if (re >= 0) {
      if (W) {
        for (int k = 0; k < 5; k++) wr[k] = W[6L * re + k];
        ar_c = W[6L * re + 5];
      } else {
        for (int k = 0; k < 5; k++) u[k] = U[5L * re + k];
        to_prim(u, wr);
      }
      if (a.second_order) {
        double dx[3];
        const double *g = a.G + 15L * re;
        for (int d = 0; d < 3; d++) dx[d] = m.vcoord[3L * v + d] - m.ccoord[3L * re + d];
        for (int k = 0; k < 5; k++)
          wr[k] = wr[k] + (g[3 * k] * dx[0] + g[3 * k + 1] * dx[1] + g[3 * k + 2] * dx[2]);
      }
    } else {
      double ul[5], ug[5];
      to_cons(wl, ul);
      const double *fc = &m.fcoord[3L * m.sf_face[s]];
      int e = ghost_state(bc, re, n, fc, a.t ? *a.t : 0.0, ul, ug);
      if (e && active) atomicOr(a.err, e);
      to_prim(ug, wr);
    }
  }
  const bool wall = is_wall(bc, re);
  const bool excl = bc.excl && L.vbound[vl];

  // ---- nodal solve
  double lam_l = 0.0, lam_r = 0.0, vbar = 0.0;
  double v_node[3] = {0.0, 0.0, 0.0};
  double p_nodal = 0.0;
  double warea = 1.0;
  // the cached sound speeds hold for unreconstructed cell states only
  const bool al_cached = W && !a.second_order, ar_cached = al_cached && re >= 0;
  if (SCHEME == kMultiPoint || SCHEME == kMultiPointIso) {
    const bool iso = (SCHEME == kMultiPointIso);
    const double rhol = wl[0], pl = wl[4], al = al_cached ? al_c : sound_speed(wl);
    const double rhor = wr[0], pr = wr[4], ar = ar_cached ? ar_c : sound_speed(wr);
    const double vn_l = wl[1] * n[0] + wl[2] * n[1] + wl[3] * n[2];
    const double vn_r = wr[1] * n[0] + wr[2] * n[1] + wr[3] * n[2];
    for (int iter = 1; iter <= 4; iter++) {
      if (iter == 1) {
        lam_l = dmax(dmax(dmax(lam_l, al * rhol), sqrt(rhol * dmax(0.0, pr - pl))), -rhol * (vn_r - vn_l));
        lam_r = dmax(dmax(dmax(lam_r, ar * rhor), sqrt(rhor * dmax(0.0, pl - pr))), -rhor * (vn_r - vn_l));
        vbar = (lam_l * vn_l + lam_r * vn_r - (pr - pl)) / (lam_r + lam_l);
        lam_l = dmax(lam_l, al * rhol * (1.0 + 1.5 * dmax(0.0, -(vbar - vn_l) / al)));
        lam_r = dmax(lam_r, ar * rhor * (1.0 + 1.5 * dmax(0.0, (vbar - vn_r) / ar)));
      } else if (wall && bc.bc_style == 1) {
        lam_l = dmax(lam_l, al * rhol * (1.0 + 1.5 * dmax(0.0, -(vbar - vn_l) / al)));
        lam_r = dmax(lam_r, ar * rhor * (1.0 + 1.5 * dmax(0.0, (vbar - vn_r) / ar)));
      } else {
        double v_et = v_node[0] * n[0] + v_node[1] * n[1] + v_node[2] * n[2];
        lam_l = dmax(lam_l, al * rhol * (1.0 + 1.5 * dmax(0.0, -(v_et - vn_l) / al)));
        lam_r = dmax(lam_r, ar * rhor * (1.0 + 1.5 * dmax(0.0, (v_et - vn_r) / ar)));
      }
      vbar = (lam_l * vn_l + lam_r * vn_r - (pr - pl)) / (lam_r + lam_l);
      // contributions to the nodal matrix (6 distinct entries) and Rp
      double cm = iso ? (lam_r + lam_l) : (lam_r + lam_l) * area;
      double cr = iso ? (lam_r + lam_l) * vbar : area * (lam_r + lam_l) * vbar;
      bool addm = true, addr = true;
      if (wall) {
        if (bc.bc_style == 0 || bc.bc_style == 2) {
          if (bc.bc_style == 2) {
            addr = bc.b2d ? (fabs(n[2]) < 1e-8) : true;
            cr = iso ? pl : area * pl;
          }
        } else {
          addm = addr = false;
        }
      }
      if (!active || excl) addm = addr = false;
      L.red[0][tid] = addm ? cm * (n[0] * n[0]) : 0.0;
      L.red[1][tid] = addm ? cm * (n[0] * n[1]) : 0.0;
      L.red[2][tid] = addm ? cm * (n[0] * n[2]) : 0.0;
      L.red[3][tid] = addm ? cm * (n[1] * n[1]) : 0.0;
      L.red[4][tid] = addm ? cm * (n[1] * n[2]) : 0.0;
      L.red[5][tid] = addm ? cm * (n[2] * n[2]) : 0.0;
      L.red[6][tid] = addr ? cr * n[0] : 0.0;
      L.red[7][tid] = addr ? cr * n[1] : 0.0;
      L.red[8][tid] = addr ? cr * n[2] : 0.0;
      __syncthreads();
      for (int p = tid; p < nv * 9; p += kGrpThreads) {
        const int vv = p / 9, c = p - vv * 9;
        L.sum[vv][c] = seg_sum(L.red[c], L.vtab[vv], L.vtab[vv + 1]);
      }
      __syncthreads();
      if (tid < nv) {  // first wavefront only (nv <= 64)
        double vn3[3] = {0.0, 0.0, 0.0};
        if (!(bc.excl && L.vbound[tid])) {
          const double B[3] = {L.bp[tid][0], L.bp[tid][1], L.bp[tid][2]};
          nodal_velocity(bc, L.sum[tid], B, vn3);
        }
        for (int d = 0; d < 3; d++) L.vn[tid][d] = vn3[d];
      }
      __syncthreads();
      for (int d = 0; d < 3; d++) v_node[d] = L.vn[vl][d];
    }
    if (a.vp && tid < nv)
      for (int d = 0; d < 3; d++) a.vp[d * m.V + vb + tid] = L.vn[tid][d];
    if (iso) {
      // warea = minval(area)/area over the vertex's sub-faces
      L.red[0][tid] = active ? area : 1e300;
      __syncthreads();
      if (tid < nv) {
        double mn = L.red[0][L.vtab[tid]];
        for (int j = L.vtab[tid] + 1; j < L.vtab[tid + 1]; j++) mn = dmin(mn, L.red[0][j]);
        L.sum[tid][0] = mn;
      }
      __syncthreads();
      warea = L.sum[vl][0] / area;
    }
  } else if (SCHEME == kMultiPointPressure || SCHEME == kMultiPointPressurePh) {
    const double rhol = wl[0], pl = wl[4], al = al_cached ? al_c : sound_speed(wl);
    const double rhor = wr[0], pr = wr[4], ar = ar_cached ? ar_c : sound_speed(wr);
    const double vn_l = wl[1] * n[0] + wl[2] * n[1] + wl[3] * n[2];
    const double vn_r = wr[1] * n[0] + wr[2] * n[1] + wr[3] * n[2];
    const bool w1 = (bc.bc_style == 1 && wall);
    for (int iter = 1; iter <= 10; iter++) {
      if (iter == 1) {
        lam_l = dmax(dmax(dmax(lam_l, al * rhol), sqrt(rhol * dmax(0.0, pr - pl))), -rhol * (vn_r - vn_l));
        lam_r = dmax(dmax(dmax(lam_r, ar * rhor), sqrt(rhor * dmax(0.0, pl - pr))), -rhor * (vn_r - vn_l));
        vbar = (lam_r * pl + lam_l * pr - lam_r * lam_l * (vn_r - vn_l)) / (lam_r + lam_l);
        lam_l = dmax(lam_l, al * rhol + 1.5 * sqrt(dmax(0.0, (vbar - pl) * rhol)));
        lam_r = dmax(lam_r, ar * rhor + 1.5 * sqrt(dmax(0.0, (vbar - pr) * rhor)));
      } else if (w1) {
        lam_l = dmax(lam_l, al * rhol + 1.5 * sqrt(dmax(0.0, (vbar - pl) * rhol)));
        lam_r = dmax(lam_r, ar * rhor + 1.5 * sqrt(dmax(0.0, (vbar - pr) * rhor)));
      } else {
        lam_l = dmax(lam_l, al * rhol + 1.5 * sqrt(dmax(0.0, (p_nodal - pl) * rhol)));
        lam_r = dmax(lam_r, ar * rhor + 1.5 * sqrt(dmax(0.0, (p_nodal - pr) * rhor)));
      }
      vbar = (lam_r * pl + lam_l * pr - lam_r * lam_l * (vn_r - vn_l)) / (lam_r + lam_l);
      double c1, c2;
      if (w1) {
        c1 = (1.0 / lam_r + 1.0 / lam_l) * area * (lam_r * pl + lam_r * lam_l * vn_l) / (lam_r + lam_l);
        c2 = (1.0 / lam_l) * area;
      } else {
        c1 = (1.0 / lam_r + 1.0 / lam_l) * area * vbar;
        c2 = (1.0 / lam_r + 1.0 / lam_l) * area;
      }
      L.red[0][tid] = active ? c1 : 0.0;
      L.red[1][tid] = active ? c2 : 0.0;
      __syncthreads();
      if (tid < nv) {
        const double s1 = seg_sum(L.red[0], L.vtab[tid], L.vtab[tid + 1]);
        const double s2 = seg_sum(L.red[1], L.vtab[tid], L.vtab[tid + 1]);
        L.sum[tid][0] = s1 / s2;
      }
      __syncthreads();
      p_nodal = L.sum[vl][0];
    }
  }

  // ---- sub-face flux
  double fx[10], sl = 0.0, sr = 0.0;
  if (SCHEME == kMultiPoint || SCHEME == kMultiPointIso) {
    if (excl) {
      flux_two_wave(wl, wr, n, fx, sl, sr);
    } else {
      double vn = wall ? 0.0 : (v_node[0] * n[0] + v_node[1] * n[1] + v_node[2] * n[2]);
      bool ok = flux_multi_point(wl, wr, n, fx, vn, lam_l, lam_r, sl, sr);
      if (!ok && active) atomicOr(a.err, (int)kErrNegStarDensity);
    }
    if (SCHEME == kMultiPointIso) {
      double f3[10];
      flux_three_wave<false>(wl, wr, n, f3, sl, sr);
      for (int k = 0; k < 10; k++) fx[k] = warea * fx[k] + (1.0 - warea) * f3[k];
    }
  } else if (SCHEME == kThreeWave) {
    flux_three_wave<false>(wl, wr, n, fx, sl, sr);
  } else if (SCHEME == kModifiedThreeWave) {
    flux_three_wave<true>(wl, wr, n, fx, sl, sr);
  } else if (SCHEME == kTwoWave) {
    flux_two_wave(wl, wr, n, fx, sl, sr);
  } else if (SCHEME == kMultiPointPressure) {
    flux_multi_point_pressure<false>(wl, wr, n, fx, p_nodal, lam_l, lam_r, sl, sr);
  } else {
    flux_multi_point_pressure<true>(wl, wr, n, fx, p_nodal, lam_l, lam_r, sl, sr);
  }

  // ---- area*flux per sub-element slot (ns_euler_module.F90:506-517)
  for (int k = 0; k < 5; k++) {
    L.red[k][tid] = area * fx[k];
    L.red[6 + k][tid] = area * fx[5 + k];
  }
  L.red[5][tid] = area * dmax(0.0, -sl);
  L.red[11][tid] = area * dmax(0.0, sr);
  L.lc[tid] = le;
  __syncthreads();
  if (has_slot) {
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int e = 0; e < kSlotSrc; e++) {
      const int j = (src >> (8 * e)) & 0xFFu;
      if (j == 0xFF) break;
      const int r0 = (L.lc[j] == slot_cell) ? 0 : 6;
#pragma unroll
      for (int c = 0; c < 6; c++) acc[c] = acc[c] + L.red[r0 + c][j];
    }
    double *out = a.P + (long)kSlotStride * (se0 + tid);
#pragma unroll
    for (int c = 0; c < 6; c++) out[c] = acc[c];
  }
}

// W[6 c + k] = conserv_to_primit(U_c)(k), W[6 c + 5] = sound_speed_w of it:
// the cell values every sub-face of the sweeps would otherwise recompute
static __global__ __launch_bounds__(256) void k_prim(long C, const double *__restrict__ U, double *__restrict__ W) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  double u[5], w[5];
  for (int k = 0; k < 5; k++) u[k] = U[5 * c + k];
  to_prim(u, w);
  for (int k = 0; k < 5; k++) W[6 * c + k] = w[k];
  W[6 * c + 5] = sound_speed(w);
}

template <int SCHEME>
__global__ __launch_bounds__(kGrpThreads, 3) void k_node_flux(FluxArgs a) {
  __shared__ FluxLds L;
  const int lb = xcd_block(a.per_xcd);
  if (lb >= a.nblocks) return;  // uniform over the workgroup
  node_flux_group<SCHEME>(a, a.U, lb, L);
}

}  // namespace sfv
