// Cell-side kernels: node->cell gather of the sweep partials, the global dt
// reduction, the stage update (add_timedisc_and_solve) with its positivity
// scaling, flux limiting, and the gradient + limiter kernels of the second
// order reconstruction.
#pragma once
#include "device_mesh.hpp"
#include "kernels_flux.hpp"
#include "physics.hpp"

namespace sfv {

// Device-resident step state: the scalars the reference keeps on the host
// (t, dt, omega) live here so a whole step runs without a host round trip.
struct StepState {
  double t, dt, dt_fl, omega, tmax, cfl;
  double dt_adv;       // dt the step advances t by: (1 - omega)*dt when omega < 1
  double dtmin;        // min(vol / sum_lambda) of the last gather
  double omega_min;    // min omega of the last update
  int err;             // sfv::Err bits, first failure sticks
  int fixup;           // last update needs the omega < 1 correction
  int done;            // t >= tmax reached (non-local time stepping)
  int lts;             // local_time_step
  unsigned int ticket; // last-block-done counter
  int steps;           // completed steps
};

constexpr int kCellThreads = 256;

// Block min of one value per thread, then the last block to finish reduces
// the per-block minima.  Returns true in the thread that holds the final
// value in *out_all (thread 0 of the last block).
SFV_DEV bool block_min_last(double v, double *blk, unsigned int *ticket, int nblk, double *out_all) {
  __shared__ double s_m[kCellThreads];
  __shared__ bool s_last;
  s_m[threadIdx.x] = v;
  __syncthreads();
  for (int o = kCellThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s_m[threadIdx.x] = dmin(s_m[threadIdx.x], s_m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    blk[blockIdx.x] = s_m[0];
    __threadfence();
    unsigned int t = atomicAdd(ticket, 1u);
    s_last = (t == (unsigned int)nblk - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  double acc = 1e300;
  for (int i = threadIdx.x; i < nblk; i += kCellThreads) acc = dmin(acc, ((volatile double *)blk)[i]);
  s_m[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kCellThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s_m[threadIdx.x] = dmin(s_m[threadIdx.x], s_m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *out_all = s_m[0];
    *ticket = 0u;
  }
  return threadIdx.x == 0;
}

struct GatherArgs {
  DevMesh m;
  const double *P;   // first-order partials [6][SE]
  const double *P2;  // second-order partials or null
  double *rhs;       // [5][C]
  double *rhs2;      // [5][C] (second order)
  double *sl;        // [C] sum_lambda
  double *blk;       // per-block minima
  StepState *st;
  int fixed_dt;      // keep st->dt (SSP-RK3 stages 2 and 3)
};

// rhs(:,c) = -sum of the partials of c's vertices in increasing vertex id,
// sum_lambda(c) = 1e-12 + ... (ns_euler_module.F90:181-183, 330-335), then
// dt = cfl*min(vol/sum_lambda) (:546-560) via the last block.
__global__ __launch_bounds__(kCellThreads) void k_gather(GatherArgs a) {
  const DevMesh &m = a.m;
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double local = 1e10;
  if (c < m.C) {
    const long SE = m.SE, C = m.C;
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double lam = 1e-12;
    for (int q = m.e_voff[c]; q < m.e_voff[c + 1]; q++) {
      int slot = m.c_slot[q];
      for (int k = 0; k < 5; k++) r[k] = r[k] - a.P[k * SE + slot];
      lam = lam + a.P[5 * SE + slot];
    }
    for (int k = 0; k < 5; k++) a.rhs[k * C + c] = r[k];
    a.sl[c] = lam;
    if (a.P2) {
      double r2[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int q = m.e_voff[c]; q < m.e_voff[c + 1]; q++) {
        int slot = m.c_slot[q];
        for (int k = 0; k < 5; k++) r2[k] = r2[k] - a.P2[k * SE + slot];
      }
      for (int k = 0; k < 5; k++) a.rhs2[k * C + c] = r2[k];
    }
    local = m.cvol[c] / lam;
  }
  double mn;
  int nblk = gridDim.x;
  if (block_min_last(local, a.blk, &a.st->ticket, nblk, &mn)) {
    StepState *st = a.st;
    double d = dmin(1e10, mn);
    st->dtmin = d;
    st->dt_fl = st->cfl * d;
    if (!a.fixed_dt) {
      double dt = st->cfl * d;
      if (!st->lts) dt = dmin(st->tmax - st->t + 1e-12, dt);
      st->dt = dt;
    }
  }
}

struct UpdateArgs {
  DevMesh m;
  const double *U;    // base state [5][C]
  const double *U0;   // combination partner (w0 != 0) or null
  double *Uout;       // [5][C]
  const double *rhs, *rhs2, *sl;
  const double *diag; // [25][C] or null (pure c*I diagonal)
  double *blk;
  StepState *st;
  double a1, a2, w0, w1;
  int flux_limit;     // second order: flux_limiting before the update
  int advance_t;      // closes a step: 1 -> t += dt_adv (reference), 2 -> t += dt
  int fixup_pass;     // 0 speculative (omega = 1), 1 correction pass
};

// flux_limiting (ns_euler_recon_module.F90:586-593) + the per-cell part of
// add_timedisc_and_solve (ns_euler_module.F90:565-580).  Returns the cell's
// omega and the increment.
SFV_DEV double stage_cell(const UpdateArgs &a, long c, double *u, double *delta) {
  const long C = a.m.C;
  const StepState *st = a.st;
  double r[5];
  for (int k = 0; k < 5; k++) {
    u[k] = a.U[k * C + c];
    r[k] = a.rhs[k * C + c];
  }
  const double vol = a.m.cvol[c];
  const double lam = a.sl[c];
  if (a.flux_limit) {
    double r2[5], x[5], y[5];
    double cf = vol / st->dt_fl + lam;
    for (int k = 0; k < 5; k++) {
      r2[k] = a.rhs2[k * C + c];
      x[k] = u[k] + r[k] / cf;
      y[k] = (r2[k] - r[k]) / cf;
    }
    double om = max_omega_pos(x, y);
    for (int k = 0; k < 5; k++) r[k] = r[k] + dmin(1.0, om) * (r2[k] - r[k]);
  }
  double cc = a.a1 * (vol / st->dt + (st->lts ? lam : 0.0));
  double nw[5];
  if (a.diag) {
    double D[25];
    for (int q = 0; q < 25; q++) D[q] = a.diag[q * C + c];
    for (int q = 0; q < 5; q++) D[q + 5 * q] = D[q + 5 * q] + cc;
    double b[5];
    for (int k = 0; k < 5; k++) b[k] = r[k] + cc * u[k];
    lu_solve<5>(D, b, nw);
  } else {
    // LU of cc*I reduces exactly to one division per component
    for (int k = 0; k < 5; k++) nw[k] = (r[k] + cc * u[k]) / cc;
  }
  double ad[5];
  for (int k = 0; k < 5; k++) {
    delta[k] = nw[k] - u[k];
    ad[k] = a.a2 * delta[k];
  }
  return max_omega_pos(u, ad);
}

SFV_DEV void store_stage(const UpdateArgs &a, long c, const double *u, const double *delta, double omega) {
  const long C = a.m.C;
  for (int k = 0; k < 5; k++) {
    double x = u[k] + a.a2 * omega * delta[k];
    if (a.U0) x = a.w0 * a.U0[k * C + c] + a.w1 * x;
    a.Uout[k * C + c] = x;
  }
}

__global__ __launch_bounds__(kCellThreads) void k_update(UpdateArgs a) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  StepState *st = a.st;
  if (st->done || st->err) {  // uniform over the grid: the state is frozen
    if (c < a.m.C)
      for (int k = 0; k < 5; k++) a.Uout[k * a.m.C + c] = a.U[k * a.m.C + c];
    return;
  }
  double om_c = 1.0;
  if (c < a.m.C) {
    double u[5], d[5];
    om_c = stage_cell(a, c, u, d);
    store_stage(a, c, u, d, 1.0);
  }
  double om;
  if (block_min_last(om_c, a.blk, &st->ticket, gridDim.x, &om)) {
    double omega = dmin(1.0, om);
    st->omega_min = om;
    st->omega = omega;
    if (omega <= 1e-6) {
      st->err |= kErrBadOmega;
      st->fixup = 0;
      return;
    }
    st->fixup = omega < 1.0;
    double dt = st->dt;
    if (!st->lts && omega < 1.0) dt = (1.0 - omega) * dt;
    st->dt_adv = dt;
    if (a.advance_t) {
      st->t = st->t + (a.advance_t == 1 ? dt : st->dt);
      st->steps += 1;
      if (!st->lts && st->t >= st->tmax) st->done = 1;
    }
  }
}

// Correction when the global omega came out below one: redo the cell with
// sol = sol + a2*omega*delta (ns_euler_module.F90:608).
__global__ __launch_bounds__(kCellThreads) void k_update_fixup(UpdateArgs a) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const StepState *st = a.st;
  if (!st->fixup || st->err) return;
  if (c >= a.m.C) return;
  double u[5], d[5];
  stage_cell(a, c, u, d);
  store_stage(a, c, u, d, st->omega);
}

// ---------------------------------------------------------------- gradients
struct GradArgs {
  DevMesh m;
  const double *U;   // [5][C]
  double *NG;        // [15][V] nodal gradients, (3k+d) major
  double *VMM;       // [4][V] rho_min, rho_max, p_min, p_max over vertex cells
  double *G;         // [15][C]
  int method;
};

// compute_nodal_grad (ns_euler_recon_module.F90:453-490) and the per-vertex
// part of the bounds of relaxed_maximum_principle (:82-97)
__global__ __launch_bounds__(kCellThreads) void k_nodal_grad(GradArgs a) {
  const DevMesh &m = a.m;
  long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= m.V) return;
  const long C = m.C, SF = m.SF;
  double ng[15], M[9];
  for (int i = 0; i < 15; i++) ng[i] = 0.0;
  for (int i = 0; i < 9; i++) M[i] = 0.0;
  for (int s = m.sf_off[v]; s < m.sf_off[v + 1]; s++) {
    int re = m.sf_re[s];
    if (re < 0) continue;
    int le = m.sf_le[s];
    double ul[5], ur[5], wl[5], wr[5];
    for (int k = 0; k < 5; k++) {
      ul[k] = a.U[k * C + le];
      ur[k] = a.U[k * C + re];
    }
    to_prim(ul, wl);
    to_prim(ur, wr);
    double n[3] = {m.sfn[s], m.sfn[SF + s], m.sfn[2 * SF + s]};
    double ar = m.sfn[3 * SF + s];
    for (int k = 0; k < 5; k++) {
      double dw = wr[k] - wl[k];
      for (int d = 0; d < 3; d++) ng[3 * k + d] = ng[3 * k + d] + ar * (n[d] * dw);
    }
    for (int b = 0; b < 3; b++) {
      double dx = m.ccoord[3L * re + b] - m.ccoord[3L * le + b];
      for (int d = 0; d < 3; d++) M[d + 3 * b] = M[d + 3 * b] + ar * (n[d] * dx);
    }
  }
  pinv3(M);
  const bool bnd = m.vbound[v];
  for (int k = 0; k < 5; k++)
    for (int d = 0; d < 3; d++) {
      double o = M[d] * ng[3 * k] + M[d + 3] * ng[3 * k + 1] + M[d + 6] * ng[3 * k + 2];
      a.NG[(3 * k + d) * m.V + v] = bnd ? 0.0 : o;
    }
  double rmin = 1e100, rmax = 0.0, pmin = 1e100, pmax = 0.0;
  for (int q = m.v_eoff[v]; q < m.v_eoff[v + 1]; q++) {
    int e = m.v_elem[q];
    double u[5], w[5];
    for (int k = 0; k < 5; k++) u[k] = a.U[k * C + e];
    to_prim(u, w);
    pmin = dmin(pmin, w[4]);
    pmax = dmax(pmax, w[4]);
    rmin = dmin(rmin, w[0]);
    rmax = dmax(rmax, w[0]);
  }
  a.VMM[v] = rmin;
  a.VMM[m.V + v] = rmax;
  a.VMM[2 * m.V + v] = pmin;
  a.VMM[3 * m.V + v] = pmax;
}

SFV_DEV double oi3(const double *g) { return sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]); }

// compute_cell_grad_method_{2..6} (:166-320), physical_slope_limiter
// (:129-164) and relaxed_maximum_principle (:61-127)
__global__ __launch_bounds__(kCellThreads) void k_cell_grad(GradArgs a) {
  const DevMesh &m = a.m;
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.C) return;
  const long C = m.C, V = m.V;
  double g[15];
  for (int q = 0; q < 15; q++) g[q] = 0.0;
  const int vb = m.e_voff[e], nv = m.e_voff[e + 1] - vb;
  if (a.method == 0 || a.method == 1) {
    for (int q = 0; q < 15; q++) a.G[q * C + e] = 0.0;
    return;
  }
  const double eps = 1e-3;
  if (a.method == 2) {
    double dvs = 0.0;
    for (int j = 0; j < nv; j++) {
      int v = m.e_vert[vb + j];
      double vv = m.vvol[v];
      for (int q = 0; q < 15; q++) g[q] = g[q] + vv * a.NG[q * V + v];
      dvs = dvs + vv;
    }
    for (int q = 0; q < 15; q++) g[q] = g[q] / dvs;
  } else if (a.method == 3) {
    int v = m.e_vert[vb];
    for (int q = 0; q < 15; q++) g[q] = a.NG[q * V + v];
    for (int j = 1; j < nv; j++) {
      v = m.e_vert[vb + j];
      for (int k = 0; k < 5; k++) {
        double ngk[3] = {a.NG[(3 * k) * V + v], a.NG[(3 * k + 1) * V + v], a.NG[(3 * k + 2) * V + v]};
        if (oi3(ngk) < oi3(&g[3 * k]))
          for (int d = 0; d < 3; d++) g[3 * k + d] = ngk[d];
      }
    }
  } else {
    double ws[5] = {0, 0, 0, 0, 0}, cg[15], cws = 0.0;
    for (int q = 0; q < 15; q++) cg[q] = 0.0;
    for (int j = 0; j < nv; j++) {
      int v = m.e_vert[vb + j];
      double vv = m.vvol[v];
      double ng[15];
      for (int q = 0; q < 15; q++) ng[q] = a.NG[q * V + v];
      for (int k = 0; k < 5; k++) {
        double den = eps + oi3(&ng[3 * k]);
        if (a.method != 4) den = den * den;
        double wgt = vv / den;
        for (int d = 0; d < 3; d++) g[3 * k + d] = g[3 * k + d] + wgt * ng[3 * k + d];
        ws[k] = ws[k] + wgt;
      }
      if (a.method == 6) {
        for (int q = 0; q < 15; q++) cg[q] = cg[q] + vv * ng[q];
        cws = cws + vv;
      }
    }
    if (a.method == 6) {
      for (int q = 0; q < 15; q++) cg[q] = cg[q] / cws;
      for (int k = 0; k < 5; k++) {
        double den = eps + oi3(&cg[3 * k]);
        den = den * den;
        for (int d = 0; d < 3; d++) g[3 * k + d] = g[3 * k + d] + cws / den * cg[3 * k + d];
        ws[k] = ws[k] + cws / den;
      }
    }
    for (int k = 0; k < 5; k++)
      for (int d = 0; d < 3; d++) g[3 * k + d] = g[3 * k + d] / ws[k];
  }
  // physical_slope_limiter
  const double tol = 1e-8;
  double u[5], w[5];
  for (int k = 0; k < 5; k++) u[k] = a.U[k * C + e];
  to_prim(u, w);
  const double *xc = &m.ccoord[3L * e];
  double omega = 1.0;
  for (int j = 0; j < nv; j++) {
    int v = m.e_vert[vb + j];
    double dx[3];
    for (int d = 0; d < 3; d++) dx[d] = m.vcoord[3L * v + d] - xc[d];
    double d0 = g[0] * dx[0] + g[1] * dx[1] + g[2] * dx[2];
    double d4 = g[12] * dx[0] + g[13] * dx[1] + g[14] * dx[2];
    if (d0 < 0.0) omega = dmin(omega, dmax(0.0, w[0] / (tol - d0)));
    if (d4 < 0.0) omega = dmin(omega, dmax(0.0, w[4] / (tol - d4)));
  }
  for (int q = 0; q < 15; q++) g[q] = omega * g[q];
  // relaxed_maximum_principle
  double rmin = 1e100, rmax = 0.0, pmin = 1e100, pmax = 0.0;
  for (int j = 0; j < nv; j++) {
    int v = m.e_vert[vb + j];
    rmin = dmin(rmin, a.VMM[v]);
    rmax = dmax(rmax, a.VMM[V + v]);
    pmin = dmin(pmin, a.VMM[2 * V + v]);
    pmax = dmax(pmax, a.VMM[3 * V + v]);
  }
  pmin = (1.0 - 1e-2) * pmin;
  rmin = (1.0 - 1e-2) * rmin;
  pmax = (1.0 + 1e-2) * pmax;
  rmax = (1.0 + 1e-2) * rmax;
  omega = 1.0;
  for (int j = 0; j < nv; j++) {
    int v = m.e_vert[vb + j];
    double dx[3];
    for (int d = 0; d < 3; d++) dx[d] = m.vcoord[3L * v + d] - xc[d];
    double d0 = g[0] * dx[0] + g[1] * dx[1] + g[2] * dx[2];
    double d4 = g[12] * dx[0] + g[13] * dx[1] + g[14] * dx[2];
    if (d0 < 0.0)
      omega = dmin(omega, dmax(0.0, (w[0] - rmin) / (tol - d0)));
    else
      omega = dmin(omega, dmax(0.0, -(w[0] - rmax) / (tol + d0)));
    if (d4 < 0.0)
      omega = dmin(omega, dmax(0.0, (w[4] - pmin) / (tol - d4)));
    else
      omega = dmin(omega, dmax(0.0, -(w[4] - pmax) / (tol + d4)));
  }
  for (int q = 0; q < 15; q++) a.G[q * C + e] = omega * g[q];
}

}  // namespace sfv
