// Cell-side kernels: node->cell gather of the sweep partials, the global dt
// reduction, the stage update (add_timedisc_and_solve) with its positivity
// scaling and flux limiting, and the gradient + limiter kernels of the
// second-order reconstruction.
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

// Global minima are taken in two launches: every block stores its minimum
// (plain stores; the kernel boundary makes them visible), then one block
// reduces them.  A device-wide fence per block would write back the whole
// L2 of the issuing XCD, so none is used.  min is order-independent, so the
// result is exact.
SFV_DEV void block_min_store(double v, double *blk, int bidx) {
  __shared__ double s_m[kCellThreads];
  s_m[threadIdx.x] = v;
  __syncthreads();
  for (int o = kCellThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s_m[threadIdx.x] = dmin(s_m[threadIdx.x], s_m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[bidx] = s_m[0];
}
SFV_DEV void block_min_store(double v, double *blk) { block_min_store(v, blk, blockIdx.x); }

SFV_DEV double reduce_min_one_block(const double *blk, int nblk) {
  __shared__ double s_m[kCellThreads];
  double acc = 1e300;
  for (int i = threadIdx.x; i < nblk; i += kCellThreads) acc = dmin(acc, blk[i]);
  s_m[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kCellThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s_m[threadIdx.x] = dmin(s_m[threadIdx.x], s_m[threadIdx.x + o]);
    __syncthreads();
  }
  return s_m[0];
}

// dt = cfl*min(vol/sum_lambda) (ns_euler_module.F90:546-560; the flux
// limiter's own dt, ns_euler_recon_module.F90:575-581, has no tmax clamp)
SFV_DEV void finalize_dt(const double *blk, int nblk, StepState *st, int fixed_dt) {
  double mn = reduce_min_one_block(blk, nblk);
  if (threadIdx.x != 0) return;
  double d = dmin(1e10, mn);
  st->dtmin = d;
  st->dt_fl = st->cfl * d;
  if (!fixed_dt) {
    double dt = st->cfl * d;
    if (!st->lts) dt = dmin(st->tmax - st->t + 1e-12, dt);
    st->dt = dt;
  }
}
__global__ __launch_bounds__(kCellThreads) void k_finalize_dt(const double *blk, int nblk, StepState *st,
                                                              int fixed_dt) {
  finalize_dt(blk, nblk, st, fixed_dt);
}

// omega checks and time advance of add_timedisc_and_solve (:595-608)
SFV_DEV void finalize_omega(const double *blk, int nblk, StepState *st, int advance_t) {
  double om = reduce_min_one_block(blk, nblk);
  if (threadIdx.x != 0) return;
  if (st->done || st->err) return;
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
  if (advance_t) {
    st->t = st->t + (advance_t == 1 ? dt : st->dt);
    st->steps += 1;
    if (!st->lts && st->t >= st->tmax) st->done = 1;
  }
}
__global__ __launch_bounds__(kCellThreads) void k_finalize_omega(const double *blk, int nblk, StepState *st,
                                                                 int advance_t) {
  finalize_omega(blk, nblk, st, advance_t);
}

struct Init1drp {
  double x1drp, wl[5], wr[5];
};

// init_1drp (ns_euler_module.F90:36-43): U = primit_to_conserv of the left
// state where the centroid's x < x1drp, of the right state elsewhere
__global__ __launch_bounds__(kCellThreads) void k_init_1drp(DevMesh m, Init1drp a, double *U) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.C) return;
  double u[5];
  to_cons(m.ccoord[3 * c] < a.x1drp ? a.wl : a.wr, u);
  for (int k = 0; k < 5; k++) U[5 * c + k] = u[k];
}

struct GatherArgs {
  DevMesh m;
  const double *P;   // first-order partials [SE][6], node-sorted slots
  const double *P2;  // second-order partials or null
  double *rhs;       // [5][C]
  double *rhs2;      // [5][C] (second order)
  double *sl;        // [C] sum_lambda
  double *blk;       // per-block minima
  StepState *st;
  int fixed_dt;      // keep st->dt (SSP-RK3 stages 2 and 3)
  int per_xcd;       // logical cell blocks per XCD slice
  int nblocks;       // logical cell blocks
};

// rhs(:,c) = -sum of the partials of c's vertices in increasing vertex id,
// sum_lambda(c) = 1e-12 + ... (ns_euler_module.F90:181-183, 330-335), and
// the per-block minima of vol/sum_lambda that k_finalize_dt turns into
// dt = cfl*min(vol/sum_lambda) (:546-560).
SFV_DEV void gather_block(const GatherArgs &a, long bidx) {
  const DevMesh &m = a.m;
  long c = bidx * kCellThreads + threadIdx.x;
  double local = 1e10;
  if (c < m.C) {
    const long C = m.C;
    const int b = m.e_voff[c], nv = m.e_voff[c + 1] - b;
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double lam = 1e-12;
    for (int j = 0; j < nv; j++) {
      const double *p = a.P + (long)kSlotStride * m.c_slot[b + j];
      for (int k = 0; k < 5; k++) r[k] = r[k] - p[k];
      lam = lam + p[5];
    }
    for (int k = 0; k < 5; k++) a.rhs[k * C + c] = r[k];
    a.sl[c] = lam;
    if (a.P2) {
      double r2[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int j = 0; j < nv; j++) {
        const double *p = a.P2 + (long)kSlotStride * m.c_slot[b + j];
        for (int k = 0; k < 5; k++) r2[k] = r2[k] - p[k];
      }
      for (int k = 0; k < 5; k++) a.rhs2[k * C + c] = r2[k];
    }
    local = m.cvol[c] / lam;
  }
  block_min_store(local, a.blk, (int)bidx);
}
// Consecutive cells share the cache lines of their vertices' slots, so each
// XCD takes a contiguous range of cell blocks.
__global__ __launch_bounds__(kCellThreads) void k_gather(GatherArgs a) {
  const int lb = xcd_block(a.per_xcd);
  if (lb >= a.nblocks) return;  // uniform over the workgroup
  gather_block(a, lb);
}

struct UpdateArgs {
  DevMesh m;
  const double *U;    // base state [C][5]
  const double *U0;   // combination partner (w0 != 0) or null
  double *Uout;       // [C][5]
  const double *rhs, *rhs2, *sl;
  const double *diag; // [C][25] column-major 5x5 blocks, or null (c*I diagonal)
  const double *rdiff;// [C][4] viscous rhs of components 2..5 (added after flux limiting), or null
  double *blk;
  StepState *st;
  double a1, a2, w0, w1;
  int flux_limit;     // second order: flux_limiting before the update
  int advance_t;      // closes a step: 1 -> t += dt_adv (reference), 2 -> t += dt
};

// flux_limiting (ns_euler_recon_module.F90:586-593) + the per-cell part of
// add_timedisc_and_solve (ns_euler_module.F90:565-580).  Returns the cell's
// omega; u and delta receive the state and the increment.
template <bool DIAG>
SFV_DEV double stage_cell(const UpdateArgs &a, const double *U, long c, double *u, double *delta) {
  const long C = a.m.C;
  const StepState *st = a.st;
  double r[5];
  for (int k = 0; k < 5; k++) {
    u[k] = U[5 * c + k];
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
  if (DIAG && a.rdiff)  // the diffusion block follows flux limiting (ns_main.F90:170-175)
    for (int k = 1; k < 5; k++) r[k] = r[k] + a.rdiff[4 * c + k - 1];
  double cc = a.a1 * (vol / st->dt + (st->lts ? lam : 0.0));
  double nw[5];
  if (DIAG) {
    double D[25];
    for (int q = 0; q < 25; q++) D[q] = a.diag[25 * c + q];
    for (int q = 0; q < 5; q++) D[q + 5 * q] = D[q + 5 * q] + cc;
    double b[5];
    for (int k = 0; k < 5; k++) b[k] = r[k] + cc * u[k];
    lu_solve<5>(D, b, nw);
  } else {
    // the LU solve of cc*I reduces exactly to one division per component
    for (int k = 0; k < 5; k++) nw[k] = (r[k] + cc * u[k]) / cc;
  }
  double ad[5];
  for (int k = 0; k < 5; k++) {
    delta[k] = nw[k] - u[k];
    ad[k] = a.a2 * delta[k];
  }
  return max_omega_pos(u, ad);
}

SFV_DEV void store_stage(const UpdateArgs &a, double *Uout, long c, const double *u, const double *delta,
                         double omega) {
  for (int k = 0; k < 5; k++) {
    double x = u[k] + a.a2 * omega * delta[k];
    if (a.U0) x = a.w0 * a.U0[5 * c + k] + a.w1 * x;
    Uout[5 * c + k] = x;
  }
}

template <bool DIAG>
SFV_DEV void update_block(const UpdateArgs &a, const double *U, double *Uout, long bidx) {
  long c = bidx * kCellThreads + threadIdx.x;
  StepState *st = a.st;
  if (st->done || st->err) {  // uniform over the grid: the state is frozen
    if (c < a.m.C)
      for (int k = 0; k < 5; k++) Uout[5 * c + k] = U[5 * c + k];
    return;
  }
  double om_c = 1.0;
  if (c < a.m.C) {
    double u[5], d[5];
    om_c = stage_cell<DIAG>(a, U, c, u, d);
    store_stage(a, Uout, c, u, d, 1.0);
  }
  block_min_store(om_c, a.blk, (int)bidx);
}

template <bool DIAG>
__global__ __launch_bounds__(kCellThreads) void k_update(UpdateArgs a) {
  update_block<DIAG>(a, a.U, a.Uout, blockIdx.x);
}

// Correction when the global omega came out below one: redo the cell with
// sol = sol + a2*omega*delta (ns_euler_module.F90:608).
template <bool DIAG>
SFV_DEV void update_fixup_block(const UpdateArgs &a, const double *U, double *Uout, long bidx) {
  long c = bidx * kCellThreads + threadIdx.x;
  const StepState *st = a.st;
  if (!st->fixup || st->err) return;
  if (c >= a.m.C) return;
  double u[5], d[5];
  stage_cell<DIAG>(a, U, c, u, d);
  store_stage(a, Uout, c, u, d, st->omega);
}

template <bool DIAG>
__global__ __launch_bounds__(kCellThreads) void k_update_fixup(UpdateArgs a) {
  update_fixup_block<DIAG>(a, a.U, a.Uout, blockIdx.x);
}

// ---------------------------------------------------------------- gradients
struct GradArgs {
  DevMesh m;
  const double *U;   // [C][5]
  double *NG;        // [V][15] nodal gradients, 3*k + d fastest
  double *VMM;       // [V][4] rho_min, rho_max, p_min, p_max over the vertex's cells
  double *G;         // [C][15]
  double *NGP;       // [V][9] pinv of sum a n (x) (x_R - x_L): geometry only, cached per mesh
  const double *W;   // [C][6] primitives of U (k_prim), or null
  int method;
  int per_xcd, nblocks;    // vertex groups (k_nodal_grad)
  int cper_xcd, cblocks;   // cell blocks (k_cell_grad)
};

// ROWS staging rows of per-sub-face terms, SUMW per-vertex sums
template <int ROWS, int SUMW>
struct GradLds {
  double red[ROWS][kRedStride];
  double sum[kGrpVert][SUMW];
  int vtab[kGrpVert + 1];
  uint8_t vbound[kGrpVert];
};

// compute_nodal_grad (ns_euler_recon_module.F90:453-490) over the vertex
// groups of the node sweep, thread = sub-face, and the per-vertex bounds of
// relaxed_maximum_principle (:82-97) taken over the cells on either side of
// the vertex's sub-faces (exactly the cells around the vertex on a
// non-periodic mesh; min and max do not depend on the order).  Sums run in
// sub-face order (seg_sum); the 3x3 pseudo-inverses of the group run in the
// first wavefront.
// MODE: kNGFull forms and pseudo-inverts the geometric matrix of every
// vertex; kNGStore does the same and keeps the pseudo-inverse in a.NGP;
// kNGCached reads it from a.NGP instead (the matrix depends on the mesh only,
// so the three give the same bits).  Cached, the gradient terms are staged
// in two rounds (rho, u, v; then w, p) in 9 rows, so the workgroup's LDS
// is 26 KB instead of 47 KB and six workgroups share a CU instead of three.
enum { kNGFull = 0, kNGStore = 1, kNGCached = 2 };

template <int MODE>
__global__ __launch_bounds__(kGrpThreads, 6) void k_nodal_grad(GradArgs a) {  // six waves per SIMD: 76 VGPRs
  constexpr bool kCached = MODE == kNGCached;
  constexpr int kRows = kCached ? 9 : 16, kB = kCached ? 0 : 9;  // rows; first bounds row
  __shared__ GradLds<kRows, kCached ? 15 : 28> L;
  const DevMesh &m = a.m;
  const int lb = xcd_block(a.per_xcd);
  if (lb >= a.nblocks) return;  // uniform over the workgroup
  const int tid = threadIdx.x;
  const int *gh = m.grp + kGrpHdr * (long)lb;
  const int vb = gh[0], nv = gh[1], sf0 = gh[2], nsf = gh[3];
  if (tid <= nv) L.vtab[tid] = m.sf_off[vb + tid] - sf0;
  if (tid < nv) L.vbound[tid] = m.vbound[vb + tid];
  const bool active = tid < nsf;
  const long s = sf0 + (active ? tid : 0);
  const long C = m.C, SF = m.SF;
  const int le = m.sf_le[s], re = m.sf_re[s];
  const bool inner = active && re >= 0;
  double wl[5], wr[5];
  if (a.W) {
    for (int k = 0; k < 5; k++) {
      wl[k] = a.W[6L * le + k];
      wr[k] = a.W[6L * (re >= 0 ? re : le) + k];
    }
  } else {
    double ul[5], ur[5];
    for (int k = 0; k < 5; k++) {
      ul[k] = a.U[5L * le + k];
      ur[k] = a.U[5L * (re >= 0 ? re : le) + k];
    }
    to_prim(ul, wl);
    to_prim(ur, wr);
  }
  const double n[3] = {m.sfn[s], m.sfn[SF + s], m.sfn[2 * SF + s]};
  const double ar = m.sfn[3 * SF + s];
  // pass 1: sum a n (x) (x_R - x_L) (unless cached) and the bounds
  if (MODE != kNGCached)
    for (int b = 0; b < 3; b++) {
      double dx = inner ? m.ccoord[3L * re + b] - m.ccoord[3L * le + b] : 0.0;
      for (int d = 0; d < 3; d++) L.red[d + 3 * b][tid] = inner ? ar * (n[d] * dx) : 0.0;
    }
  // bounds: the left cell always, the right cell when it is a cell
  L.red[kB][tid] = active ? dmin(wl[0], inner ? wr[0] : 1e100) : 1e100;
  L.red[kB + 1][tid] = active ? dmax(wl[0], inner ? wr[0] : 0.0) : 0.0;
  L.red[kB + 2][tid] = active ? dmin(wl[4], inner ? wr[4] : 1e100) : 1e100;
  L.red[kB + 3][tid] = active ? dmax(wl[4], inner ? wr[4] : 0.0) : 0.0;
  __syncthreads();
  if (MODE != kNGCached)
    for (int p = tid; p < nv * 9; p += kGrpThreads) {
      const int vv = p / 9, c = p - vv * 9;
      L.sum[vv][15 + c] = seg_sum(L.red[c], L.vtab[vv], L.vtab[vv + 1]);
    }
  for (int p = tid; p < nv * 4; p += kGrpThreads) {
    const int vv = p >> 2, c = p & 3;
    double mm;
    if (m.sf_n0) {
      // periodic mesh: a stitched sub-face reaches a cell of the partner
      // vertex, so the bounds come from the vertex's own cells
      mm = (c & 1) ? 0.0 : 1e100;
      for (int q = m.v_eoff[vb + vv]; q < m.v_eoff[vb + vv + 1]; q++) {
        double u[5], w[5];
        for (int k = 0; k < 5; k++) u[k] = a.U[5L * m.v_elem[q] + k];
        to_prim(u, w);
        const double x = (c < 2) ? w[0] : w[4];
        mm = (c & 1) ? dmax(mm, x) : dmin(mm, x);
      }
    } else {
      const int jb = L.vtab[vv], je = L.vtab[vv + 1];
      const double *row = L.red[kB + c];
      mm = row[jb];
      if (c & 1)
        for (int j = jb + 1; j < je; j++) mm = dmax(mm, row[j]);
      else
        for (int j = jb + 1; j < je; j++) mm = dmin(mm, row[j]);
      // the bounds start from 1e100 / 0.0 (ns_euler_recon_module.F90:83-86)
      mm = (c & 1) ? dmax(0.0, mm) : dmin(1e100, mm);
    }
    a.VMM[4L * (vb + vv) + c] = mm;
  }
  __syncthreads();
  // pass 2: sum a n (x) (W_R - W_L), variables k0 .. k1 - 1 per round
#pragma unroll
  for (int k0 = 0; k0 < 5; k0 += kCached ? 3 : 5) {
    const int k1 = kCached ? min(k0 + 3, 5) : 5;
    if (k0 > 0) __syncthreads();  // the previous round's sums have read the rows
#pragma unroll
    for (int k = k0; k < k1; k++) {
      double dw = wr[k] - wl[k];
      for (int d = 0; d < 3; d++) L.red[3 * (k - k0) + d][tid] = inner ? ar * (n[d] * dw) : 0.0;
    }
    __syncthreads();
    const int nc = 3 * (k1 - k0);
    for (int p = tid; p < nv * nc; p += kGrpThreads) {
      const int vv = p / nc, c = p - vv * nc;
      L.sum[vv][3 * k0 + c] = seg_sum(L.red[c], L.vtab[vv], L.vtab[vv + 1]);
    }
  }
  __syncthreads();
  if (tid < nv) {  // first wavefront only (nv <= 64)
    const double *S = L.sum[tid];
    const bool bnd = L.vbound[tid];
    double M[9];
    if (MODE == kNGCached) {
      for (int q = 0; q < 9; q++) M[q] = bnd ? 0.0 : a.NGP[9L * (vb + tid) + q];
    } else {
      for (int q = 0; q < 9; q++) M[q] = S[15 + q];
      pinv3(M);
      if (MODE == kNGStore)
        for (int q = 0; q < 9; q++) a.NGP[9L * (vb + tid) + q] = M[q];
    }
    double *o = a.NG + 15L * (vb + tid);
    for (int k = 0; k < 5; k++)
      for (int d = 0; d < 3; d++) {
        double x = M[d] * S[3 * k] + M[d + 3] * S[3 * k + 1] + M[d + 6] * S[3 * k + 2];
        o[3 * k + d] = bnd ? 0.0 : x;
      }
  }
}

SFV_DEV double oi3(const double *g) { return sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]); }

// compute_cell_grad_method_{2..6} (:166-320), physical_slope_limiter
// (:129-164) and relaxed_maximum_principle (:61-127)
// METHOD >= 0: compiled for that method alone (fewer registers); -1: method
template <int METHOD>
__global__ __launch_bounds__(kCellThreads) void k_cell_grad(GradArgs a) {
  const int method = METHOD >= 0 ? METHOD : a.method;
  const DevMesh &m = a.m;
  const int lb = xcd_block(a.cper_xcd);  // neighbouring cells share their vertices' gradients
  if (lb >= a.cblocks) return;
  long e = (long)lb * blockDim.x + threadIdx.x;
  if (e >= m.C) return;
  const long C = m.C;
  double g[15];
  for (int q = 0; q < 15; q++) g[q] = 0.0;
  const int vb = m.e_voff[e], nv = m.e_voff[e + 1] - vb;
  double *go = a.G + 15L * e;
  if (method == 0 || method == 1) {
    for (int q = 0; q < 15; q++) go[q] = 0.0;
    return;
  }
  const double eps = 1e-3;
  if (method == 2) {
    double dvs = 0.0;
    for (int j = 0; j < nv; j++) {
      int v = m.e_vert[vb + j];
      double vv = m.vvol[v];
      const double *ng = a.NG + 15L * v;
      for (int q = 0; q < 15; q++) g[q] = g[q] + vv * ng[q];
      dvs = dvs + vv;
    }
    for (int q = 0; q < 15; q++) g[q] = g[q] / dvs;
  } else if (method == 3) {
    int v = m.e_vert[vb];
    for (int q = 0; q < 15; q++) g[q] = a.NG[15L * v + q];
    for (int j = 1; j < nv; j++) {
      v = m.e_vert[vb + j];
      const double *ng = a.NG + 15L * v;
      for (int k = 0; k < 5; k++)
        if (oi3(&ng[3 * k]) < oi3(&g[3 * k]))
          for (int d = 0; d < 3; d++) g[3 * k + d] = ng[3 * k + d];
    }
  } else {
    double ws[5] = {0, 0, 0, 0, 0}, cg[15], cws = 0.0;
    for (int q = 0; q < 15; q++) cg[q] = 0.0;
    for (int j = 0; j < nv; j++) {
      int v = m.e_vert[vb + j];
      double vv = m.vvol[v];
      double ng[15];
      for (int q = 0; q < 15; q++) ng[q] = a.NG[15L * v + q];
      for (int k = 0; k < 5; k++) {
        double den = eps + oi3(&ng[3 * k]);
        if (method != 4) den = den * den;
        double wgt = vv / den;
        for (int d = 0; d < 3; d++) g[3 * k + d] = g[3 * k + d] + wgt * ng[3 * k + d];
        ws[k] = ws[k] + wgt;
      }
      if (method == 6) {
        for (int q = 0; q < 15; q++) cg[q] = cg[q] + vv * ng[q];
        cws = cws + vv;
      }
    }
    if (method == 6) {
      for (int q = 0; q < 15; q++) cg[q] = cg[q] / cws;
      // the reference passes the whole 3 x 5 central gradient to OI, whose
      // dummy is dimension(3): sequence association hands it the first
      // variable's gradient, so one weight serves all five variables
      const double den1 = eps + oi3(&cg[0]);
      for (int k = 0; k < 5; k++) {
        double den = den1;
        den = den * den;
        for (int d = 0; d < 3; d++) g[3 * k + d] = g[3 * k + d] + cws / den * cg[3 * k + d];
        ws[k] = ws[k] + cws / den;
      }
    }
    for (int k = 0; k < 5; k++)
      for (int d = 0; d < 3; d++) g[3 * k + d] = g[3 * k + d] / ws[k];
  }
  // physical_slope_limiter, with the neighbourhood bounds gathered alongside
  const double tol = 1e-8;
  double u[5], w[5];
  for (int k = 0; k < 5; k++) u[k] = a.U[5 * e + k];
  to_prim(u, w);
  const double xc[3] = {m.ccoord[3L * e], m.ccoord[3L * e + 1], m.ccoord[3L * e + 2]};
  double omega = 1.0;
  double rmin = 1e100, rmax = 0.0, pmin = 1e100, pmax = 0.0;
  for (int j = 0; j < nv; j++) {
    int v = m.e_vert[vb + j];
    double dx[3];
    for (int d = 0; d < 3; d++) dx[d] = m.vcoord[3L * v + d] - xc[d];
    double d0 = g[0] * dx[0] + g[1] * dx[1] + g[2] * dx[2];
    double d4 = g[12] * dx[0] + g[13] * dx[1] + g[14] * dx[2];
    if (d0 < 0.0) omega = dmin(omega, dmax(0.0, w[0] / (tol - d0)));
    if (d4 < 0.0) omega = dmin(omega, dmax(0.0, w[4] / (tol - d4)));
    const double *mm = a.VMM + 4L * v;
    rmin = dmin(rmin, mm[0]);
    rmax = dmax(rmax, mm[1]);
    pmin = dmin(pmin, mm[2]);
    pmax = dmax(pmax, mm[3]);
  }
  for (int q = 0; q < 15; q++) g[q] = omega * g[q];
  // relaxed_maximum_principle on the limited gradient
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
  for (int q = 0; q < 15; q++) go[q] = omega * g[q];
}

}  // namespace sfv
