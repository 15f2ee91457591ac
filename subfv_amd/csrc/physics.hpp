// Device-side FP64 physics of subfv's explicit residual: primitives, BC ghost
// states, the subface Riemann fluxes and the small dense solves of the nodal
// solver.  Expression order follows the cited reference lines so that, built
// without floating-point contraction, results match the reference
// arithmetic operation for operation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SFV_DEV __device__ __forceinline__

namespace sfv {

constexpr double kGamma = 1.4;   // ns_global_data_module.F90:9
constexpr double kGm1 = kGamma - 1;

enum Scheme : int {
  kMultiPoint = 1,
  kMultiPointIso = 2,
  kThreeWave = 3,
  kTwoWave = 4,
  kModifiedThreeWave = 5,
  kMultiPointPressure = 6,
  kMultiPointPressurePh = 8
};
enum BCKind : int {
  kBCWall = 1,
  kBCAdherenceWall = 2,
  kBCFreestream = 3,
  kBCOutflowSupersonic = 4,
  kBCInflowPond = 5,
  kBCInoutDoubleMach = 6,
  kBCDoubleMachBottom = 7,
  kBCPotentialFlow2D = 8,
  kBCPotentialFlow3D = 9
};
enum Err : int { kErrNegStarDensity = 1, kErrBadOmega = 2, kErrBadBC = 4 };

struct BCParams {
  int scheme, bc_style, b2d, excl, n_bc, lts;
  int bc_id[11];
  double bc_w[11][5];
};

SFV_DEV double dmax(double a, double b) { return a > b ? a : b; }
SFV_DEV double dmin(double a, double b) { return a < b ? a : b; }

// conserv_to_primit / primit_to_conserv, ns_euler_primitives_module.F90:7-31
SFV_DEV void to_prim(const double *u, double *w) {
  w[0] = u[0];
  w[1] = u[1] / u[0];
  w[2] = u[2] / u[0];
  w[3] = u[3] / u[0];
  w[4] = kGm1 * (u[4] - 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]));
}
SFV_DEV void to_cons(const double *w, double *u) {
  u[0] = w[0];
  u[1] = w[1] * w[0];
  u[2] = w[2] * w[0];
  u[3] = w[3] * w[0];
  u[4] = w[4] / kGm1 + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
}
SFV_DEV double sound_speed(const double *w) { return sqrt(kGamma * w[4] / w[0]); }

// max_omega_for_positivity, ns_euler_primitives_module.F90:92-111
SFV_DEV double max_omega_pos(const double *U, const double *dU) {
  double om = 1.0, us[5], ws[5];
  for (int k = 0; k < 5; k++) us[k] = U[k] + om * dU[k];
  to_prim(us, ws);
  bool pos = (ws[0] > 1e-17 && ws[4] > 1e-17);
  while (!pos && om > 1e-6) {
    om = om * 0.9;
    for (int k = 0; k < 5; k++) us[k] = U[k] + om * dU[k];
    to_prim(us, ws);
    pos = (ws[0] > 1e-17 && ws[4] > 1e-17);
  }
  return om;
}

// right-neighbour code on the device: >= 0 cell index, -1 default wall,
// <= -2 boundary condition (-code - 1).  is_wall: ns_euler_recon_module.F90:411-425
SFV_DEV bool is_wall(const BCParams &bc, int re) {
  if (re >= 0) return false;
  if (re == -1) return true;
  return bc.bc_id[-re - 1] == kBCWall;
}

SFV_DEV void reflect(const double *n, const double *ul, double *ur) {
  for (int k = 0; k < 5; k++) ur[k] = ul[k];
  double vn = ur[1] * n[0] + ur[2] * n[1] + ur[3] * n[2];
  for (int d = 0; d < 3; d++) ur[1 + d] = ur[1 + d] - 2.0 * vn * n[d];
}

// compute_right_state, ns_euler_recon_module.F90:331-409 (conserved in/out)
// sol_potential_flow_2d / _3d (ns_euler_exact_sol_module.F90:105-170):
// potential flow around a cylinder (2-D, r0 = 0.5, outer radius 5.5, through
// atan2 / cos / sin) or a sphere (3-D, r0 = 0.5) at Mach 1e-4, primitive state
// at the face centroid x; integer powers as products (r^5 = ((r r)(r r)) r).
SFV_DEV void potential_flow(bool three_d, const double *x, double *w) {
  const double rhoinf = 1.4, mach = 1e-4, pinf = 1.0, r0 = 0.5;
  const double vinf = mach * sqrt(kGamma * pinf / rhoinf);
  double ux, uy, uz = 0.0, vmag2;
  if (!three_d) {
    const double r1 = 5.5;
    const double r = sqrt(x[0] * x[0] + x[1] * x[1]);
    const double theta = atan2(x[1], x[0]);
    const double k = (r1 * r1) / (r1 * r1 - r0 * r0), q = (r0 / r) * (r0 / r);
    const double vr = k * (1.0 - q) * cos(theta);
    const double vt = -k * (1.0 + q) * sin(theta);
    ux = vinf * (cos(theta) * vr - sin(theta) * vt);
    uy = vinf * (sin(theta) * vr + cos(theta) * vt);
    vmag2 = ux * ux + uy * uy;
  } else {
    const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const double f = (r0 * r0 * r0) / (2.0 * (((r * r) * (r * r)) * r));
    ux = vinf * (1.0 + f * (r * r - 3.0 * (x[0] * x[0])));
    uy = -3.0 * vinf * f * x[0] * x[1];
    uz = -3.0 * vinf * f * x[0] * x[2];
    vmag2 = ux * ux + uy * uy + uz * uz;
  }
  w[0] = rhoinf;
  w[1] = ux;
  w[2] = uy;
  w[3] = uz;
  w[4] = pinf + 0.5 * rhoinf * (vinf * vinf - vmag2);
}

SFV_DEV int ghost_state(const BCParams &bc, int re, const double *n, const double *fc, double t,
                        const double *ul, double *ur) {
  if (re == -1) {
    reflect(n, ul, ur);
    return 0;
  }
  int ib = -re - 1;
  switch (bc.bc_id[ib]) {
    case kBCOutflowSupersonic:
      for (int k = 0; k < 5; k++) ur[k] = ul[k];
      return 0;
    case kBCFreestream:
      to_cons(bc.bc_w[ib], ur);
      return 0;
    case kBCInflowPond:
      to_cons(bc.bc_w[ib], ur);
      if (ur[1] * n[0] + ur[2] * n[1] + ur[3] * n[2] > 0.0)
        for (int k = 0; k < 5; k++) ur[k] = ul[k];
      return 0;
    case kBCWall:
      reflect(n, ul, ur);
      return 0;
    case kBCAdherenceWall: {
      for (int k = 0; k < 5; k++) ur[k] = ul[k];
      for (int d = 0; d < 3; d++) ur[1 + d] = 2.0 * ul[0] * bc.bc_w[ib][1 + d] - ul[1 + d];
      double a0 = ul[1] / ul[0], a1 = ul[2] / ul[0], a2 = ul[3] / ul[0];
      double b0 = ur[1] / ur[0], b1 = ur[2] / ur[0], b2 = ur[3] / ur[0];
      double na = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
      double nb = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
      ur[4] = ul[4] - 0.5 * ul[0] * (na * na) + 0.5 * ul[0] * (nb * nb);
      return 0;
    }
    case kBCInoutDoubleMach: {
      const double wa[5] = {8.0, 7.145, -4.125, 0.0, 116.5};
      const double wb[5] = {1.4, 0.0, 0.0, 0.0, 1.0};
      if (fc[1] > 1.732 * (fc[0] - 0.1667 - 10.0 * t))
        to_cons(wa, ur);
      else
        to_cons(wb, ur);
      return 0;
    }
    case kBCDoubleMachBottom: {
      const double wa[5] = {8.0, 7.145, -4.125, 0.0, 116.5};
      if (fc[0] < 0.1667)
        to_cons(wa, ur);
      else
        reflect(n, ul, ur);
      return 0;
    }
    case kBCPotentialFlow2D:
    case kBCPotentialFlow3D: {
      double w[5];
      potential_flow(bc.bc_id[ib] == kBCPotentialFlow3D, fc, w);
      to_cons(w, ur);
      return 0;
    }
    default:
      for (int k = 0; k < 5; k++) ur[k] = ul[k];
      return kErrBadBC;
  }
}

// One side of a subface: normal velocity, conserved state, physical flux.
struct Side {
  double rho, vn, p, e, a, u[5], f[5];
};
SFV_DEV void side(const double *w, const double *n, Side &s, bool need_a) {
  s.rho = w[0];
  s.vn = w[1] * n[0] + w[2] * n[1] + w[3] * n[2];
  s.p = w[4];
  to_cons(w, s.u);
  s.e = s.u[4] / s.rho;
  s.a = need_a ? sound_speed(w) : 0.0;
  s.f[0] = s.vn * s.u[0];
  for (int d = 0; d < 3; d++) s.f[1 + d] = s.vn * s.u[1 + d] + s.p * n[d];
  s.f[4] = (s.u[4] + s.p) * s.vn;
}

SFV_DEV void two_point_lambdas(const Side &L, const Side &R, double &ll, double &lr) {
  ll = dmax(dmax(L.a * L.rho, sqrt(L.rho * dmax(0.0, R.p - L.p))), -L.rho * (R.vn - L.vn));
  lr = dmax(dmax(R.a * R.rho, sqrt(R.rho * dmax(0.0, L.p - R.p))), -R.rho * (R.vn - L.vn));
}

// two_wave, ns_euler_rs_module.F90:154-198; out[0..4] left, out[5..9] right
SFV_DEV void flux_two_wave(const double *wl, const double *wr, const double *n, double *out, double &sl,
                           double &sr) {
  Side L, R;
  side(wl, n, L, true);
  side(wr, n, R, true);
  double ll, lr;
  two_point_lambdas(L, R, ll, lr);
  sl = L.vn - ll / L.rho;
  sr = R.vn + lr / R.rho;
  for (int k = 0; k < 5; k++) {
    double et = (sr * R.u[k] - sl * L.u[k] - (R.f[k] - L.f[k])) / (sr - sl);
    out[k] = 0.5 * (L.f[k] + R.f[k]) - 0.5 * (fabs(sl) * (et - L.u[k]) + fabs(sr) * (R.u[k] - et));
    out[5 + k] = -out[k];
  }
}

// three_wave (:9-75) and modified_three_wave (:77-152)
template <bool MOD>
SFV_DEV void flux_three_wave(const double *wl, const double *wr, const double *n, double *out, double &sl,
                             double &sr) {
  Side L, R;
  side(wl, n, L, true);
  side(wr, n, R, true);
  double ll, lr;
  two_point_lambdas(L, R, ll, lr);
  double v_et = (ll * L.vn + lr * R.vn - (R.p - L.p)) / (lr + ll);
  double rl = 1.0 / (1.0 / L.rho + (v_et - L.vn) / ll);
  double pl_bar = L.p - ll * (v_et - L.vn);
  double rr = 1.0 / (1.0 / R.rho + (R.vn - v_et) / lr);
  double pr_bar = R.p + lr * (v_et - R.vn);
  double ql[5], qr[5];
  ql[0] = rl;
  qr[0] = rr;
  if (MOD) {
    double vtl[3], vtr[3], vte[3], dvt[3];
    for (int d = 0; d < 3; d++) {
      vtl[d] = wl[1 + d] - L.vn * n[d];
      vtr[d] = wr[1 + d] - R.vn * n[d];
    }
    for (int d = 0; d < 3; d++) {
      vte[d] = (ll * vtl[d] + lr * vtr[d]) / (ll + lr);
      dvt[d] = vtr[d] - vtl[d];
    }
    double tc = -ll * lr / (ll + lr);
    double te = tc * (dvt[0] * vte[0] + dvt[1] * vte[1] + dvt[2] * vte[2]);
    for (int d = 0; d < 3; d++) {
      ql[1 + d] = rl * (v_et * n[d] + vte[d]);
      qr[1 + d] = rr * (v_et * n[d] + vte[d]);
    }
    ql[4] = rl * (L.e + (L.p * L.vn - pl_bar * v_et - te) / ll);
    qr[4] = rr * (R.e + (pr_bar * v_et - R.p * R.vn + te) / lr);
  } else {
    for (int d = 0; d < 3; d++) {
      ql[1 + d] = rl * (wl[1 + d] + (v_et - L.vn) * n[d]);
      qr[1 + d] = rr * (wr[1 + d] + (v_et - R.vn) * n[d]);
    }
    ql[4] = rl * (L.e + (L.p * L.vn - pl_bar * v_et) / ll);
    qr[4] = rr * (R.e + (pr_bar * v_et - R.p * R.vn) / lr);
  }
  sl = L.vn - ll / L.rho;
  sr = R.vn + lr / R.rho;
  for (int k = 0; k < 5; k++) {
    out[k] = 0.5 * (L.f[k] + R.f[k]) -
             0.5 * (fabs(sl) * (ql[k] - L.u[k]) + fabs(v_et) * (qr[k] - ql[k]) + fabs(sr) * (R.u[k] - qr[k]));
    out[5 + k] = -out[k];
  }
}

// multi_point, ns_euler_rs_module.F90:200-286.  Returns false on a negative
// star density (the reference stops there).
SFV_DEV bool flux_multi_point(const double *wl, const double *wr, const double *n, double *out, double v_et,
                              double ll, double lr, double &sl, double &sr) {
  Side L, R;
  side(wl, n, L, false);
  side(wr, n, R, false);
  double rl = 1.0 / (1.0 / L.rho + (v_et - L.vn) / ll);
  double pl_bar = L.p - ll * (v_et - L.vn);
  double rr = 1.0 / (1.0 / R.rho + (R.vn - v_et) / lr);
  double pr_bar = R.p + lr * (v_et - R.vn);
  double ql[5], qr[5];
  ql[0] = rl;
  qr[0] = rr;
  for (int d = 0; d < 3; d++) {
    ql[1 + d] = rl * (wl[1 + d] + (v_et - L.vn) * n[d]);
    qr[1 + d] = rr * (wr[1 + d] + (v_et - R.vn) * n[d]);
  }
  ql[4] = rl * (L.e + (L.p * L.vn - pl_bar * v_et) / ll);
  qr[4] = rr * (R.e + (pr_bar * v_et - R.p * R.vn) / lr);
  sl = L.vn - ll / L.rho;
  sr = R.vn + lr / R.rho;
  const double e[5] = {0.0, n[0], n[1], n[2], v_et};
  for (int k = 0; k < 5; k++) {
    double c = 0.5 * (L.f[k] + R.f[k]) -
               0.5 * (fabs(sl) * (ql[k] - L.u[k]) + fabs(v_et) * (qr[k] - ql[k]) + fabs(sr) * (R.u[k] - qr[k]));
    out[k] = c - 0.5 * (pr_bar - pl_bar) * e[k];
    out[5 + k] = -(c + 0.5 * (pr_bar - pl_bar) * e[k]);
  }
  return !(rl < 0.0 || rr < 0.0);
}

// multi_point_pressure (:555-616) and _ph (:618-679)
template <bool PH>
SFV_DEV void flux_multi_point_pressure(const double *wl, const double *wr, const double *n, double *out,
                                       double p_nodal, double ll, double lr, double &sl, double &sr) {
  Side L, R;
  side(wl, n, L, false);
  side(wr, n, R, false);
  double vtl[3], vtr[3];
  for (int d = 0; d < 3; d++) {
    vtl[d] = wl[1 + d] - L.vn * n[d];
    vtr[d] = wr[1 + d] - R.vn * n[d];
  }
  double vn_bar = (ll * L.vn + lr * R.vn - (R.p - L.p)) / (lr + ll);
  double vle = L.vn - (p_nodal - L.p) / ll;
  double rl = 1.0 / (1.0 / L.rho + (vn_bar - L.vn) / ll);
  double vre = R.vn + (p_nodal - R.p) / lr;
  double rr = 1.0 / (1.0 / R.rho + (R.vn - vn_bar) / lr);
  double ql[5], qr[5];
  ql[0] = rl;
  qr[0] = rr;
  for (int d = 0; d < 3; d++) {
    ql[1 + d] = rl * (vtl[d] + vle * n[d]);
    qr[1 + d] = rr * (vtr[d] + vre * n[d]);
  }
  if (PH) {
    ql[4] = rl * (L.e + (L.p * L.vn - p_nodal * vle) / ll);
    qr[4] = rr * (R.e + (p_nodal * vre - R.p * R.vn) / lr);
  } else {
    ql[4] = rl * (L.e + (L.p * L.vn - p_nodal * vn_bar) / ll);
    qr[4] = rr * (R.e + (p_nodal * vn_bar - R.p * R.vn) / lr);
  }
  sl = L.vn - ll / L.rho;
  sr = R.vn + lr / R.rho;
  for (int k = 0; k < 5; k++) {
    out[k] = 0.5 * (L.f[k] + R.f[k]) -
             0.5 * (fabs(sl) * (ql[k] - L.u[k]) + fabs(vn_bar) * (qr[k] - ql[k]) + fabs(sr) * (R.u[k] - qr[k]));
    out[5 + k] = -out[k];
  }
}

// ------------------------------------------------------------ dense solves
// LU with partial pivoting in the order of reference LAPACK dgetrf2/dgetrs
// (pivot = first largest |a|; L scaled by the reciprocal of the pivot;
// rank-1 updates; dtrsm-ordered substitutions).  A column-major, n <= 5.
template <int N>
SFV_DEV void lu_solve(double *A, const double *b, double *x) {
  int piv[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    int p = k;
    double amax = fabs(A[k + N * k]);
#pragma unroll
    for (int i = k + 1; i < N; i++)
      if (fabs(A[i + N * k]) > amax) {
        amax = fabs(A[i + N * k]);
        p = i;
      }
    piv[k] = p;
    if (amax != 0.0) {  // amax = |A(p, k)|
      // row interchange with static indices only (no scratch for A)
#pragma unroll
      for (int i = k + 1; i < N; i++)
        if (i == p) {
#pragma unroll
          for (int j = 0; j < N; j++) {
            double t = A[k + N * j];
            A[k + N * j] = A[i + N * j];
            A[i + N * j] = t;
          }
        }
      double pv = A[k + N * k];
      if (fabs(pv) >= 2.2250738585072014e-308) {
        double r = 1.0 / pv;
#pragma unroll
        for (int i = k + 1; i < N; i++) A[i + N * k] = A[i + N * k] * r;
      } else {
#pragma unroll
        for (int i = k + 1; i < N; i++) A[i + N * k] = A[i + N * k] / pv;
      }
    }
#pragma unroll
    for (int j = k + 1; j < N; j++)
#pragma unroll
      for (int i = k + 1; i < N; i++) A[i + N * j] = A[i + N * j] - A[i + N * k] * A[k + N * j];
  }
  double w[N];
#pragma unroll
  for (int i = 0; i < N; i++) w[i] = b[i];
#pragma unroll
  for (int k = 0; k < N; k++)
    if (piv[k] != k) {
      double t = w[k];
      // dynamic index into a small register array
      double q = w[0];
#pragma unroll
      for (int i = 0; i < N; i++)
        if (i == piv[k]) q = w[i];
      w[k] = q;
#pragma unroll
      for (int i = 0; i < N; i++)
        if (i == piv[k]) w[i] = t;
    }
#pragma unroll
  for (int k = 0; k < N; k++)
    if (w[k] != 0.0)
#pragma unroll
      for (int i = k + 1; i < N; i++) w[i] = w[i] - w[k] * A[i + N * k];
#pragma unroll
  for (int k = N - 1; k >= 0; k--)
    if (w[k] != 0.0) {
      w[k] = w[k] / A[k + N * k];
#pragma unroll
      for (int i = 0; i < k; i++) w[i] = w[i] - w[k] * A[i + N * k];
    }
#pragma unroll
  for (int i = 0; i < N; i++) x[i] = w[i];
}

// Cofactor inverse of the 3x3 matrix with rows (a b c), (d e f), (g h i),
// written column-major into P, when it is provably far from the
// pseudo-inverse cut-off; false (P untouched) otherwise.  With F the
// Frobenius norm (F2 = F^2), s_max <= F and s_min = |det| / (s_max s_mid) >=
// 2 |det| / F^2, so |det| > 2e-12 F^3 gives s_min > 3.9e-12 s_max: every
// singular value passes the 3e-12 * max cut-off of
// linear_solver_module.F90:703-711 with a margin far above the rounding of
// det (< 3e-15 F^3), and the pseudo-inverse is the inverse.
SFV_DEV bool inv3_regular(double a, double b, double c, double d, double e, double f, double g, double h, double i,
                          double F2, double *P) {
  const double c00 = e * i - f * h, c01 = f * g - d * i, c02 = d * h - e * g;
  const double det = a * c00 + b * c01 + c * c02;
  if (!(fabs(det) > 2e-12 * (F2 * sqrt(F2)))) return false;
  const double r = 1.0 / det;
  P[0] = c00 * r;
  P[1] = c01 * r;
  P[2] = c02 * r;
  P[3] = (c * h - b * i) * r;
  P[4] = (a * i - c * g) * r;
  P[5] = (b * g - a * h) * r;
  P[6] = (b * f - c * e) * r;
  P[7] = (c * d - a * f) * r;
  P[8] = (a * e - b * d) * r;
  return true;
}

// Pseudo-inverse V S^+ U^T of a column-major 3x3 matrix, singular values
// below 3e-12 * max(S) dropped (linear_solver_module.F90:683-715): the
// cofactor inverse when the determinant shows no value near the cut-off,
// otherwise an SVD by one-sided Jacobi rotations, sorted by decreasing value.
SFV_DEV void pinv3(double *A) {
  {
    double F2 = 0.0, P[9];
#pragma unroll
    for (int k = 0; k < 9; k++) F2 = F2 + A[k] * A[k];
    if (inv3_regular(A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8], F2, P)) {
#pragma unroll
      for (int k = 0; k < 9; k++) A[k] = P[k];
      return;
    }
  }
  double a[9], v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
  for (int i = 0; i < 9; i++) a[i] = A[i];
  for (int sweep = 0; sweep < 16; sweep++) {
    bool rotated = false;
#pragma unroll
    for (int pq = 0; pq < 3; pq++) {
      const int p = (pq == 2) ? 1 : 0;
      const int q = (pq == 0) ? 1 : 2;
      double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        al = al + a[i + 3 * p] * a[i + 3 * p];
        be = be + a[i + 3 * q] * a[i + 3 * q];
        ga = ga + a[i + 3 * p] * a[i + 3 * q];
      }
      if (ga == 0.0) continue;
      if (fabs(ga) <= 1e-15 * (sqrt(al) * sqrt(be))) continue;
      rotated = true;
      double zeta = (be - al) / (2.0 * ga);
      double t;
      if (fabs(zeta) > 1e100)
        t = 0.5 / zeta;
      else
        t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
      double c = 1.0 / sqrt(1.0 + t * t);
      double s = c * t;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        double ap = a[i + 3 * p], aq = a[i + 3 * q];
        a[i + 3 * p] = c * ap - s * aq;
        a[i + 3 * q] = s * ap + c * aq;
        double vp = v[i + 3 * p], vq = v[i + 3 * q];
        v[i + 3 * p] = c * vp - s * vq;
        v[i + 3 * q] = s * vp + c * vq;
      }
    }
    if (!rotated) break;
  }
  double sv[3], u[9];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    sv[k] = sqrt(a[3 * k] * a[3 * k] + a[3 * k + 1] * a[3 * k + 1] + a[3 * k + 2] * a[3 * k + 2]);
#pragma unroll
    for (int i = 0; i < 3; i++) u[i + 3 * k] = (sv[k] > 0.0) ? a[i + 3 * k] / sv[k] : 0.0;
  }
  // stable order by decreasing singular value (insertion sort of 3)
  int o0 = 0, o1 = 1, o2 = 2;
  if (sv[o0] < sv[o1]) { int t = o0; o0 = o1; o1 = t; }
  if (sv[o1] < sv[o2]) {
    int t = o1; o1 = o2; o2 = t;
    if (sv[o0] < sv[o1]) { t = o0; o0 = o1; o1 = t; }
  }
  const int ord[3] = {o0, o1, o2};
  double tol = 3.0 * 1e-12 * sv[o0];
  double sp[3], uu[9], vv[9];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int ok = ord[k];
    double s = (ok == 0) ? sv[0] : ((ok == 1) ? sv[1] : sv[2]);
    sp[k] = (s > tol) ? 1.0 / s : 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      uu[i + 3 * k] = (ok == 0) ? u[i] : ((ok == 1) ? u[i + 3] : u[i + 6]);
      vv[i + 3 * k] = (ok == 0) ? v[i] : ((ok == 1) ? v[i + 3] : v[i + 6]);
    }
  }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++) acc = acc + vv[i + 3 * k] * (sp[k] * uu[j + 3 * k]);
      A[i + 3 * j] = acc;
    }
}

// Pseudo-inverse of a symmetric 3x3 matrix (a00, a01, a02, a11, a12, a22),
// written column-major into P: the cofactor inverse when the determinant
// shows no eigenvalue near the cut-off; otherwise cyclic Jacobi rotations give
// M = Q L Q^T, the singular values are |L|, and eigenvalues with
// |l| <= 3e-12 * max |l| are dropped, the cut-off of
// linear_solver_module.F90:703-711.
SFV_DEV void pinv3_sym(const double *S6, double *P) {
  double a00 = S6[0], a01 = S6[1], a02 = S6[2], a11 = S6[3], a12 = S6[4], a22 = S6[5];
  {
    const double F2 = ((a00 * a00 + a11 * a11) + a22 * a22) + 2.0 * ((a01 * a01 + a02 * a02) + a12 * a12);
    if (inv3_regular(a00, a01, a02, a01, a11, a12, a02, a12, a22, F2, P)) return;
  }
  double q[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int sweep = 0; sweep < 16; sweep++) {
    bool rotated = false;
#pragma unroll
    for (int pq = 0; pq < 3; pq++) {
      // (p, q, r) = (0,1,2), (0,2,1), (1,2,0)
      double app, aqq, apq, arp, arq;
      if (pq == 0) { app = a00; aqq = a11; apq = a01; arp = a02; arq = a12; }
      else if (pq == 1) { app = a00; aqq = a22; apq = a02; arp = a01; arq = a12; }
      else { app = a11; aqq = a22; apq = a12; arp = a01; arq = a02; }
      if (apq == 0.0) continue;
      if (fabs(apq) <= 1e-15 * sqrt(fabs(app) * fabs(aqq))) continue;
      rotated = true;
      double theta = (aqq - app) / (2.0 * apq);
      double t;
      if (fabs(theta) > 1e100)
        t = 0.5 / theta;
      else
        t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
      double c = 1.0 / sqrt(1.0 + t * t);
      double s = t * c;
      double npp = app - t * apq, nqq = aqq + t * apq;
      double nrp = c * arp - s * arq, nrq = s * arp + c * arq;
      const int p = (pq == 2) ? 1 : 0, qq = (pq == 0) ? 1 : 2;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        double vp = q[i + 3 * p], vq = q[i + 3 * qq];
        q[i + 3 * p] = c * vp - s * vq;
        q[i + 3 * qq] = s * vp + c * vq;
      }
      if (pq == 0) { a00 = npp; a11 = nqq; a01 = 0.0; a02 = nrp; a12 = nrq; }
      else if (pq == 1) { a00 = npp; a22 = nqq; a02 = 0.0; a01 = nrp; a12 = nrq; }
      else { a11 = npp; a22 = nqq; a12 = 0.0; a01 = nrp; a02 = nrq; }
    }
    if (!rotated) break;
  }
  double l[3] = {a00, a11, a22};
  // order by decreasing |l| (stable)
  int o0 = 0, o1 = 1, o2 = 2;
  if (fabs(l[o0]) < fabs(l[o1])) { int t = o0; o0 = o1; o1 = t; }
  if (fabs(l[o1]) < fabs(l[o2])) {
    int t = o1; o1 = o2; o2 = t;
    if (fabs(l[o0]) < fabs(l[o1])) { t = o0; o0 = o1; o1 = t; }
  }
  const int ord[3] = {o0, o1, o2};
  const double tol = 3.0 * 1e-12 * fabs(l[o0]);
  double ip[3], qs[9];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int ok = ord[k];
    const double lk = (ok == 0) ? l[0] : ((ok == 1) ? l[1] : l[2]);
    ip[k] = (fabs(lk) > tol) ? 1.0 / lk : 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) qs[i + 3 * k] = (ok == 0) ? q[i] : ((ok == 1) ? q[i + 3] : q[i + 6]);
  }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++) acc = acc + qs[i + 3 * k] * (ip[k] * qs[j + 3 * k]);
      P[i + 3 * j] = acc;
    }
}

}  // namespace sfv
