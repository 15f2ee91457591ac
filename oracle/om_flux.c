/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see om_internal.h).
 *
 * Restatement of the explicit residual path of subfv:
 *   primitives          src/ns/ns_euler_primitives_module.F90:7-111,162-190
 *   Riemann fluxes      src/ns/ns_euler_rs_module.F90:9-286,555-679
 *   nodal solvers       src/ns/ns_euler_rs_module.F90:288-553,681-758
 *   reconstruction/BCs  src/ns/ns_euler_recon_module.F90:331-451
 *   node sweep+scatter  src/ns/ns_euler_module.F90:151-339 (+ :341-519)
 *   gradients/limiters  src/ns/ns_euler_recon_module.F90:8-329,453-490
 *   flux_limiting       src/ns/ns_euler_recon_module.F90:560-601
 *   stage update        src/ns/ns_euler_module.F90:521-609
 *   time loop           src/ns/ns_main.F90:83-107, compute_rhs_ns :130-179
 * State arrays keep the reference's Fortran layout: sol(5,C) is sol[5*c+k],
 * grad(3,5,C) is grad[15*c + 3*k + d], diag(5,5,C) is diag[25*c + i + 5*j].
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "om_dense.h"
#include "om_internal.h"

#define GAMMA 1.4
#define VC(m, i, d) ((m)->vcoord[3 * (i) + (d)])
#define EC(m, i, d) ((m)->ecoord[3 * (i) + (d)])
#define FC(m, i, d) ((m)->fcoord[3 * (i) + (d)])
#define SN(m, i, d) ((m)->sf_norm[3 * (i) + (d)])
#define SOL(s, c, k) ((s)[5 * ((c)-1) + (k)])

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* sol_potential_flow_2d / _3d (ns_euler_exact_sol_module.F90:105-170): the
 * primitive state of potential flow around a cylinder (2-D, r0 = 0.5 inside an
 * outer radius 5.5, through atan2 / cos / sin) or a sphere (3-D, r0 = 0.5) at
 * Mach 1e-4; integer powers as products (r^5 = ((r r)(r r)) r). */
void om_potential_flow(int three_d, const double *x, double *w) {
  const double rhoinf = 1.4, mach = 1e-4, pinf = 1.0, r0 = 0.5;
  const double vinf = mach * sqrt(GAMMA * pinf / rhoinf);
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

/* ------------------------------------------------------------- primitives */
static void primit_to_conserv(const double *w, double *u) {
  u[0] = w[0];
  u[1] = w[1] * w[0];
  u[2] = w[2] * w[0];
  u[3] = w[3] * w[0];
  u[4] = w[4] / (GAMMA - 1) + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
}
static void conserv_to_primit(const double *u, double *w) {
  w[0] = u[0];
  w[1] = u[1] / u[0];
  w[2] = u[2] / u[0];
  w[3] = u[3] / u[0];
  w[4] = (GAMMA - 1) * (u[4] - 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]));
}
static double sound_speed_w(const double *w) { return sqrt(GAMMA * w[4] / w[0]); }

/* max_omega_for_positivity, ns_euler_primitives_module.F90:92-111 */
double om_max_omega_for_positivity(const double *U, const double *dU) {
  double omega = 1.0, us[5], ws[5];
  for (int k = 0; k < 5; k++) us[k] = U[k] + omega * dU[k];
  conserv_to_primit(us, ws);
  int pos = (ws[0] > 1e-17 && ws[4] > 1e-17);
  while (!pos && omega > 1e-6) {
    omega = omega * 0.9;
    for (int k = 0; k < 5; k++) us[k] = U[k] + omega * dU[k];
    conserv_to_primit(us, ws);
    pos = (ws[0] > 1e-17 && ws[4] > 1e-17);
  }
  return omega;
}

/* is_wall, ns_euler_recon_module.F90:411-425 */
static int is_wall(const om_config *cfg, int re) {
  if (re > 0) return 0;
  if (re == 0) return 1;
  return cfg->bc_euler_id[-re] == OM_BC_WALL;
}

/* wall_normal, ns_euler_primitives_module.F90:162-190 */
static void wall_normal(const om_mesh *m, const om_config *cfg, int v, double *Bp) {
  Bp[0] = Bp[1] = Bp[2] = 0.0;
  for (int j = 1; j <= m->v_nsfn[v]; j++) {
    int sf = m->v_sf[v][j];
    int re = m->f_right[m->sf_face[sf]];
    if (re <= 0) {
      const double *n = &SN(m, sf, 0);
      if (!(cfg->boundary_2d && fabs(n[2]) > 1e-12)) {
        if (re == 0 || cfg->bc_euler_id[-re] == OM_BC_WALL)
          for (int d = 0; d < 3; d++) Bp[d] = Bp[d] + m->sf_area[sf] * n[d];
      }
    }
  }
}

/* ------------------------------------------------------------ BC ghosts */
static void wall_reflect(const double *n, const double *sl, double *sr) {
  for (int k = 0; k < 5; k++) sr[k] = sl[k];
  double vn = sr[1] * n[0] + sr[2] * n[1] + sr[3] * n[2];
  for (int d = 0; d < 3; d++) sr[1 + d] = sr[1 + d] - 2.0 * vn * n[d];
}

/* compute_right_state, ns_euler_recon_module.F90:331-409 (potential-flow
 * boundaries are not restated). Returns nonzero for an unsupported BC. */
static int compute_right_state(const om_mesh *m, const om_config *cfg, int sf, int re,
                               const double *sol_l, double *sol_r) {
  const double *n = &SN(m, sf, 0);
  if (re > 0) return OM_ERR_BAD_INPUT;
  if (re == 0) {
    wall_reflect(n, sol_l, sol_r);
    return OM_OK;
  }
  int id = cfg->bc_euler_id[-re];
  switch (id) {
    case OM_BC_OUTFLOWSUPERSONIC:
      for (int k = 0; k < 5; k++) sol_r[k] = sol_l[k];
      return OM_OK;
    case OM_BC_FREESTREAM:
      primit_to_conserv(cfg->bc_val[-re], sol_r);
      return OM_OK;
    case OM_BC_INFLOW_POND:
      primit_to_conserv(cfg->bc_val[-re], sol_r);
      if (sol_r[1] * n[0] + sol_r[2] * n[1] + sol_r[3] * n[2] > 0.0)
        for (int k = 0; k < 5; k++) sol_r[k] = sol_l[k];
      return OM_OK;
    case OM_BC_WALL:
      wall_reflect(n, sol_l, sol_r);
      return OM_OK;
    case OM_BC_ADHERENCE_WALL: {
      for (int k = 0; k < 5; k++) sol_r[k] = sol_l[k];
      for (int d = 0; d < 3; d++) sol_r[1 + d] = 2.0 * sol_l[0] * cfg->bc_val[-re][1 + d] - sol_l[1 + d];
      double a[3], b[3];
      for (int d = 0; d < 3; d++) {
        a[d] = sol_l[1 + d] / sol_l[0];
        b[d] = sol_r[1 + d] / sol_r[0];
      }
      double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      sol_r[4] = sol_l[4] - 0.5 * sol_l[0] * (na * na) + 0.5 * sol_l[0] * (nb * nb);
      return OM_OK;
    }
    case OM_BC_INOUT_DOUBLE_MACH: {
      const double *fc = &FC(m, m->sf_face[sf], 0);
      static const double wa[5] = {8.0, 7.145, -4.125, 0.0, 116.5};
      static const double wb[5] = {1.4, 0.0, 0.0, 0.0, 1.0};
      if (fc[1] > 1.732 * (fc[0] - 0.1667 - 10.0 * cfg->t))
        primit_to_conserv(wa, sol_r);
      else
        primit_to_conserv(wb, sol_r);
      return OM_OK;
    }
    case OM_BC_DOUBLE_MACH_BOTTOM: {
      const double *fc = &FC(m, m->sf_face[sf], 0);
      static const double wa[5] = {8.0, 7.145, -4.125, 0.0, 116.5};
      if (fc[0] < 0.1667)
        primit_to_conserv(wa, sol_r);
      else
        wall_reflect(n, sol_l, sol_r);
      return OM_OK;
    }
    case OM_BC_POTENTIAL_FLOW_2D:
    case OM_BC_POTENTIAL_FLOW_3D: {
      double w[5];
      om_potential_flow(id == OM_BC_POTENTIAL_FLOW_3D, &FC(m, m->sf_face[sf], 0), w);
      primit_to_conserv(w, sol_r);
      return OM_OK;
    }
    default:
      return OM_ERR_BAD_INPUT;
  }
}

/* reconstruct_lr_w, ns_euler_recon_module.F90:427-451 */
static int reconstruct_lr_w(const om_mesh *m, const om_config *cfg, const double *sol,
                            const double *grad, int v, int sf, int le, int re, int second_order,
                            double *wl, double *wr) {
  conserv_to_primit(&SOL(sol, le, 0), wl);
  if (second_order) {
    const double *g = &grad[15 * (le - 1)];
    double dx[3];
    for (int d = 0; d < 3; d++) dx[d] = VC(m, v, d) - EC(m, le, d);
    for (int k = 0; k < 5; k++) wl[k] = wl[k] + (g[3 * k] * dx[0] + g[3 * k + 1] * dx[1] + g[3 * k + 2] * dx[2]);
  }
  if (re > 0) {
    conserv_to_primit(&SOL(sol, re, 0), wr);
    if (second_order) {
      const double *g = &grad[15 * (re - 1)];
      double dx[3];
      for (int d = 0; d < 3; d++) dx[d] = VC(m, v, d) - EC(m, re, d);
      for (int k = 0; k < 5; k++)
        wr[k] = wr[k] + (g[3 * k] * dx[0] + g[3 * k + 1] * dx[1] + g[3 * k + 2] * dx[2]);
    }
  } else {
    double ul[5], ug[5];
    primit_to_conserv(wl, ul);
    int st = compute_right_state(m, cfg, sf, re, ul, ug);
    if (st) return st;
    conserv_to_primit(ug, wr);
  }
  return OM_OK;
}

/* --------------------------------------------------------- subface fluxes */
typedef struct {
  double rho, vn, p, e, a, sol[5], f[5];
} side_t;

static void side_setup(const double *w, const double *n, side_t *s, int need_a) {
  s->rho = w[0];
  s->vn = w[1] * n[0] + w[2] * n[1] + w[3] * n[2];
  s->p = w[4];
  primit_to_conserv(w, s->sol);
  s->e = s->sol[4] / s->rho;
  s->a = need_a ? sound_speed_w(w) : 0.0;
  s->f[0] = s->vn * s->sol[0];
  for (int d = 0; d < 3; d++) s->f[1 + d] = s->vn * s->sol[1 + d] + s->p * n[d];
  s->f[4] = (s->sol[4] + s->p) * s->vn;
}

/* three_wave, ns_euler_rs_module.F90:9-75 */
static void three_wave(const double *wl, const double *wr, const double *n, double *flux, double *sl,
                       double *sr) {
  side_t L, R;
  side_setup(wl, n, &L, 1);
  side_setup(wr, n, &R, 1);
  double lambda_l = dmax(dmax(L.a * L.rho, sqrt(L.rho * dmax(0.0, R.p - L.p))), -L.rho * (R.vn - L.vn));
  double lambda_r = dmax(dmax(R.a * R.rho, sqrt(R.rho * dmax(0.0, L.p - R.p))), -R.rho * (R.vn - L.vn));
  double v_bar = (lambda_l * L.vn + lambda_r * R.vn - (R.p - L.p)) / (lambda_r + lambda_l);
  double v_et = v_bar;
  double rhol_et = 1.0 / (1.0 / L.rho + (v_et - L.vn) / lambda_l);
  double pl_bar = L.p - lambda_l * (v_et - L.vn);
  double ql[5], qr[5];
  ql[0] = rhol_et;
  for (int d = 0; d < 3; d++) ql[1 + d] = rhol_et * (wl[1 + d] + (v_et - L.vn) * n[d]);
  ql[4] = rhol_et * (L.e + (L.p * L.vn - pl_bar * v_et) / lambda_l);
  double rhor_et = 1.0 / (1.0 / R.rho + (R.vn - v_et) / lambda_r);
  double pr_bar = R.p + lambda_r * (v_et - R.vn);
  qr[0] = rhor_et;
  for (int d = 0; d < 3; d++) qr[1 + d] = rhor_et * (wr[1 + d] + (v_et - R.vn) * n[d]);
  qr[4] = rhor_et * (R.e + (pr_bar * v_et - R.p * R.vn) / lambda_r);
  *sl = L.vn - lambda_l / L.rho;
  *sr = R.vn + lambda_r / R.rho;
  for (int k = 0; k < 5; k++) {
    flux[k] = 0.5 * (L.f[k] + R.f[k]) -
              0.5 * (fabs(*sl) * (ql[k] - L.sol[k]) + fabs(v_et) * (qr[k] - ql[k]) + fabs(*sr) * (R.sol[k] - qr[k]));
    flux[5 + k] = -flux[k];
  }
}

/* modified_three_wave, ns_euler_rs_module.F90:77-152 */
static void modified_three_wave(const double *wl, const double *wr, const double *n, double *flux,
                                double *sl, double *sr) {
  side_t L, R;
  side_setup(wl, n, &L, 1);
  side_setup(wr, n, &R, 1);
  double lambda_l = dmax(dmax(L.a * L.rho, sqrt(L.rho * dmax(0.0, R.p - L.p))), -L.rho * (R.vn - L.vn));
  double lambda_r = dmax(dmax(R.a * R.rho, sqrt(R.rho * dmax(0.0, L.p - R.p))), -R.rho * (R.vn - L.vn));
  double v_bar = (lambda_l * L.vn + lambda_r * R.vn - (R.p - L.p)) / (lambda_r + lambda_l);
  double v_et = v_bar;
  double vt_l[3], vt_r[3], vt_et[3], dvt[3];
  for (int d = 0; d < 3; d++) {
    vt_l[d] = wl[1 + d] - L.vn * n[d];
    vt_r[d] = wr[1 + d] - R.vn * n[d];
  }
  for (int d = 0; d < 3; d++) {
    vt_et[d] = (lambda_l * vt_l[d] + lambda_r * vt_r[d]) / (lambda_l + lambda_r);
    dvt[d] = vt_r[d] - vt_l[d];
  }
  double tang_coeff = -lambda_l * lambda_r / (lambda_l + lambda_r);
  double tang_energy = tang_coeff * dot3(dvt, vt_et);
  double rhol_et = 1.0 / (1.0 / L.rho + (v_et - L.vn) / lambda_l);
  double pl_bar = L.p - lambda_l * (v_et - L.vn);
  double ql[5], qr[5];
  ql[0] = rhol_et;
  for (int d = 0; d < 3; d++) ql[1 + d] = rhol_et * (v_et * n[d] + vt_et[d]);
  ql[4] = rhol_et * (L.e + (L.p * L.vn - pl_bar * v_et - tang_energy) / lambda_l);
  double rhor_et = 1.0 / (1.0 / R.rho + (R.vn - v_et) / lambda_r);
  double pr_bar = R.p + lambda_r * (v_et - R.vn);
  qr[0] = rhor_et;
  for (int d = 0; d < 3; d++) qr[1 + d] = rhor_et * (v_et * n[d] + vt_et[d]);
  qr[4] = rhor_et * (R.e + (pr_bar * v_et - R.p * R.vn + tang_energy) / lambda_r);
  *sl = L.vn - lambda_l / L.rho;
  *sr = R.vn + lambda_r / R.rho;
  for (int k = 0; k < 5; k++) {
    flux[k] = 0.5 * (L.f[k] + R.f[k]) -
              0.5 * (fabs(*sl) * (ql[k] - L.sol[k]) + fabs(v_et) * (qr[k] - ql[k]) + fabs(*sr) * (R.sol[k] - qr[k]));
    flux[5 + k] = -flux[k];
  }
}

/* two_wave, ns_euler_rs_module.F90:154-198 */
static void two_wave(const double *wl, const double *wr, const double *n, double *flux, double *sl,
                     double *sr) {
  side_t L, R;
  side_setup(wl, n, &L, 1);
  side_setup(wr, n, &R, 1);
  double lambda_l = dmax(dmax(L.a * L.rho, sqrt(L.rho * dmax(0.0, R.p - L.p))), -L.rho * (R.vn - L.vn));
  double lambda_r = dmax(dmax(R.a * R.rho, sqrt(R.rho * dmax(0.0, L.p - R.p))), -R.rho * (R.vn - L.vn));
  *sl = L.vn - lambda_l / L.rho;
  *sr = R.vn + lambda_r / R.rho;
  for (int k = 0; k < 5; k++) {
    double et = (*sr * R.sol[k] - *sl * L.sol[k] - (R.f[k] - L.f[k])) / (*sr - *sl);
    flux[k] = 0.5 * (L.f[k] + R.f[k]) - 0.5 * (fabs(*sl) * (et - L.sol[k]) + fabs(*sr) * (R.sol[k] - et));
    flux[5 + k] = -flux[k];
  }
}

/* multi_point, ns_euler_rs_module.F90:200-286 */
static int multi_point(const double *wl, const double *wr, const double *n, double *flux, double vn_nodal,
                       double lambda_l, double lambda_r, double *sl, double *sr) {
  side_t L, R;
  side_setup(wl, n, &L, 0);
  side_setup(wr, n, &R, 0);
  double v_et = vn_nodal;
  double rhol_et = 1.0 / (1.0 / L.rho + (v_et - L.vn) / lambda_l);
  double pl_bar = L.p - lambda_l * (v_et - L.vn);
  double ql[5], qr[5];
  ql[0] = rhol_et;
  for (int d = 0; d < 3; d++) ql[1 + d] = rhol_et * (wl[1 + d] + (v_et - L.vn) * n[d]);
  ql[4] = rhol_et * (L.e + (L.p * L.vn - pl_bar * v_et) / lambda_l);
  double rhor_et = 1.0 / (1.0 / R.rho + (R.vn - v_et) / lambda_r);
  double pr_bar = R.p + lambda_r * (v_et - R.vn);
  qr[0] = rhor_et;
  for (int d = 0; d < 3; d++) qr[1 + d] = rhor_et * (wr[1 + d] + (v_et - R.vn) * n[d]);
  qr[4] = rhor_et * (R.e + (pr_bar * v_et - R.p * R.vn) / lambda_r);
  if (rhol_et < 0.0 || rhor_et < 0.0) return OM_ERR_NEG_STAR_DENSITY;
  *sl = L.vn - lambda_l / L.rho;
  *sr = R.vn + lambda_r / R.rho;
  double e[5] = {0.0, n[0], n[1], n[2], v_et};
  for (int k = 0; k < 5; k++) {
    double common = 0.5 * (L.f[k] + R.f[k]) -
                    0.5 * (fabs(*sl) * (ql[k] - L.sol[k]) + fabs(v_et) * (qr[k] - ql[k]) + fabs(*sr) * (R.sol[k] - qr[k]));
    flux[k] = common - 0.5 * (pr_bar - pl_bar) * e[k];
    flux[5 + k] = -(common + 0.5 * (pr_bar - pl_bar) * e[k]);
  }
  return OM_OK;
}

/* multi_point_pressure (ph=0) and multi_point_pressure_ph (ph=1),
 * ns_euler_rs_module.F90:555-679 */
static void multi_point_pressure(const double *wl, const double *wr, const double *n, double *flux,
                                 double p_nodal, double lambda_l, double lambda_r, double *sl,
                                 double *sr, int ph) {
  side_t L, R;
  side_setup(wl, n, &L, 0);
  side_setup(wr, n, &R, 0);
  double vt_l[3], vt_r[3];
  for (int d = 0; d < 3; d++) {
    vt_l[d] = wl[1 + d] - L.vn * n[d];
    vt_r[d] = wr[1 + d] - R.vn * n[d];
  }
  double vn_bar = (lambda_l * L.vn + lambda_r * R.vn - (R.p - L.p)) / (lambda_r + lambda_l);
  double vn_l_et = L.vn - (p_nodal - L.p) / lambda_l;
  double rhol_et = 1.0 / (1.0 / L.rho + (vn_bar - L.vn) / lambda_l);
  double ql[5], qr[5];
  ql[0] = rhol_et;
  for (int d = 0; d < 3; d++) ql[1 + d] = rhol_et * (vt_l[d] + vn_l_et * n[d]);
  double vn_r_et = R.vn + (p_nodal - R.p) / lambda_r;
  double rhor_et = 1.0 / (1.0 / R.rho + (R.vn - vn_bar) / lambda_r);
  qr[0] = rhor_et;
  for (int d = 0; d < 3; d++) qr[1 + d] = rhor_et * (vt_r[d] + vn_r_et * n[d]);
  if (ph) {
    ql[4] = rhol_et * (L.e + (L.p * L.vn - p_nodal * vn_l_et) / lambda_l);
    qr[4] = rhor_et * (R.e + (p_nodal * vn_r_et - R.p * R.vn) / lambda_r);
  } else {
    ql[4] = rhol_et * (L.e + (L.p * L.vn - p_nodal * vn_bar) / lambda_l);
    qr[4] = rhor_et * (R.e + (p_nodal * vn_bar - R.p * R.vn) / lambda_r);
  }
  *sl = L.vn - lambda_l / L.rho;
  *sr = R.vn + lambda_r / R.rho;
  for (int k = 0; k < 5; k++) {
    flux[k] = 0.5 * (L.f[k] + R.f[k]) -
              0.5 * (fabs(*sl) * (ql[k] - L.sol[k]) + fabs(vn_bar) * (qr[k] - ql[k]) + fabs(*sr) * (R.sol[k] - qr[k]));
    flux[5 + k] = -flux[k];
  }
}

/* ---------------------------------------------------------- nodal solvers */
/* compute_lambdas_and_solve_nodal_velocity (iso=0, ns_euler_rs_module.F90:288-422)
 * and its _iso variant without area weights (iso=1, :424-553).
 * W holds the left/right primitive states of the vertex's subfaces,
 * W[10*(j-1) + 5*side + k]. */
static void solve_nodal_velocity(const om_mesh *m, const om_config *cfg, int v, const double *W,
                                 double *lam, double *v_bars, double *v_node, int iso) {
  double Bp[3];
  wall_normal(m, cfg, v, Bp);
  v_node[0] = v_node[1] = v_node[2] = 0.0;
  int nsfn = m->v_nsfn[v];
  for (int iter = 1; iter <= 4; iter++) {
    double mat[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double Rp[3] = {0, 0, 0};
    for (int j = 1; j <= nsfn; j++) {
      int sf = m->v_sf[v][j];
      int re = m->f_right[m->sf_face[sf]];
      const double *n = &SN(m, sf, 0);
      const double *wl = &W[10 * (j - 1)], *wr = &W[10 * (j - 1) + 5];
      double area = m->sf_area[sf];
      double rhol = wl[0], vn_l = wl[1] * n[0] + wl[2] * n[1] + wl[3] * n[2], pl = wl[4];
      double al = sound_speed_w(wl);
      double rhor = wr[0], vn_r = wr[1] * n[0] + wr[2] * n[1] + wr[3] * n[2], pr = wr[4];
      double ar = sound_speed_w(wr);
      double lambda_l = lam[2 * (j - 1)], lambda_r = lam[2 * (j - 1) + 1];
      if (iter == 1) {
        lambda_l = dmax(dmax(dmax(lambda_l, al * rhol), sqrt(rhol * dmax(0.0, pr - pl))), -rhol * (vn_r - vn_l));
        lambda_r = dmax(dmax(dmax(lambda_r, ar * rhor), sqrt(rhor * dmax(0.0, pl - pr))), -rhor * (vn_r - vn_l));
        v_bars[j - 1] = (lambda_l * vn_l + lambda_r * vn_r - (pr - pl)) / (lambda_r + lambda_l);
        lambda_l = dmax(lambda_l, al * rhol * (1.0 + 1.5 * dmax(0.0, -(v_bars[j - 1] - vn_l) / al)));
        lambda_r = dmax(lambda_r, ar * rhor * (1.0 + 1.5 * dmax(0.0, (v_bars[j - 1] - vn_r) / ar)));
      } else {
        if (is_wall(cfg, re) && cfg->bc_style == 1) {
          lambda_l = dmax(lambda_l, al * rhol * (1.0 + 1.5 * dmax(0.0, -(v_bars[j - 1] - vn_l) / al)));
          lambda_r = dmax(lambda_r, ar * rhor * (1.0 + 1.5 * dmax(0.0, (v_bars[j - 1] - vn_r) / ar)));
        } else {
          double v_et = v_node[0] * n[0] + v_node[1] * n[1] + v_node[2] * n[2];
          lambda_l = dmax(lambda_l, al * rhol * (1.0 + 1.5 * dmax(0.0, -(v_et - vn_l) / al)));
          lambda_r = dmax(lambda_r, ar * rhor * (1.0 + 1.5 * dmax(0.0, (v_et - vn_r) / ar)));
        }
      }
      lam[2 * (j - 1)] = lambda_l;
      lam[2 * (j - 1) + 1] = lambda_r;
      v_bars[j - 1] = (lambda_l * vn_l + lambda_r * vn_r - (pr - pl)) / (lambda_r + lambda_l);
      /* (lambda_r + lambda_l)*area*n(a)*n(b) for the area-weighted form,
       * (lambda_r + lambda_l)*n(a)*n(b) for the iso form */
      double cm = iso ? (lambda_r + lambda_l) : (lambda_r + lambda_l) * area;
      if (is_wall(cfg, re)) {
        if (cfg->bc_style == 0 || cfg->bc_style == 2) {
          for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) mat[a + 3 * b] = mat[a + 3 * b] + cm * (n[a] * n[b]);
          if (cfg->bc_style == 0) {
            double cr = iso ? (lambda_r + lambda_l) * v_bars[j - 1] : area * (lambda_r + lambda_l) * v_bars[j - 1];
            for (int d = 0; d < 3; d++) Rp[d] = Rp[d] + cr * n[d];
          } else {
            int add = cfg->boundary_2d ? (fabs(n[2]) < 1e-8) : 1;
            if (add) {
              double cr = iso ? pl : area * pl;
              for (int d = 0; d < 3; d++) Rp[d] = Rp[d] + cr * n[d];
            }
          }
        }
      } else {
        for (int b = 0; b < 3; b++)
          for (int a = 0; a < 3; a++) mat[a + 3 * b] = mat[a + 3 * b] + cm * (n[a] * n[b]);
        double cr = iso ? (lambda_r + lambda_l) * v_bars[j - 1] : area * (lambda_r + lambda_l) * v_bars[j - 1];
        for (int d = 0; d < 3; d++) Rp[d] = Rp[d] + cr * n[d];
      }
    }
    if (cfg->bc_style == 2) {
      double mb = dmax(dmax(fabs(Bp[0]), fabs(Bp[1])), fabs(Bp[2]));
      if (mb > 1e-8) {
        om_pinv3_sym(mat, mat);
        double mR[3], mB[3];
        for (int a = 0; a < 3; a++) {
          mR[a] = mat[a] * Rp[0] + mat[a + 3] * Rp[1] + mat[a + 6] * Rp[2];
          mB[a] = mat[a] * Bp[0] + mat[a + 3] * Bp[1] + mat[a + 6] * Bp[2];
        }
        double p_bound = dot3(mR, Bp) / dot3(mB, Bp);
        double rr[3];
        for (int d = 0; d < 3; d++) rr[d] = Rp[d] - p_bound * Bp[d];
        for (int a = 0; a < 3; a++) v_node[a] = mat[a] * rr[0] + mat[a + 3] * rr[1] + mat[a + 6] * rr[2];
        double sc = dot3(v_node, Bp) / dot3(Bp, Bp);
        for (int d = 0; d < 3; d++) v_node[d] = v_node[d] - sc * Bp[d];
      } else {
        om_lu_solve(3, mat, Rp, v_node);
      }
    } else {
      om_pinv3_sym(mat, mat);
      double vn[3];
      for (int a = 0; a < 3; a++) vn[a] = mat[a] * Rp[0] + mat[a + 3] * Rp[1] + mat[a + 6] * Rp[2];
      for (int d = 0; d < 3; d++) v_node[d] = vn[d];
    }
  }
}

/* compute_lambdas_and_solve_nodal_pressure, ns_euler_rs_module.F90:681-758 */
static double solve_nodal_pressure(const om_mesh *m, const om_config *cfg, int v, const double *W,
                                   double *lam, double *p_bars) {
  double p_node = 0.0;
  int nsfn = m->v_nsfn[v];
  for (int iter = 1; iter <= 10; iter++) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 1; j <= nsfn; j++) {
      int sf = m->v_sf[v][j];
      int re = m->f_right[m->sf_face[sf]];
      const double *n = &SN(m, sf, 0);
      const double *wl = &W[10 * (j - 1)], *wr = &W[10 * (j - 1) + 5];
      double area = m->sf_area[sf];
      double rhol = wl[0], vn_l = wl[1] * n[0] + wl[2] * n[1] + wl[3] * n[2], pl = wl[4];
      double al = sound_speed_w(wl);
      double rhor = wr[0], vn_r = wr[1] * n[0] + wr[2] * n[1] + wr[3] * n[2], pr = wr[4];
      double ar = sound_speed_w(wr);
      double lambda_l = lam[2 * (j - 1)], lambda_r = lam[2 * (j - 1) + 1];
      double *pb = &p_bars[j - 1];
      if (iter == 1) {
        lambda_l = dmax(dmax(dmax(lambda_l, al * rhol), sqrt(rhol * dmax(0.0, pr - pl))), -rhol * (vn_r - vn_l));
        lambda_r = dmax(dmax(dmax(lambda_r, ar * rhor), sqrt(rhor * dmax(0.0, pl - pr))), -rhor * (vn_r - vn_l));
        *pb = (lambda_r * pl + lambda_l * pr - lambda_r * lambda_l * (vn_r - vn_l)) / (lambda_r + lambda_l);
        lambda_l = dmax(lambda_l, al * rhol + 1.5 * sqrt(dmax(0.0, (*pb - pl) * rhol)));
        lambda_r = dmax(lambda_r, ar * rhor + 1.5 * sqrt(dmax(0.0, (*pb - pr) * rhor)));
      } else {
        if (cfg->bc_style == 1 && is_wall(cfg, re)) {
          lambda_l = dmax(lambda_l, al * rhol + 1.5 * sqrt(dmax(0.0, (*pb - pl) * rhol)));
          lambda_r = dmax(lambda_r, ar * rhor + 1.5 * sqrt(dmax(0.0, (*pb - pr) * rhor)));
        } else {
          lambda_l = dmax(lambda_l, al * rhol + 1.5 * sqrt(dmax(0.0, (p_node - pl) * rhol)));
          lambda_r = dmax(lambda_r, ar * rhor + 1.5 * sqrt(dmax(0.0, (p_node - pr) * rhor)));
        }
      }
      *pb = (lambda_r * pl + lambda_l * pr - lambda_r * lambda_l * (vn_r - vn_l)) / (lambda_r + lambda_l);
      if (cfg->bc_style == 1 && is_wall(cfg, re)) {
        s1 = s1 + (1.0 / lambda_r + 1.0 / lambda_l) * area * (lambda_r * pl + lambda_r * lambda_l * vn_l) /
                      (lambda_r + lambda_l);
        s2 = s2 + (1.0 / lambda_l) * area;
      } else {
        s1 = s1 + (1.0 / lambda_r + 1.0 / lambda_l) * area * *pb;
        s2 = s2 + (1.0 / lambda_r + 1.0 / lambda_l) * area;
      }
      lam[2 * (j - 1)] = lambda_l;
      lam[2 * (j - 1) + 1] = lambda_r;
    }
    p_node = s1 / s2;
  }
  return p_node;
}

/* --------------------------------------------------- the node-centred sweep */
typedef struct {
  double *W, *lam, *vb, *flux, *sflux, *slam, *warea;
} om_scratch;

static void scratch_alloc(om_scratch *s, int nmax, int nsen_max) {
  s->W = (double *)malloc(sizeof(double) * 10 * nmax);
  s->lam = (double *)malloc(sizeof(double) * 2 * nmax);
  s->vb = (double *)malloc(sizeof(double) * nmax);
  s->flux = (double *)malloc(sizeof(double) * 10 * nmax);
  s->warea = (double *)malloc(sizeof(double) * nmax);
  s->sflux = (double *)malloc(sizeof(double) * 5 * nsen_max);
  s->slam = (double *)malloc(sizeof(double) * nsen_max);
}
static void scratch_free(om_scratch *s) {
  free(s->W);
  free(s->lam);
  free(s->vb);
  free(s->flux);
  free(s->warea);
  free(s->sflux);
  free(s->slam);
}

/* compute_rhs_around_vert, ns_euler_module.F90:341-519 */
static int rhs_around_vert(const om_mesh *m, const om_config *cfg, const double *sol, const double *grad,
                           int second_order, int v, double *vp, om_scratch *s) {
  int nsfn = m->v_nsfn[v], nsen = m->v_nsen[v];
  double v_vert[3] = {0, 0, 0};
  double p_nodal = 0.0;
  for (int j = 0; j < nsen; j++) {
    s->slam[j] = 0.0;
    for (int k = 0; k < 5; k++) s->sflux[5 * j + k] = 0.0;
  }
  double amin = 0.0;
  for (int j = 1; j <= nsfn; j++) {
    int sf = m->v_sf[v][j];
    int f = m->sf_face[sf];
    int st = reconstruct_lr_w(m, cfg, sol, grad, v, sf, m->f_left[f], m->f_right[f], second_order,
                              &s->W[10 * (j - 1)], &s->W[10 * (j - 1) + 5]);
    if (st) return st;
    s->warea[j - 1] = m->sf_area[sf];
    amin = (j == 1) ? s->warea[0] : dmin(amin, s->warea[j - 1]);
    s->lam[2 * (j - 1)] = s->lam[2 * (j - 1) + 1] = 0.0;
    s->vb[j - 1] = 0.0;
  }
  for (int j = 0; j < nsfn; j++) s->warea[j] = amin / s->warea[j];

  int excl = m->vbound[v] && cfg->exclude_bound_vert;
  switch (cfg->scheme_id) {
    case OM_SCHEME_MULTI_POINT:
    case OM_SCHEME_MULTI_POINT_ISO:
      if (!excl)
        solve_nodal_velocity(m, cfg, v, s->W, s->lam, s->vb, v_vert, cfg->scheme_id == OM_SCHEME_MULTI_POINT_ISO);
      if (vp)
        for (int d = 0; d < 3; d++) vp[3 * (v - 1) + d] = v_vert[d];
      break;
    case OM_SCHEME_MULTI_POINT_PRESSURE:
    case OM_SCHEME_MULTI_POINT_PRESSURE_PH:
      p_nodal = solve_nodal_pressure(m, cfg, v, s->W, s->lam, s->vb);
      break;
    default:
      break;
  }

  for (int j = 1; j <= nsfn; j++) {
    int sf = m->v_sf[v][j];
    int f = m->sf_face[sf];
    int re = m->f_right[f];
    const double *n = &SN(m, sf, 0);
    const double *wl = &s->W[10 * (j - 1)], *wr = &s->W[10 * (j - 1) + 5];
    double *fl = &s->flux[10 * (j - 1)];
    double sl = 0.0, sr = 0.0;
    int lse = m->sf_lse[sf];
    int lse_loc = m->se_loc[lse];
    int rse = m->sf_rse[sf];
    int rse_loc = (rse > 0) ? m->se_loc[rse] : 0;
    switch (cfg->scheme_id) {
      case OM_SCHEME_MULTI_POINT:
      case OM_SCHEME_MULTI_POINT_ISO: {
        if (excl) {
          two_wave(wl, wr, n, fl, &sl, &sr);
        } else {
          double vn = is_wall(cfg, re) ? 0.0 : (v_vert[0] * n[0] + v_vert[1] * n[1] + v_vert[2] * n[2]);
          int st = multi_point(wl, wr, n, fl, vn, s->lam[2 * (j - 1)], s->lam[2 * (j - 1) + 1], &sl, &sr);
          if (st) return st;
        }
        if (cfg->scheme_id == OM_SCHEME_MULTI_POINT_ISO) {
          double f3[10];
          three_wave(wl, wr, n, f3, &sl, &sr);
          double wa = s->warea[j - 1];
          for (int k = 0; k < 10; k++) fl[k] = wa * fl[k] + (1.0 - wa) * f3[k];
        }
        break;
      }
      case OM_SCHEME_THREE_WAVE: three_wave(wl, wr, n, fl, &sl, &sr); break;
      case OM_SCHEME_TWO_WAVE: two_wave(wl, wr, n, fl, &sl, &sr); break;
      case OM_SCHEME_MODIFIED_THREE_WAVE: modified_three_wave(wl, wr, n, fl, &sl, &sr); break;
      case OM_SCHEME_MULTI_POINT_PRESSURE:
        multi_point_pressure(wl, wr, n, fl, p_nodal, s->lam[2 * (j - 1)], s->lam[2 * (j - 1) + 1], &sl, &sr, 0);
        break;
      case OM_SCHEME_MULTI_POINT_PRESSURE_PH:
        multi_point_pressure(wl, wr, n, fl, p_nodal, s->lam[2 * (j - 1)], s->lam[2 * (j - 1) + 1], &sl, &sr, 1);
        break;
      default:
        return OM_ERR_BAD_INPUT;
    }
    double area = m->sf_area[sf];
    if (m->se_vert[lse] == v) {
      s->slam[lse_loc - 1] = s->slam[lse_loc - 1] + area * dmax(0.0, -sl);
      for (int k = 0; k < 5; k++) s->sflux[5 * (lse_loc - 1) + k] = s->sflux[5 * (lse_loc - 1) + k] + area * fl[k];
      if (rse > 0 && m->se_vert[rse] == v) {
        s->slam[rse_loc - 1] = s->slam[rse_loc - 1] + area * dmax(0.0, sr);
        for (int k = 0; k < 5; k++)
          s->sflux[5 * (rse_loc - 1) + k] = s->sflux[5 * (rse_loc - 1) + k] + area * fl[5 + k];
      }
    }
  }
  return OM_OK;
}

static int max_nsen(const om_mesh *m) {
  int mx = 0;
  for (int i = 1; i <= m->nv; i++)
    if (m->v_nsen[i] > mx) mx = m->v_nsen[i];
  return mx;
}

/* compute_euler_flux, ns_euler_module.F90:151-339 (non-ZB schemes). */
int om_compute_euler_flux(const om_mesh *m, const om_config *cfg, const double *sol, double *rhs,
                          const double *grad, double *sum_lambda, int second_order, double *vp) {
  for (long i = 0; i < 5L * m->ne; i++) rhs[i] = 0.0;
  for (int i = 0; i < m->ne; i++) sum_lambda[i] = 1e-12;
  om_scratch s;
  int nsenm = max_nsen(m);
  scratch_alloc(&s, m->nmax_sfn > 0 ? m->nmax_sfn : 1, nsenm > 0 ? nsenm : 1);
  int status = OM_OK;
  for (int v = 1; v <= m->nv && status == OM_OK; v++) {
    status = rhs_around_vert(m, cfg, sol, grad, second_order, v, vp, &s);
    if (status) break;
    for (int j = 1; j <= m->v_nsen[v]; j++) {
      int se = m->v_se[v][j];
      int e = m->se_elem[se];
      for (int k = 0; k < 5; k++) SOL(rhs, e, k) = SOL(rhs, e, k) - s.sflux[5 * (j - 1) + k];
      sum_lambda[e - 1] = sum_lambda[e - 1] + s.slam[j - 1];
    }
  }
  scratch_free(&s);
  return status;
}

/* ----------------------------------------------------------- gradients */
static double oi3(const double *g) { return sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]); }

/* compute_nodal_grad, ns_euler_recon_module.F90:453-490; ng is (3,5) */
static void nodal_grad(const om_mesh *m, int v, const double *sol, double *ng) {
  double mat[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 15; i++) ng[i] = 0.0;
  for (int i = 1; i <= m->v_nsfn[v]; i++) {
    int sf = m->v_sf[v][i];
    int f = m->sf_face[sf];
    int le = m->f_left[f], re = m->f_right[f];
    if (re > 0) {
      double wl[5], wr[5];
      conserv_to_primit(&SOL(sol, le, 0), wl);
      conserv_to_primit(&SOL(sol, re, 0), wr);
      const double *n = &SN(m, sf, 0);
      double a = m->sf_area[sf];
      for (int k = 0; k < 5; k++) {
        double dw = wr[k] - wl[k];
        for (int d = 0; d < 3; d++) ng[3 * k + d] = ng[3 * k + d] + a * (n[d] * dw);
      }
      for (int b = 0; b < 3; b++) {
        double dx = EC(m, re, b) - EC(m, le, b);
        for (int d = 0; d < 3; d++) mat[d + 3 * b] = mat[d + 3 * b] + a * (n[d] * dx);
      }
    }
  }
  om_pinv3(mat);
  double out[15];
  for (int k = 0; k < 5; k++)
    for (int d = 0; d < 3; d++)
      out[3 * k + d] = mat[d] * ng[3 * k] + mat[d + 3] * ng[3 * k + 1] + mat[d + 6] * ng[3 * k + 2];
  for (int i = 0; i < 15; i++) ng[i] = m->vbound[v] ? 0.0 : out[i];
}

/* delta_sol_w = matmul(transpose(grad(:,:,i)), dx) */
static void grad_dot(const double *g, const double *dx, double *dw) {
  for (int k = 0; k < 5; k++) dw[k] = g[3 * k] * dx[0] + g[3 * k + 1] * dx[1] + g[3 * k + 2] * dx[2];
}

/* physical_slope_limiter, ns_euler_recon_module.F90:129-164 */
static void physical_slope_limiter(const om_mesh *m, const double *sol, int i, double *g) {
  const double tol = 1e-8;
  double w[5], dw[5], dx[3];
  conserv_to_primit(&SOL(sol, i, 0), w);
  double omega = 1.0;
  for (int j = 1; j <= m->e_nv[i]; j++) {
    int v = m->e_vert[i][j];
    for (int d = 0; d < 3; d++) dx[d] = VC(m, v, d) - EC(m, i, d);
    grad_dot(g, dx, dw);
    if (dw[0] < 0.0) omega = dmin(omega, dmax(0.0, w[0] / (tol - dw[0])));
    if (dw[4] < 0.0) omega = dmin(omega, dmax(0.0, w[4] / (tol - dw[4])));
  }
  for (int q = 0; q < 15; q++) g[q] = omega * g[q];
}

/* relaxed_maximum_principle, ns_euler_recon_module.F90:61-127 */
static void relaxed_maximum_principle(const om_mesh *m, const double *sol, int i, double *g) {
  const double tol = 1e-8;
  double rho_min = 1e100, rho_max = 0.0, p_min = 1e100, p_max = 0.0;
  double w[5], dw[5], dx[3];
  for (int j = 1; j <= m->e_nv[i]; j++) {
    int v = m->e_vert[i][j];
    for (int k = 1; k <= m->v_nen[v]; k++) {
      conserv_to_primit(&SOL(sol, m->v_elem[v][k], 0), w);
      p_min = dmin(p_min, w[4]);
      p_max = dmax(p_max, w[4]);
      rho_min = dmin(rho_min, w[0]);
      rho_max = dmax(rho_max, w[0]);
    }
  }
  p_min = (1.0 - 1e-2) * p_min;
  rho_min = (1.0 - 1e-2) * rho_min;
  p_max = (1.0 + 1e-2) * p_max;
  rho_max = (1.0 + 1e-2) * rho_max;
  conserv_to_primit(&SOL(sol, i, 0), w);
  double omega = 1.0;
  for (int j = 1; j <= m->e_nv[i]; j++) {
    int v = m->e_vert[i][j];
    for (int d = 0; d < 3; d++) dx[d] = VC(m, v, d) - EC(m, i, d);
    grad_dot(g, dx, dw);
    if (dw[0] < 0.0)
      omega = dmin(omega, dmax(0.0, (w[0] - rho_min) / (tol - dw[0])));
    else
      omega = dmin(omega, dmax(0.0, -(w[0] - rho_max) / (tol + dw[0])));
    if (dw[4] < 0.0)
      omega = dmin(omega, dmax(0.0, (w[4] - p_min) / (tol - dw[4])));
    else
      omega = dmin(omega, dmax(0.0, -(w[4] - p_max) / (tol + dw[4])));
  }
  for (int q = 0; q < 15; q++) g[q] = omega * g[q];
}

/* compute_cell_grad_method_{2..6}, ns_euler_recon_module.F90:166-320 */
static void cell_grad_method(const om_mesh *m, const double *ang, int e, int method, double *g) {
  const double eps = 1e-3;
  for (int q = 0; q < 15; q++) g[q] = 0.0;
  int nvv = m->e_nv[e];
  if (method == 2) {
    double dvs = 0.0;
    for (int j = 1; j <= nvv; j++) {
      int v = m->e_vert[e][j];
      const double *ng = &ang[15 * (v - 1)];
      for (int q = 0; q < 15; q++) g[q] = g[q] + m->vvol[v] * ng[q];
      dvs = dvs + m->vvol[v];
    }
    for (int q = 0; q < 15; q++) g[q] = g[q] / dvs;
  } else if (method == 3) {
    int v = m->e_vert[e][1];
    for (int q = 0; q < 15; q++) g[q] = ang[15 * (v - 1) + q];
    for (int j = 2; j <= nvv; j++) {
      v = m->e_vert[e][j];
      const double *ng = &ang[15 * (v - 1)];
      for (int k = 0; k < 5; k++)
        if (oi3(&ng[3 * k]) < oi3(&g[3 * k]))
          for (int d = 0; d < 3; d++) g[3 * k + d] = ng[3 * k + d];
    }
  } else if (method == 4 || method == 5 || method == 6) {
    double ws[5] = {0, 0, 0, 0, 0};
    double cg[15], cws = 0.0;
    for (int q = 0; q < 15; q++) cg[q] = 0.0;
    for (int j = 1; j <= nvv; j++) {
      int v = m->e_vert[e][j];
      const double *ng = &ang[15 * (v - 1)];
      for (int k = 0; k < 5; k++) {
        double den = eps + oi3(&ng[3 * k]);
        if (method != 4) den = den * den;
        double wgt = m->vvol[v] / den;
        for (int d = 0; d < 3; d++) g[3 * k + d] = g[3 * k + d] + wgt * ng[3 * k + d];
        ws[k] = ws[k] + wgt;
      }
      if (method == 6) {
        for (int q = 0; q < 15; q++) cg[q] = cg[q] + m->vvol[v] * ng[q];
        cws = cws + m->vvol[v];
      }
    }
    if (method == 6) {
      for (int q = 0; q < 15; q++) cg[q] = cg[q] / cws;
      /* OI(central_grad(:, :)) at :312,315: the 3 x 5 actual is sequence-
         associated with OI's dimension(3) dummy, which sees the first
         variable's gradient only; one weight serves all five variables */
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
}

/* compute_cell_grad_from_nodal_grad, ns_euler_recon_module.F90:8-59 */
int om_compute_cell_grad(const om_mesh *m, const double *sol, double *grad, double *all_nodal_grad,
                         int method) {
  for (int v = 1; v <= m->nv; v++) nodal_grad(m, v, sol, &all_nodal_grad[15 * (v - 1)]);
  if (method == 0 || method == 1) {
    for (long q = 0; q < 15L * m->ne; q++) grad[q] = 0.0;
    return OM_OK;
  }
  if (method < 0 || method > 6) return OM_ERR_BAD_INPUT;
  for (int e = 1; e <= m->ne; e++) {
    double *g = &grad[15 * (e - 1)];
    cell_grad_method(m, all_nodal_grad, e, method, g);
    physical_slope_limiter(m, sol, e, g);
    relaxed_maximum_principle(m, sol, e, g);
  }
  return OM_OK;
}

/* flux_limiting, ns_euler_recon_module.F90:560-601 */
void om_flux_limiting(const om_mesh *m, const double *sol, double *rhs, const double *rhs_o2,
                      const double *sum_lambda, double cfl) {
  double dt = 1e10;
  for (int i = 1; i <= m->ne; i++) dt = dmin(dt, m->evol[i] / sum_lambda[i - 1]);
  dt = cfl * dt;
  for (int i = 1; i <= m->ne; i++) {
    double c = m->evol[i] / dt + sum_lambda[i - 1];
    double a[5], b[5];
    for (int k = 0; k < 5; k++) {
      a[k] = SOL(sol, i, k) + SOL(rhs, i, k) / c;
      b[k] = (SOL(rhs_o2, i, k) - SOL(rhs, i, k)) / c;
    }
    double omega = om_max_omega_for_positivity(a, b);
    for (int k = 0; k < 5; k++)
      SOL(rhs, i, k) = SOL(rhs, i, k) + dmin(1.0, omega) * (SOL(rhs_o2, i, k) - SOL(rhs, i, k));
  }
}

/* add_timedisc_and_solve, ns_euler_module.F90:521-609 */
int om_add_timedisc_and_solve(const om_mesh *m, const om_config *cfg, double *sol, double *rhs,
                              double *diag, double *sum_lambda, double *delta_sol, double *res,
                              double *res_vect, double a1, double a2, double cfl, double t, double *dt_out,
                              double *omega_out) {
  double dt = 1e10;
  for (int i = 1; i <= m->ne; i++) dt = dmin(dt, m->evol[i] / sum_lambda[i - 1]);
  dt = cfl * dt;
  if (!cfg->local_time_step) {
    dt = dmin(cfg->tmax - t + 1e-12, dt);
    for (int i = 0; i < m->ne; i++) sum_lambda[i] = 0.0;
  }
  *res = 0.0;
  for (int k = 0; k < 5; k++) res_vect[k] = 0.0;
  double omega = 1.0;
  for (int i = 1; i <= m->ne; i++) {
    for (int k = 0; k < 5; k++) res_vect[k] = res_vect[k] + SOL(rhs, i, k) * SOL(rhs, i, k);
    double c = a1 * (m->evol[i] / dt + sum_lambda[i - 1]);
    double *D = &diag[25 * (i - 1)];
    for (int q = 0; q < 5; q++) D[q + 5 * q] = D[q + 5 * q] + c;
    for (int k = 0; k < 5; k++) SOL(rhs, i, k) = SOL(rhs, i, k) + c * SOL(sol, i, k);
    double nw[5];
    om_lu_solve(5, D, &SOL(rhs, i, 0), nw);
    double dd[5], ad[5];
    for (int k = 0; k < 5; k++) {
      dd[k] = nw[k] - SOL(sol, i, k);
      SOL(delta_sol, i, k) = dd[k];
      ad[k] = a2 * dd[k];
    }
    omega = dmin(omega, om_max_omega_for_positivity(&SOL(sol, i, 0), ad));
    *res = *res + sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] + dd[3] * dd[3] + dd[4] * dd[4]);
  }
  for (int k = 0; k < 5; k++) res_vect[k] = sqrt(res_vect[k]);
  if (omega_out) *omega_out = omega;
  if (omega <= 1e-6) {
    *dt_out = dt;
    return OM_ERR_BAD_OMEGA;
  }
  if (!cfg->local_time_step && omega < 1.0) dt = (1.0 - omega) * dt;
  for (int i = 1; i <= m->ne; i++)
    for (int k = 0; k < 5; k++) SOL(sol, i, k) = SOL(sol, i, k) + a2 * omega * SOL(delta_sol, i, k);
  *dt_out = dt;
  return OM_OK;
}

/* compute_rhs_ns, ns_main.F90:130-179 (diffusion: om_diffusion.c) */
int om_compute_rhs_ns(const om_mesh *m, const om_config *cfg, const double *sol, double *rhs, double *diag,
                      double *sum_lambda, double *cell_grad, double *nodal_grad, double *vp) {
  long ne = m->ne;
  for (long i = 0; i < 25 * ne; i++) diag[i] = 0.0;
  for (long i = 0; i < 15 * ne; i++) cell_grad[i] = 0.0;
  for (long i = 0; i < 15L * m->nv; i++) nodal_grad[i] = 0.0;
  int st = om_compute_euler_flux(m, cfg, sol, rhs, cell_grad, sum_lambda, 0, vp);
  if (st) return st;
  if (cfg->second_order) {
    double *rhs_o2 = (double *)calloc(5 * ne, sizeof(double));
    double *sl_o2 = (double *)calloc(ne, sizeof(double));
    st = om_compute_cell_grad(m, sol, cell_grad, nodal_grad, cfg->method);
    if (!st) st = om_compute_euler_flux(m, cfg, sol, rhs_o2, cell_grad, sl_o2, 1, vp);
    if (!st) om_flux_limiting(m, sol, rhs, rhs_o2, sum_lambda, cfg->cfl);
    free(rhs_o2);
    free(sl_o2);
  }
  if (!st) st = om_add_diffusion(m, cfg, sol, rhs, diag);
  return st;
}

/* dt of add_timedisc_and_solve, ns_euler_module.F90:546-560 */
static double stage_dt(const om_mesh *m, const om_config *cfg, const double *sum_lambda, double t) {
  double dt = 1e10;
  for (int i = 1; i <= m->ne; i++) dt = dmin(dt, m->evol[i] / sum_lambda[i - 1]);
  dt = cfg->cfl * dt;
  if (!cfg->local_time_step) dt = dmin(cfg->tmax - t + 1e-12, dt);
  return dt;
}

/* Forward-Euler stage with a prescribed dt: the per-cell arithmetic of
 * add_timedisc_and_solve (ns_euler_module.F90:565-608) with a1 = a2 = 1 and
 * sum_lambda = 0, including the positivity scaling of the increment. */
int om_stage_fixed_dt(const om_mesh *m, double *sol, double *rhs, double *diag, double *delta_sol,
                      double dt, double *omega_out) {
  double omega = 1.0;
  for (int i = 1; i <= m->ne; i++) {
    double c = 1.0 * (m->evol[i] / dt + 0.0);
    double *D = &diag[25 * (i - 1)];
    for (int q = 0; q < 5; q++) D[q + 5 * q] = D[q + 5 * q] + c;
    for (int k = 0; k < 5; k++) SOL(rhs, i, k) = SOL(rhs, i, k) + c * SOL(sol, i, k);
    double nw[5], dd[5];
    om_lu_solve(5, D, &SOL(rhs, i, 0), nw);
    for (int k = 0; k < 5; k++) {
      dd[k] = nw[k] - SOL(sol, i, k);
      SOL(delta_sol, i, k) = dd[k];
    }
    omega = dmin(omega, om_max_omega_for_positivity(&SOL(sol, i, 0), dd));
  }
  if (omega_out) *omega_out = omega;
  if (omega <= 1e-6) return OM_ERR_BAD_OMEGA;
  for (int i = 1; i <= m->ne; i++)
    for (int k = 0; k < 5; k++) SOL(sol, i, k) = SOL(sol, i, k) + omega * SOL(delta_sol, i, k);
  return OM_OK;
}

/* Time loop of ns_main.F90:83-107 (time_scheme 0: forward Euler, or the
 * two-stage scheme when second_order is set) and an SSP-RK3 variant
 * (time_scheme 1, requires local_time_step off) built from the same stage
 * operator L with one dt per step taken from the residual at U^n:
 * U1 = L(U^n), U2 = 0.75 U^n + 0.25 L(U1), U^{n+1} = (1/3) U^n + (2/3) L(U2).
 * out[0] = final t, out[1] = last dt, out[2] = last res. */
int om_run(const om_mesh *m, om_config *cfg, double *sol, int n_steps, int time_scheme, double *out) {
  long ne = m->ne, nv = m->nv;
  if (time_scheme == 1 && cfg->local_time_step) return OM_ERR_BAD_INPUT;
  double *rhs = (double *)calloc(5 * ne, sizeof(double));
  double *diag = (double *)calloc(25 * ne, sizeof(double));
  double *sl = (double *)calloc(ne, sizeof(double));
  double *cg = (double *)calloc(15 * ne, sizeof(double));
  double *ng = (double *)calloc(15 * nv, sizeof(double));
  double *vp = (double *)calloc(3 * nv, sizeof(double));
  double *dsol = (double *)calloc(5 * ne, sizeof(double));
  double *soln = (double *)calloc(5 * ne, sizeof(double));
  double t = cfg->t, dt = 0.0, res = 0.0, rv[5];
  int st = OM_OK;
  for (int it = 0; it < n_steps && !st; it++) {
    cfg->t = t;
    st = om_compute_rhs_ns(m, cfg, sol, rhs, diag, sl, cg, ng, vp);
    if (st) break;
    if (time_scheme == 0) {
      if (cfg->second_order) {
        memcpy(soln, sol, sizeof(double) * 5 * ne);
        st = om_add_timedisc_and_solve(m, cfg, sol, rhs, diag, sl, dsol, &res, rv, 2.0, 1.0, cfg->cfl, t, &dt, NULL);
        if (st) break;
        st = om_compute_rhs_ns(m, cfg, sol, rhs, diag, sl, cg, ng, vp);
        if (st) break;
        st = om_add_timedisc_and_solve(m, cfg, soln, rhs, diag, sl, dsol, &res, rv, 1.0, 1.0, cfg->cfl, t, &dt, NULL);
        if (st) break;
        memcpy(sol, soln, sizeof(double) * 5 * ne);
      } else {
        st = om_add_timedisc_and_solve(m, cfg, sol, rhs, diag, sl, dsol, &res, rv, 1.0, 1.0, cfg->cfl, t, &dt, NULL);
        if (st) break;
      }
    } else {
      dt = stage_dt(m, cfg, sl, t);
      memcpy(soln, sol, sizeof(double) * 5 * ne);
      st = om_stage_fixed_dt(m, sol, rhs, diag, dsol, dt, NULL);
      if (st) break;
      st = om_compute_rhs_ns(m, cfg, sol, rhs, diag, sl, cg, ng, vp);
      if (st) break;
      st = om_stage_fixed_dt(m, sol, rhs, diag, dsol, dt, NULL);
      if (st) break;
      for (long q = 0; q < 5 * ne; q++) sol[q] = 0.75 * soln[q] + 0.25 * sol[q];
      st = om_compute_rhs_ns(m, cfg, sol, rhs, diag, sl, cg, ng, vp);
      if (st) break;
      st = om_stage_fixed_dt(m, sol, rhs, diag, dsol, dt, NULL);
      if (st) break;
      for (long q = 0; q < 5 * ne; q++) sol[q] = (1.0 / 3.0) * soln[q] + (2.0 / 3.0) * sol[q];
    }
    t = t + dt;
    if (!cfg->local_time_step && t >= cfg->tmax) break;
  }
  cfg->t = t;
  if (out) {
    out[0] = t;
    out[1] = dt;
    out[2] = res;
  }
  free(rhs);
  free(diag);
  free(sl);
  free(cg);
  free(ng);
  free(vp);
  free(dsol);
  free(soln);
  return st;
}

/* test hooks for the dense solvers (column-major inputs) */
void om_test_lu_solve(int n, const double *A, const double *b, double *x) {
  double W[64];
  for (int i = 0; i < n * n; i++) W[i] = A[i];
  om_lu_solve(n, W, b, x);
}
void om_test_pinv3(double *A) { om_pinv3(A); }
void om_test_lu_solve_noswp(int n, const double *A, const double *b, double *x) { om_lu_solve_noswp(n, A, x, b); }
void om_test_pinv3_sym(double *A) { om_pinv3_sym(A, A); }
void om_test_primitives(const double *u, double *w, double *u_back, double *a) {
  conserv_to_primit(u, w);
  primit_to_conserv(w, u_back);
  *a = sound_speed_w(w);
}
