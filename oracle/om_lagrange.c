/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see om_internal.h).
 *
 * The node sweeps of src/lagrange and src/acoustic, restated on the same
 * mesh connectivity:
 *   compute_rhs_lagrange   src/lagrange/lagrange_module.F90:14-147
 *                          (pressure :429-436, inverse_3_by_3
 *                          src/linear_algebra/linear_solver_module.F90:125-172)
 *   acoustic time step     src/acoustic/acoustic_main.F90:171-290 with
 *                          multi_point_pressure acoustic_module.F90:11-35 and
 *                          rotate :232-262
 * Both scatter into the cells with one running sum per cell, in vertex order
 * and, within a vertex, in sub-face order, as the reference loops do.
 * norm2 is the square root of the sequential sum of squares.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

#define GAMMA_L (7.0 / 5.0)
#define FN(m, f, d) ((m)->fnorm[3 * (f) + (d)])
#define SOLL(s, c, k) ((s)[5 * ((long)(c)-1) + (k)])
#define SOLA(s, c, k) ((s)[4 * ((long)(c)-1) + (k)])

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }
static double norm2_3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* pressure(u) of the Lagrangian state (rho, u, v, w, E) */
static double pressure_l(const double *u) {
  const double nrm = norm2_3(u + 1);
  return (GAMMA_L - 1.0) / u[0] * (u[4] - 0.5 * (nrm * nrm));
}

static double det2(double a, double b, double c, double d) { return a * d - c * b; }

/* inverse_3_by_3: adjugate of the transpose over the cofactor determinant;
 * column-major in and out */
void om_inverse_3_by_3(const double *mat, double *inv) {
#define M(i, j) mat[((i)-1) + 3 * ((j)-1)]
#define T(i, j) M(j, i)
  double adj[3][3];
  adj[0][0] = det2(T(2, 2), T(2, 3), T(3, 2), T(3, 3));
  adj[0][1] = -det2(T(2, 1), T(2, 3), T(3, 1), T(3, 3));
  adj[0][2] = det2(T(2, 1), T(2, 2), T(3, 1), T(3, 2));
  adj[1][0] = -det2(T(1, 2), T(1, 3), T(3, 2), T(3, 3));
  adj[1][1] = det2(T(1, 1), T(1, 3), T(3, 1), T(3, 3));
  adj[1][2] = -det2(T(1, 1), T(1, 2), T(3, 1), T(3, 2));
  adj[2][0] = det2(T(1, 2), T(1, 3), T(2, 2), T(2, 3));
  adj[2][1] = -det2(T(1, 1), T(1, 3), T(2, 1), T(2, 3));
  adj[2][2] = det2(T(1, 1), T(1, 2), T(2, 1), T(2, 2));
  const double det = M(1, 1) * adj[0][0] + M(1, 2) * adj[1][0] + M(1, 3) * adj[2][0];
  const double r = 1.0 / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) inv[i + 3 * j] = r * adj[i][j];
#undef T
#undef M
}

static void matvec3(const double *A, const double *x, double *y) {
  for (int i = 0; i < 3; i++) {
    double acc = 0.0;
    for (int j = 0; j < 3; j++) acc = acc + A[i + 3 * j] * x[j];
    y[i] = acc;
  }
}

/* compute_rhs_lagrange: sol(5, ne) = (rho, u, v, w, E) per cell, mass(ne);
 * rhs(5, ne) and vp(3, nv) out (vp[3*(v-1) + d]). */
void om_compute_rhs_lagrange(const om_mesh *m, const double *sol, const double *mass, int b2d, double *rhs,
                             double *vp) {
  for (long q = 0; q < 5L * m->ne; q++) rhs[q] = 0.0;
  for (long q = 0; q < 3L * m->nv; q++) vp[q] = 0.0;
  const int nmax = m->nmax_sfn > 0 ? m->nmax_sfn : 1;
  double *lam = (double *)malloc(sizeof(double) * 2 * (nmax + 1));
  for (int i = 1; i <= m->nv; i++) {
    const int nsfn = m->v_nsfn[i];
    double Mp[9] = {0}, Rp[3] = {0}, Bp[3] = {0};
    for (int j = 1; j <= nsfn; j++) {
      const int sf = m->v_sf[i][j], f = m->sf_face[sf];
      const int idl = m->f_left[f], idr = m->f_right[f];
      const double *n = &FN(m, f, 0), area = m->sf_area[sf];
      const double *ul = &SOLL(sol, idl, 0);
      const double pl = pressure_l(ul), vl = dot3(ul + 1, n);
      double l1, l2 = 0.0;
      if (idr > 0) {
        const double *ur = &SOLL(sol, idr, 0);
        const double pr = pressure_l(ur), vr = dot3(ur + 1, n);
        l1 = dmax(dmax(sqrt(GAMMA_L * pl * ul[0]) / ul[0], sqrt(dmax(0.0, pr - pl) / ul[0])), -(vr - vl) / ul[0]);
        l2 = dmax(dmax(sqrt(GAMMA_L * pr * ur[0]) / ur[0], sqrt(dmax(0.0, pl - pr) / ur[0])), -(vr - vl) / ur[0]);
        const double vbar = (l1 * vl + l2 * vr - (pr - pl)) / (l2 + l1);
        const double s = area * (l1 + l2);
        for (int b = 0; b < 3; b++)
          for (int a = 0; a < 3; a++) Mp[a + 3 * b] = Mp[a + 3 * b] + s * (n[a] * n[b]);
        for (int a = 0; a < 3; a++) Rp[a] = Rp[a] + s * n[a] * vbar;
      } else if (b2d && fabs(n[2]) > 1e-8) {
        l1 = sqrt(GAMMA_L * pl * ul[0]) / ul[0];
        const double s = area * 2 * l1;
        for (int b = 0; b < 3; b++)
          for (int a = 0; a < 3; a++) Mp[a + 3 * b] = Mp[a + 3 * b] + s * (n[a] * n[b]);
      } else {
        l1 = sqrt(GAMMA_L * pl * ul[0]) / ul[0];
        for (int a = 0; a < 3; a++) Bp[a] = Bp[a] + area * n[a];
        const double s = area * l1;
        for (int b = 0; b < 3; b++)
          for (int a = 0; a < 3; a++) Mp[a + 3 * b] = Mp[a + 3 * b] + s * (n[a] * n[b]);
        const double sr = area * (pl + l1 * vl);
        for (int a = 0; a < 3; a++) Rp[a] = Rp[a] + sr * n[a];
      }
      lam[2 * j] = l1;
      lam[2 * j + 1] = l2;
    }
    double Mi[9], v3[3];
    om_inverse_3_by_3(Mp, Mi);
    if (dmax(dmax(fabs(Bp[0]), fabs(Bp[1])), fabs(Bp[2])) > 1e-8) {
      double mb[3], w[3];
      matvec3(Mi, Bp, mb);
      const double Pp = dot3(Rp, mb) / dot3(Bp, mb);
      for (int a = 0; a < 3; a++) w[a] = Rp[a] - Pp * Bp[a];
      matvec3(Mi, w, v3);
    } else {
      matvec3(Mi, Rp, v3);
    }
    for (int a = 0; a < 3; a++) vp[3L * (i - 1) + a] = v3[a];
    for (int j = 1; j <= nsfn; j++) {
      const int sf = m->v_sf[i][j], f = m->sf_face[sf];
      const int idl = m->f_left[f], idr = m->f_right[f];
      const double *n = &FN(m, f, 0), area = m->sf_area[sf];
      const double *ul = &SOLL(sol, idl, 0);
      const double v_et = dot3(v3, n), vl = dot3(ul + 1, n), pl = pressure_l(ul);
      const double pl_et = pl - lam[2 * j] * (v_et - vl);
      double flux[5] = {-v_et, pl_et * n[0], pl_et * n[1], pl_et * n[2], pl_et * v_et};
      for (int k = 0; k < 5; k++) SOLL(rhs, idl, k) = SOLL(rhs, idl, k) - area / mass[idl - 1] * flux[k];
      if (idr > 0) {
        const double *ur = &SOLL(sol, idr, 0);
        const double vr = dot3(ur + 1, n), pr = pressure_l(ur);
        const double pr_et = pr + lam[2 * j + 1] * (v_et - vr);
        double fr[5] = {-v_et, pr_et * n[0], pr_et * n[1], pr_et * n[2], pr_et * v_et};
        for (int k = 0; k < 5; k++) SOLL(rhs, idr, k) = SOLL(rhs, idr, k) + area / mass[idr - 1] * fr[k];
      }
    }
  }
  free(lam);
}

/* ---------------------------------------------------------------- acoustic */
#define RHO_A 2.0
#define A_A 1.0

static void rotate(double *v, const double *n1, int forward) {
  double n2[3] = {n1[2], 0.0, -n1[0]}, n3[3], vt[3];
  double nn2 = norm2_3(n2);
  if (nn2 > 1e-10) {
    for (int d = 0; d < 3; d++) n2[d] = n2[d] / nn2;
  } else {
    n2[0] = 0.0;
    n2[1] = -n1[2];
    n2[2] = n1[1];
    const double q = norm2_3(n2);
    for (int d = 0; d < 3; d++) n2[d] = n2[d] / q;
  }
  n3[0] = n1[1] * n2[2] - n1[2] * n2[1];
  n3[1] = -(n1[0] * n2[2] - n1[2] * n2[0]);
  n3[2] = n1[0] * n2[1] - n1[1] * n2[0];
  const double q3 = norm2_3(n3);
  for (int d = 0; d < 3; d++) n3[d] = n3[d] / q3;
  if (forward) {
    vt[0] = dot3(v, n1);
    vt[1] = dot3(v, n2);
    vt[2] = dot3(v, n3);
  } else {
    for (int d = 0; d < 3; d++) vt[d] = v[0] * n1[d] + v[1] * n2[d] + v[2] * n3[d];
  }
  for (int d = 0; d < 3; d++) v[d] = vt[d];
}

/* One step of the acoustic loop: sol(4, ne) = (p, u, v, w) updated in place,
 * *dt_out the step.  bc_kind[tag] for tag 1..n_bc: 0 wall, 1 freestream (bc_val
 * [4*tag + k]), 2 outflow; right_elem_neigh 0 is a wall.  Returns 0, or 1 for a
 * tag with no supported kind. */
int om_acoustic_step(const om_mesh *m, double *sol, int n_bc, const int *bc_kind, const double *bc_val, double cfl,
                     double *dt_out) {
  double dt = 1e30;
  double *flux = (double *)calloc(4L * m->ne, sizeof(double));
  const int nmax = m->nmax_sfn > 0 ? m->nmax_sfn : 1;
  double *lr = (double *)malloc(sizeof(double) * 8 * (nmax + 1));
  int st = 0;
  for (int v = 1; v <= m->nv && !st; v++) {
    const int nsfn = m->v_nsfn[v];
    for (int j = 1; j <= nsfn; j++) {
      const int sf = m->v_sf[v][j], f = m->sf_face[sf];
      const int il = m->sf_le[sf], ir = m->sf_re[sf];
      const double *n = &FN(m, f, 0);
      double sl[4], sr[4] = {0.0, 0.0, 0.0, 0.0};
      for (int k = 0; k < 4; k++) sl[k] = SOLA(sol, il, k);
      int kind = 0;
      if (ir > 0) {
        for (int k = 0; k < 4; k++) sr[k] = SOLA(sol, ir, k);
        kind = -1;
      } else if (ir < 0) {
        if (-ir > n_bc) {
          st = 1;
          break;
        }
        kind = bc_kind[-ir];
      }
      if (kind == 0) {
        const double un = dot3(sl + 1, n);
        sr[0] = sl[0];
        for (int d = 0; d < 3; d++) sr[1 + d] = sl[1 + d] - 2.0 * un * n[d];
      } else if (kind == 1) {
        for (int k = 0; k < 4; k++) sr[k] = bc_val[4 * (-ir) + k];
      } else if (kind == 2) {
        for (int k = 0; k < 4; k++) sr[k] = sl[k];
      } else if (kind != -1) {
        st = 1;
        break;
      }
      rotate(sl + 1, n, 1);
      rotate(sr + 1, n, 1);
      for (int k = 0; k < 4; k++) {
        lr[8 * j + k] = sl[k];
        lr[8 * j + 4 + k] = sr[k];
      }
    }
    if (st) break;
    double p_nodal = 0.0, tot = 0.0;
    for (int j = 1; j <= nsfn; j++) {
      const int sf = m->v_sf[v][j];
      const double pt = 0.5 * (lr[8 * j] + lr[8 * j + 4]) - 0.5 * RHO_A * A_A * (lr[8 * j + 5] - lr[8 * j + 1]);
      p_nodal = p_nodal + m->sf_area[sf] * pt;
      tot = tot + m->sf_area[sf];
    }
    p_nodal = p_nodal / tot;
    for (int j = 1; j <= nsfn; j++) {
      const int sf = m->v_sf[v][j], f = m->sf_face[sf];
      const int il = m->sf_le[sf], ir = m->sf_re[sf];
      const double *n = &FN(m, f, 0), area = m->sf_area[sf];
      const double *sl = &lr[8 * j], *sr = &lr[8 * j + 4];
      const double smax = A_A;
      const double ul_et = sl[1] - (p_nodal - sl[0]) / (RHO_A * A_A);
      const double ur_et = sr[1] + (p_nodal - sr[0]) / (RHO_A * A_A);
      double fl[4] = {-(RHO_A * (A_A * A_A) * ul_et), -(p_nodal / RHO_A), -0.0, -0.0};
      double fr[4] = {RHO_A * (A_A * A_A) * ur_et, p_nodal / RHO_A, 0.0, 0.0};
      rotate(fl + 1, n, 0);
      rotate(fr + 1, n, 0);
      for (int k = 0; k < 4; k++) SOLA(flux, il, k) = SOLA(flux, il, k) + area * fl[k];
      if (ir > 0)
        for (int k = 0; k < 4; k++) SOLA(flux, ir, k) = SOLA(flux, ir, k) + area * fr[k];
      dt = dmin(dt, m->evol[il] / (m->farea[f] * smax));
      if (ir > 0) dt = dmin(dt, m->evol[ir] / (m->farea[f] * smax));
    }
  }
  if (!st)
    for (int c = 1; c <= m->ne; c++)
      for (int k = 0; k < 4; k++) SOLA(sol, c, k) = SOLA(sol, c, k) + (cfl * dt / m->evol[c]) * SOLA(flux, c, k);
  *dt_out = dt;
  free(flux);
  free(lr);
  return st;
}
