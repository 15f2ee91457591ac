/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see om_internal.h).
 *
 * Viscous terms of compute_rhs_ns (src/ns/ns_main.F90:170-175) with the
 * viscosity of compute_mu (ns_vectorial_diffusion_module.F90:322-359):
 * mu_arr = mu_p and mu0corr = mu_p, or with use_sutherland the per-cell
 * mu_sutherland(temp_u(sol)) and mu0corr = their minimum:
 *   build_momentum_diff_nodal_system        ns_vectorial_diffusion_module.F90:7-198
 *   build_momentum_diff_mat_into_global_mat_diag                          :200-320
 *   build_diff_nodal_system                 ns_diffusion_module.F90:7-110
 *   build_diff_mat_into_global_mat_diag                                   :112-184
 *   kappa, dtheta_dU                                                      :186-206
 *   temp_u                                  ns_euler_primitives_module.F90:52-63
 * Dense helpers follow linear_solver_module.F90: inv_lapack (:629-641) is
 * the LU factorisation of om_lu_solve's order followed by LAPACK's
 * unblocked dgetri (inverse of U by columns, back-solve against L, column
 * interchanges); dgemm/dgemv sums run over k in increasing order from zero,
 * as reference BLAS loops them; matmul3_blas (:832-856) picks (A B) C when
 * m n k + m p n <= k n p + m p k.  norm2 is the plain square root of the sum
 * of squares.  Loops follow the reference's vertex / sub-face / sub-element
 * order, so every rhs and diag entry receives its terms in the same order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "om_internal.h"

#define GAMMA 1.4
#define SOLV(a, i, k) ((a)[5 * ((long)(i)-1) + (k)])
#define SN(m, s, d) ((m)->sf_norm[3 * (s) + (d)])
#define DIAG(D, c, i, j) ((D)[25 * ((long)(c)-1) + (i) + 5 * (j)])

/* LU with partial pivoting in om_lu_solve's order; A column-major n x n */
static void lu_factor(int n, double *A, int *ipiv) {
  for (int k = 0; k < n; k++) {
    int p = k;
    double amax = fabs(A[k + (long)n * k]);
    for (int i = k + 1; i < n; i++)
      if (fabs(A[i + (long)n * k]) > amax) {
        amax = fabs(A[i + (long)n * k]);
        p = i;
      }
    ipiv[k] = p;
    if (A[p + (long)n * k] != 0.0) {
      if (p != k)
        for (int j = 0; j < n; j++) {
          double t = A[k + (long)n * j];
          A[k + (long)n * j] = A[p + (long)n * j];
          A[p + (long)n * j] = t;
        }
      double piv = A[k + (long)n * k];
      if (fabs(piv) >= 2.2250738585072014e-308) {
        double r = 1.0 / piv;
        for (int i = k + 1; i < n; i++) A[i + (long)n * k] = A[i + (long)n * k] * r;
      } else {
        for (int i = k + 1; i < n; i++) A[i + (long)n * k] = A[i + (long)n * k] / piv;
      }
    }
    for (int j = k + 1; j < n; j++)
      for (int i = k + 1; i < n; i++) A[i + (long)n * j] = A[i + (long)n * j] - A[i + (long)n * k] * A[k + (long)n * j];
  }
}

/* inv_lapack: A := A^-1 (dgetrf + unblocked dgetri / dtrti2) */
void om_inv(int n, double *A) {
  int *ipiv = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  double *work = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  lu_factor(n, A, ipiv);
#define AA(i, j) A[(i) + (long)n * (j)]
  /* dtrti2, upper, non-unit: inverse of U in place */
  for (int j = 0; j < n; j++) {
    AA(j, j) = 1.0 / AA(j, j);
    double ajj = -AA(j, j);
    /* dtrmv upper no-transpose non-unit on x = A(0:j-1, j) */
    for (int jj = 0; jj < j; jj++) {
      if (AA(jj, j) != 0.0) {
        double temp = AA(jj, j);
        for (int i = 0; i < jj; i++) AA(i, j) = AA(i, j) + temp * AA(i, jj);
        AA(jj, j) = AA(jj, j) * AA(jj, jj);
      }
    }
    for (int i = 0; i < j; i++) AA(i, j) = ajj * AA(i, j);
  }
  /* solve inv(A) L = inv(U), unblocked */
  for (int j = n - 1; j >= 0; j--) {
    for (int i = j + 1; i < n; i++) {
      work[i] = AA(i, j);
      AA(i, j) = 0.0;
    }
    if (j < n - 1) {
      /* dgemv 'N': A(:, j) = A(:, j) - A(:, j+1:n) * work(j+1:n) */
      for (int jj = j + 1; jj < n; jj++) {
        double temp = -1.0 * work[jj];
        for (int i = 0; i < n; i++) AA(i, j) = AA(i, j) + temp * AA(i, jj);
      }
    }
  }
  for (int j = n - 2; j >= 0; j--) {
    int jp = ipiv[j];
    if (jp != j)
      for (int i = 0; i < n; i++) {
        double t = AA(i, j);
        AA(i, j) = AA(i, jp);
        AA(i, jp) = t;
      }
  }
#undef AA
  free(ipiv);
  free(work);
}

/* C(m,n) = A(m,k) B(k,n), column-major, sums over k in order */
static void gemm(int m, int n, int k, const double *A, const double *B, double *C) {
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < m; i++) C[i + (long)m * j] = 0.0;
    for (int l = 0; l < k; l++) {
      double temp = B[l + (long)k * j];
      for (int i = 0; i < m; i++) C[i + (long)m * j] = C[i + (long)m * j] + temp * A[i + (long)m * l];
    }
  }
}

/* y(m) = A(m,n) x(n) */
static void gemv(int m, int n, const double *A, const double *x, double *y) {
  for (int i = 0; i < m; i++) y[i] = 0.0;
  for (int j = 0; j < n; j++) {
    double temp = x[j];
    for (int i = 0; i < m; i++) y[i] = y[i] + temp * A[i + (long)m * j];
  }
}

/* matmul3_blas(A, B, C, D, m, k, n, p): D = A B C */
static void matmul3(const double *A, const double *B, const double *C, double *D, int m, int k, int n, int p) {
  long cost1 = (long)m * n * k + (long)m * p * n;
  long cost2 = (long)k * n * p + (long)m * p * k;
  if (cost1 <= cost2) {
    double *T1 = (double *)malloc(sizeof(double) * m * n);
    gemm(m, n, k, A, B, T1);
    gemm(m, p, n, T1, C, D);
    free(T1);
  } else {
    double *T2 = (double *)malloc(sizeof(double) * k * p);
    gemm(k, p, n, B, C, T2);
    gemm(m, p, k, A, T2, D);
    free(T2);
  }
}

/* matvec2_blas(A, B, C, D, m, n, p): D = A (B C) */
static void matvec2(const double *A, const double *B, const double *C, double *D, int m, int n, int p) {
  double *tmp = (double *)malloc(sizeof(double) * n);
  gemv(n, p, B, C, tmp);
  gemv(m, n, A, tmp, D);
  free(tmp);
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* temp_u */
static double temp_u(const om_config *cfg, const double *u) {
  double tau = 1.0 / u[0];
  return (u[4] * tau - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * (tau * tau)) / cfg->Cv_p;
}

/* dtheta_dU */
static void dtheta_du(const om_config *cfg, const double *u, double *P) {
  double v1 = u[1] / u[0], v2 = u[2] / u[0], v3 = u[3] / u[0];
  double nrm = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
  P[0] = (nrm * nrm - u[4] / u[0]) / (u[0] * cfg->Cv_p);
  for (int k = 1; k < 4; k++) P[k] = -u[k] / ((u[0] * u[0]) * cfg->Cv_p);
  P[4] = 1.0 / (u[0] * cfg->Cv_p);
}

static double kappa(const om_config *cfg, double mu) { return mu * cfg->Cv_p * GAMMA / cfg->prandtl; }

/* mu_sutherland (ns_vectorial_diffusion_module.F90:349-359), C0 = 110.4
 * (ns_global_data_module.F90:20); T**1.5 through pow */
double om_mu_sutherland(const om_config *cfg, const double *u) {
  const double T = temp_u(cfg, u);
  return cfg->mu0 * pow(T / cfg->T0, 1.5) * (cfg->T0 + 110.4) / (T + 110.4);
}

/* compute_mu (:322-347): mu of every cell (mu[0..ne-1]) and the return value
 * mu0corr, the minimum (from 1e12) with use_sutherland, else mu_p */
double om_compute_mu(int ne, const om_config *cfg, const double *sol, double *mu) {
  if (!cfg->use_sutherland) {
    for (int c = 0; c < ne; c++) mu[c] = cfg->mu_p;
    return cfg->mu_p;
  }
  double mn = 1e12;
  for (int c = 0; c < ne; c++) {
    mu[c] = om_mu_sutherland(cfg, sol + 5L * c);
    mn = fmin(mn, mu[c]);
  }
  return mn;
}

/* -------------------------------------------------------------- heat */
static void heat_nodal_system(const om_mesh *m, const om_config *cfg, const double *mu_arr, int v, double *N,
                              double *S, double *StT, double *B) {
  const int nf = m->v_nsfn[v], ns = m->v_nsen[v];
  memset(N, 0, sizeof(double) * nf * nf);
  memset(S, 0, sizeof(double) * nf * ns);
  memset(StT, 0, sizeof(double) * ns * nf);
  memset(B, 0, sizeof(double) * nf);
  for (int l = 1; l <= nf; l++) {
    const int isfl = m->v_sf[v][l];
    const int lo = m->sf_loc[isfl] - 1;
    const int ifl = m->sf_face[isfl];
    for (int side = 0; side < 2; side++) {
      int ise = side == 0 ? m->sf_lse[isfl] : m->sf_rse[isfl];
      double norml[3];
      for (int d = 0; d < 3; d++) norml[d] = side == 0 ? SN(m, isfl, d) : -SN(m, isfl, d);
      if (side == 1 && ise <= 0) {
        int id_elem = m->f_right[ifl];
        if (id_elem < 0 && cfg->bc_T_id[-id_elem] == 2) {
          B[lo] = m->sf_area[isfl] * cfg->bc_val_T[-id_elem];
          for (int c = 0; c < ns; c++) S[lo + nf * c] = 0.0;
          for (int c = 0; c < nf; c++) N[lo + nf * c] = 0.0;
          N[lo + nf * lo] = m->sf_area[isfl];
        }
        continue;
      }
      const int id_elem = side == 0 ? m->f_left[ifl] : m->f_right[ifl];
      const int so = m->se_loc[ise] - 1;
      const double kap = kappa(cfg, mu_arr[id_elem]);
      for (int p = 1; p <= m->se_nsf[ise]; p++) {
        const int isfp = m->se_sf[ise][p];
        const int po = m->sf_loc[isfp] - 1;
        double normp[3];
        for (int d = 0; d < 3; d++) normp[d] = (m->sf_lse[isfp] == ise) ? SN(m, isfp, d) : -SN(m, isfp, d);
        double blc = kap * m->sf_area[isfl] * m->sf_area[isfp] * (1.0 / m->se_vol[ise]) * dot3(normp, norml);
        N[lo + nf * po] = N[lo + nf * po] + blc;
        S[lo + nf * so] = S[lo + nf * so] + blc;
        StT[so + ns * lo] = StT[so + ns * lo] + blc;
      }
    }
  }
}

static void heat_into_global(const om_mesh *m, const om_config *cfg, const double *mu_arr, const double *sol,
                             double *rhs, double *diag) {
  const int cap = m->nmax_sfn > 0 ? m->nmax_sfn : 1;
  double *N = (double *)malloc(sizeof(double) * cap * cap);
  double *S = (double *)malloc(sizeof(double) * cap * cap);
  double *StT = (double *)malloc(sizeof(double) * cap * cap);
  double *B = (double *)malloc(sizeof(double) * cap);
  double *SNS = (double *)malloc(sizeof(double) * cap * cap);
  double *SB = (double *)malloc(sizeof(double) * cap);
  double *theta = (double *)malloc(sizeof(double) * cap);
  double *P = (double *)malloc(sizeof(double) * 5 * cap);
  double *b1 = (double *)malloc(sizeof(double) * cap);
  for (int v = 1; v <= m->nv; v++) {
    const int nf = m->v_nsfn[v], ns = m->v_nsen[v];
    heat_nodal_system(m, cfg, mu_arr, v, N, S, StT, B);
    om_inv(nf, N);
    matmul3(StT, N, S, SNS, ns, nf, nf, ns);
    matvec2(StT, N, B, SB, ns, nf, nf);
    for (int l = 0; l < ns; l++) {
      const int iel = m->se_elem[m->v_se[v][l + 1]];
      theta[l] = temp_u(cfg, &SOLV(sol, iel, 0));
      dtheta_du(cfg, &SOLV(sol, iel, 0), &P[5 * l]);
      double s = 0.0;
      for (int c = 0; c < nf; c++) s = s + StT[l + ns * c];
      b1[l] = s;
    }
    for (int l = 0; l < ns; l++) {
      const int iel = m->se_elem[m->v_se[v][l + 1]];
      for (int p = 0; p < ns; p++) SOLV(rhs, iel, 4) = SOLV(rhs, iel, 4) + SNS[l + ns * p] * theta[p];
      const double sll = SNS[l + ns * l];
      for (int j = 0; j < 5; j++) DIAG(diag, iel, 4, j) = DIAG(diag, iel, 4, j) + (b1[l] - sll) * P[5 * l + j];
      double pu = 0.0;
      for (int k = 0; k < 5; k++) pu = pu + P[5 * l + k] * SOLV(sol, iel, k);
      SOLV(rhs, iel, 4) = SOLV(rhs, iel, 4) - sll * pu + (b1[l] - sll) * pu - b1[l] * theta[l] + SB[l];
    }
  }
  free(N), free(S), free(StT), free(B), free(SNS), free(SB), free(theta), free(P), free(b1);
}

/* ---------------------------------------------------------- momentum */
/* blc_3 of the mu-formula: a_l a_p (1/vol) (mu (nl.np) I + (mu - mu0) np nl^T
 * + (mu0 - 2/3 mu) nl np^T); tp(i,j) = np(i) nl(j) */
static void blc_mu(double al, double ap, double vol, double mu, double mu0, const double *nl, const double *np,
                   double *blc) {
  const double s = al * ap * (1.0 / vol);
  const double d = dot3(nl, np);
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) {
      const double tp = np[i] * nl[j], tpt = np[j] * nl[i];
      blc[i + 3 * j] = s * (mu * d * (i == j ? 1.0 : 0.0) + (mu - mu0) * tp + (mu0 - (2.0 / 3.0) * mu) * tpt);
    }
}

static void add3(double *M, int ld, int i0, int j0, const double *b) {
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) M[(i0 + i) + (long)ld * (j0 + j)] = M[(i0 + i) + (long)ld * (j0 + j)] + b[i + 3 * j];
}
static void add3t(double *M, int ld, int i0, int j0, const double *b) {
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) M[(i0 + i) + (long)ld * (j0 + j)] = M[(i0 + i) + (long)ld * (j0 + j)] + b[j + 3 * i];
}

static void mom_nodal_system(const om_mesh *m, const om_config *cfg, const double *mu_arr, double mu0, int v,
                             double *N, double *S, double *StT, double *B) {
  const int nf = m->v_nsfn[v], ns = m->v_nsen[v];
  const int F = 3 * nf, E = 3 * ns;
  memset(N, 0, sizeof(double) * F * F);
  memset(S, 0, sizeof(double) * F * E);
  memset(StT, 0, sizeof(double) * E * F);
  memset(B, 0, sizeof(double) * F);
  double blc[9];
  for (int l = 1; l <= nf; l++) {
    const int isfl = m->v_sf[v][l];
    const int lo = m->sf_loc[isfl] - 1;
    const int ifl0 = 3 * lo;
    const int id_elem_r = m->sf_re[isfl];
    int neumann, dirichlet;
    if (id_elem_r > 0) {
      neumann = dirichlet = 0;
    } else if (id_elem_r < 0) {
      dirichlet = cfg->bc_V_id[-id_elem_r] == 2;
      neumann = cfg->bc_V_id[-id_elem_r] == 1;
    } else {
      neumann = 1;
      dirichlet = 0;
    }
    const double al = m->sf_area[isfl];
    if (!neumann && !dirichlet) {
      for (int side = 0; side < 2; side++) {
        const int ise = side == 0 ? m->sf_lse[isfl] : m->sf_rse[isfl];
        const int ise0 = 3 * (m->se_loc[ise] - 1);
        double nl[3];
        for (int d = 0; d < 3; d++) nl[d] = side == 0 ? SN(m, isfl, d) : -SN(m, isfl, d);
        for (int p = 1; p <= m->se_nsf[ise]; p++) {
          const int isfp = m->se_sf[ise][p];
          const int ifp0 = 3 * (m->sf_loc[isfp] - 1);
          double np[3];
          for (int d = 0; d < 3; d++) np[d] = (m->sf_lse[isfp] == ise) ? SN(m, isfp, d) : -SN(m, isfp, d);
          blc_mu(al, m->sf_area[isfp], m->se_vol[ise], mu_arr[m->se_elem[ise]], mu0, nl, np, blc);
          add3t(StT, E, ise0, ifl0, blc);
          add3(N, F, ifl0, ifp0, blc);
          add3(S, F, ifl0, ise0, blc);
        }
      }
    } else if (neumann) {
      const int ise = m->sf_lse[isfl];
      const int ise0 = 3 * (m->se_loc[ise] - 1);
      double nl[3];
      for (int d = 0; d < 3; d++) nl[d] = SN(m, isfl, d);
      for (int p = 1; p <= m->se_nsf[ise]; p++) {
        const int isfp = m->se_sf[ise][p];
        const int ifp0 = 3 * (m->sf_loc[isfp] - 1);
        double np[3];
        for (int d = 0; d < 3; d++) np[d] = (m->sf_lse[isfp] == ise) ? SN(m, isfp, d) : -SN(m, isfp, d);
        const double s = al * m->sf_area[isfp] * (1.0 / m->se_vol[ise]);
        const double dd = dot3(nl, np);
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) blc[i + 3 * j] = s * (dd * (i == j ? 1.0 : 0.0) + (np[j] * nl[i]) / 3.0);
        add3(N, F, ifl0, ifp0, blc);
        add3(S, F, ifl0, ise0, blc);
        blc_mu(al, m->sf_area[isfp], m->se_vol[ise], mu_arr[m->se_elem[ise]], mu0, nl, np, blc);
        add3t(StT, E, ise0, ifl0, blc);
      }
    } else {
      const int ise = m->sf_lse[isfl];
      const int ise0 = 3 * (m->se_loc[ise] - 1);
      double nl[3];
      for (int d = 0; d < 3; d++) nl[d] = SN(m, isfl, d);
      for (int p = 1; p <= m->se_nsf[ise]; p++) {
        const int isfp = m->se_sf[ise][p];
        double np[3];
        for (int d = 0; d < 3; d++) np[d] = (m->sf_lse[isfp] == ise) ? SN(m, isfp, d) : -SN(m, isfp, d);
        blc_mu(al, m->sf_area[isfp], m->se_vol[ise], mu_arr[m->se_elem[ise]], mu0, nl, np, blc);
        add3t(StT, E, ise0, ifl0, blc);
      }
      for (int d = 0; d < 3; d++) B[ifl0 + d] = al * cfg->bc_val_V[-id_elem_r][d];
      for (int r = ifl0; r < ifl0 + 3; r++) {
        for (int c = 0; c < E; c++) S[r + (long)F * c] = 0.0;
        for (int c = 0; c < F; c++) N[r + (long)F * c] = 0.0;
      }
      for (int d = 0; d < 3; d++) N[(ifl0 + d) + (long)F * (ifl0 + d)] = al * 1.0;
    }
  }
}

static void mom_into_global(const om_mesh *m, const om_config *cfg, const double *mu_arr, double mu0corr,
                            const double *sol, double *rhs, double *diag) {
  const int capF = 3 * (m->nmax_sfn > 0 ? m->nmax_sfn : 1);
  double *N = (double *)malloc(sizeof(double) * capF * capF);
  double *S = (double *)malloc(sizeof(double) * capF * capF);
  double *StT = (double *)malloc(sizeof(double) * capF * capF);
  double *SNS = (double *)malloc(sizeof(double) * capF * capF);
  double *B = (double *)malloc(sizeof(double) * capF);
  double *SB = (double *)malloc(sizeof(double) * capF);
  double *Ve = (double *)malloc(sizeof(double) * capF);
  double *Vf = (double *)malloc(sizeof(double) * capF);
  double *SNSVe = (double *)malloc(sizeof(double) * capF);
  double *blc3 = (double *)malloc(sizeof(double) * 9 * capF);
  for (int v = 1; v <= m->nv; v++) {
    const int nf = m->v_nsfn[v], ns = m->v_nsen[v];
    const int F = 3 * nf, E = 3 * ns;
    mom_nodal_system(m, cfg, mu_arr, mu0corr, v, N, S, StT, B);
    for (int l = 0; l < ns; l++) {
      const int iel = m->se_elem[m->v_se[v][l + 1]];
      for (int k = 0; k < 3; k++) Ve[3 * l + k] = SOLV(sol, iel, k + 1) / SOLV(sol, iel, 0);
    }
    for (int p = 0; p < ns; p++) {
      const int iep = m->se_elem[m->v_se[v][p + 1]];
      double Spc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int l = 0; l < nf; l++)
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) Spc[i + 3 * j] = Spc[i + 3 * j] + StT[(3 * p + i) + (long)E * (3 * l + j)];
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) blc3[9 * p + i + 3 * j] = Spc[j + 3 * i] / SOLV(sol, iep, 0);
    }
    gemv(F, E, S, Ve, Vf);
    for (int i = 0; i < F; i++) Vf[i] = Vf[i] + B[i];
    om_inv(F, N);
    matmul3(StT, N, S, SNS, E, F, F, E);
    matvec2(StT, N, B, SB, E, F, F);
    gemv(F, F, N, Vf, B);
    for (int i = 0; i < F; i++) Vf[i] = B[i];
    gemv(E, E, SNS, Ve, SNSVe);
    for (int l = 0; l < ns; l++) {
      const int iel = m->se_elem[m->v_se[v][l + 1]];
      const double rho = SOLV(sol, iel, 0);
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++)
          DIAG(diag, iel, 1 + i, 1 + j) =
              DIAG(diag, iel, 1 + i, 1 + j) - SNS[(3 * l + i) + (long)E * (3 * l + j)] / rho + blc3[9 * l + i + 3 * j];
      for (int i = 0; i < 3; i++) {
        double mv = 0.0;
        for (int j = 0; j < 3; j++) mv = mv + SNS[(3 * l + i) + (long)E * (3 * l + j)] * Ve[3 * l + j];
        SOLV(rhs, iel, 1 + i) = SOLV(rhs, iel, 1 + i) + SNSVe[3 * l + i] - mv + SB[3 * l + i];
      }
    }
    /* viscous work */
    for (int p = 0; p < ns; p++) {
      const int isep = m->v_se[v][p + 1];
      const int iep = m->se_elem[isep];
      double Spc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int l = 1; l <= m->se_nsf[isep]; l++) {
        const int isfl = m->se_sf[isep][l];
        const int lo = m->sf_loc[isfl] - 1;
        double nl[3], vm[3];
        for (int d = 0; d < 3; d++) nl[d] = (m->sf_lse[isfl] == isep) ? SN(m, isfl, d) : -SN(m, isfl, d);
        for (int d = 0; d < 3; d++) vm[d] = Vf[3 * lo + d] - Ve[3 * p + d];
        const double dv = dot3(vm, nl);
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) {
            const double tp = vm[i] * nl[j], tpt = vm[j] * nl[i];
            Spc[i + 3 * j] = Spc[i + 3 * j] + m->sf_area[isfl] * (tp + tpt - (2.0 / 3.0) * dv * (i == j ? 1.0 : 0.0));
          }
      }
      for (int q = 0; q < 9; q++) Spc[q] = Spc[q] * mu_arr[iep] / m->se_vol[isep];
      for (int l = 1; l <= m->se_nsf[isep]; l++) {
        const int isfl = m->se_sf[isep][l];
        const int lo = m->sf_loc[isfl] - 1;
        double nl[3], sv[3];
        for (int d = 0; d < 3; d++) nl[d] = (m->sf_lse[isfl] == isep) ? SN(m, isfl, d) : -SN(m, isfl, d);
        for (int i = 0; i < 3; i++)
          sv[i] = Spc[i] * Vf[3 * lo] + Spc[i + 3] * Vf[3 * lo + 1] + Spc[i + 6] * Vf[3 * lo + 2];
        SOLV(rhs, iep, 4) = SOLV(rhs, iep, 4) + m->sf_area[isfl] * dot3(sv, nl);
      }
    }
  }
  free(N), free(S), free(StT), free(SNS), free(B), free(SB), free(Ve), free(Vf), free(SNSVe), free(blc3);
}

/* the diffusion block of compute_rhs_ns: momentum first, then heat */
int om_add_diffusion(const om_mesh *m, const om_config *cfg, const double *sol, double *rhs, double *diag) {
  if (!cfg->activate_diffusion) return OM_OK;
  double *mu_arr = (double *)malloc(sizeof(double) * ((size_t)m->ne + 1));
  if (!mu_arr) return OM_ERR_BAD_INPUT;
  const double mu0corr = om_compute_mu(m->ne, cfg, sol, mu_arr + 1);
  mom_into_global(m, cfg, mu_arr, mu0corr, sol, rhs, diag);
  heat_into_global(m, cfg, mu_arr, sol, rhs, diag);
  free(mu_arr);
  return OM_OK;
}

/* the two builders for given cell viscosities mu_arr[0..ne) and mu0corr:
 * parts 1 build_momentum_diff_mat_into_global_mat_diag, 2
 * build_diff_mat_into_global_mat_diag, 3 both (momentum first) */
int om_add_diffusion_mu(const om_mesh *m, const om_config *cfg, const double *sol, const double *mu_cells,
                        double mu0corr, int parts, double *rhs, double *diag) {
  double *mu_arr = (double *)malloc(sizeof(double) * ((size_t)m->ne + 1));
  if (!mu_arr) return OM_ERR_BAD_INPUT;
  mu_arr[0] = 0.0;
  memcpy(mu_arr + 1, mu_cells, sizeof(double) * (size_t)m->ne);
  if (parts & 1) mom_into_global(m, cfg, mu_arr, mu0corr, sol, rhs, diag);
  if (parts & 2) heat_into_global(m, cfg, mu_arr, sol, rhs, diag);
  free(mu_arr);
  return OM_OK;
}

int om_diffusion_available(void) { return 1; }

/* test hook: temp_u with Cv_p */
double om_test_temp_u(const double *u, double cv) {
  om_config c;
  memset(&c, 0, sizeof(c));
  c.Cv_p = cv;
  return temp_u(&c, u);
}

/* test hook: inverse of a column-major n x n matrix */
void om_test_inv(int n, const double *A, double *out) {
  memcpy(out, A, sizeof(double) * n * n);
  om_inv(n, out);
}
