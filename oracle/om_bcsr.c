/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see om_internal.h).
 *
 * Restatement of the BCSR routines of
 * src/linear_algebra/sparse_bcsr_linear_module.F90:
 *   bcsr_matmul            :13-29
 *   bcsr_diag              :1414-1436 (block-Jacobi via lu_solve_noswp)
 *   modifiedArnoldi_diag   :1568-1590
 *   bcsr_gmres_diag        :1592-1691
 * Block storage follows bcsr_matrix_type (:6-11): v(b,b,nnzb) column-major,
 * 1-based row(m+1) and col(nnzb).  norm2 of a vector is sqrt of the
 * sequential sum of squares; dot_product is the sequential sum.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "om_dense.h"

typedef struct {
  int m, n, b;
  const int *row, *col;
  const double *v;
} om_bcsr;

/* y = A x */
void om_bcsr_matmul(int m, int b, const int *row, const int *col, const double *v, const double *x,
                    double *y) {
  for (long i = 0; i < (long)b * m; i++) y[i] = 0.0;
  for (int i = 1; i <= m; i++)
    for (int j = row[i - 1]; j <= row[i] - 1; j++) {
      const double *blk = &v[(long)(j - 1) * b * b];
      const double *xb = &x[(long)b * (col[j - 1] - 1)];
      double *yb = &y[(long)b * (i - 1)];
      for (int r = 0; r < b; r++) {
        double acc = 0.0;
        for (int c = 0; c < b; c++) acc = acc + blk[r + b * c] * xb[c];
        yb[r] = yb[r] + acc;
      }
    }
}

/* x = D^{-1} b with D the diagonal blocks */
void om_bcsr_diag(int m, int b, const int *row, const int *col, const double *v, double *x,
                  const double *rhs) {
  for (int i = 1; i <= m; i++)
    for (int j = row[i - 1]; j <= row[i] - 1; j++)
      if (col[j - 1] == i) {
        om_lu_solve_noswp(b, &v[(long)(j - 1) * b * b], &x[(long)b * (i - 1)], &rhs[(long)b * (i - 1)]);
        break;
      }
}

static double vdot(long n, const double *a, const double *c) {
  double s = 0.0;
  for (long i = 0; i < n; i++) s = s + a[i] * c[i];
  return s;
}
static double vnorm(long n, const double *a) { return sqrt(vdot(n, a, a)); }

/* bcsr_gmres_diag.  V is n x (m+1) column-major, H is (m+1) x m.
 * Returns 1 when converged.  iters[0] = restarts used, iters[1] = final rho. */
int om_bcsr_gmres_diag(int mrows, int b, const int *row, const int *col, const double *v, double *x,
                       const double *rhs, double tol, int m, double *iters_out) {
  const int maxit = 40;
  long n = (long)b * mrows;
  double *V = (double *)malloc(sizeof(double) * n * (m + 1));
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *r = (double *)malloc(sizeof(double) * n);
  double *z = (double *)malloc(sizeof(double) * n);
  double *zp = (double *)malloc(sizeof(double) * n);
  double *y = (double *)calloc(m + 2, sizeof(double));
  double *c = (double *)calloc(m + 2, sizeof(double));
  double *s = (double *)calloc(m + 2, sizeof(double));
#define HH(i, j) H[((i)-1) + (long)(m + 1) * ((j)-1)]
#define VV(j) (&V[n * ((j)-1)])
  int conv = 0, jrest = 0;
  for (long i = 0; i < n; i++) x[i] = 0.0;
  om_bcsr_matmul(mrows, b, row, col, v, x, r);
  for (long i = 0; i < n; i++) r[i] = rhs[i] - r[i];
  double rho = vnorm(n, r);
  if (rho < tol) {
    conv = 1;
  } else {
    for (jrest = 1; jrest <= maxit; jrest++) {
      for (int i = 0; i <= m; i++) y[i + 1] = 0.0;
      y[1] = rho;
      for (long i = 0; i < n; i++) r[i] = r[i] / y[1];
      int nr = m;
      /* modifiedArnoldi_diag */
      memcpy(VV(1), r, sizeof(double) * n);
      for (int j = 2; j <= m + 1; j++) {
        om_bcsr_diag(mrows, b, row, col, v, VV(j), VV(j - 1));
        om_bcsr_matmul(mrows, b, row, col, v, VV(j), zp);
        memcpy(VV(j), zp, sizeof(double) * n);
        for (int i = 1; i <= j - 1; i++) {
          HH(i, j - 1) = vdot(n, VV(i), VV(j));
          double h = HH(i, j - 1);
          double *vj = VV(j), *vi = VV(i);
          for (long q = 0; q < n; q++) vj[q] = vj[q] - h * vi[q];
        }
        HH(j, j - 1) = vnorm(n, VV(j));
        double h = HH(j, j - 1);
        double *vj = VV(j);
        for (long q = 0; q < n; q++) vj[q] = vj[q] / h;
      }
      /* Givens rotations */
      int i;
      int exited = 0;
      for (i = 1; i <= m; i++) {
        for (int k = 2; k <= i; k++) {
          double tmp = HH(k - 1, i);
          HH(k - 1, i) = c[k - 1] * HH(k - 1, i) + s[k - 1] * HH(k, i);
          HH(k, i) = -s[k - 1] * tmp + c[k - 1] * HH(k, i);
        }
        double delta = sqrt(HH(i, i) * HH(i, i) + HH(i + 1, i) * HH(i + 1, i));
        c[i] = HH(i, i) / delta;
        s[i] = HH(i + 1, i) / delta;
        HH(i, i) = c[i] * HH(i, i) + s[i] * HH(i + 1, i);
        for (int k = i + 1; k <= m + 1; k++) HH(k, i) = 0.0;
        y[i + 1] = -s[i] * y[i];
        y[i] = c[i] * y[i];
        rho = fabs(y[i + 1]);
        if (rho < tol) {
          nr = i + 1;
          exited = 1;
          break;
        }
      }
      if (!exited) i = m + 1;
      /* back substitution into z(1:nr) */
      for (long q = 0; q < n; q++) z[q] = 0.0;
      if (i >= m) {
        nr = m;
        z[nr - 1] = y[nr] / HH(nr, nr);
      }
      for (int k = nr - 1; k >= 1; k--) {
        z[k - 1] = y[k];
        for (int l = k + 1; l <= nr; l++) z[k - 1] = z[k - 1] - HH(k, l) * z[l - 1];
        z[k - 1] = z[k - 1] / HH(k, k);
      }
      for (long q = 0; q < n; q++) r[q] = 0.0;
      for (int ii = 1; ii <= nr; ii++) {
        double zi = z[ii - 1];
        double *vi = VV(ii);
        for (long q = 0; q < n; q++) r[q] = r[q] + zi * vi[q];
      }
      om_bcsr_diag(mrows, b, row, col, v, zp, r);
      for (long q = 0; q < n; q++) x[q] = x[q] + zp[q];
      om_bcsr_matmul(mrows, b, row, col, v, x, r);
      for (long q = 0; q < n; q++) r[q] = rhs[q] - r[q];
      rho = vnorm(n, r);
      if (rho < tol) {
        conv = 1;
        break;
      }
    }
  }
  if (iters_out) {
    iters_out[0] = (double)(jrest > maxit ? maxit : jrest);
    iters_out[1] = rho;
  }
#undef HH
#undef VV
  free(V);
  free(H);
  free(r);
  free(z);
  free(zp);
  free(y);
  free(c);
  free(s);
  return conv;
}
