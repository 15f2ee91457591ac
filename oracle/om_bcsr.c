/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see om_internal.h).
 *
 * Restatement of the BCSR routines of
 * src/linear_algebra/sparse_bcsr_linear_module.F90:
 *   bcsr_matmul            :13-29
 *   bcsr_diag              :1414-1436 (block-Jacobi via lu_solve_noswp)
 *   modifiedArnoldi_diag   :1568-1590
 *   bcsr_gmres_diag        :1592-1691
 *   add_block_to_bcsr      :56-71
 *   bcsr_gmres / modifiedArnoldi                              :194-340
 *   bcsr_gmres_lusgs / modifiedArnoldi_lusgs                  :447-548, 642-665
 *   bcsr_gmres_parallel_lusgs / modifiedArnoldi_parallel_lusgs :885-1037, 583-611
 *   bcsr_pcgm              :1110-1169 (lu_solve(3, ...) on the diagonal block)
 *   bcsr_lusgs             :1268-1335 (two forward sweeps, as written)
 *   bcsr_parallel_lusgs    :1438-1502 (colour sets given by the caller)
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
/* rows of the lower triangle in the order bcsr_lusgs visits them: the
 * entries j = row(i), ... while col(j) < i (columns assumed increasing, as the
 * reference's do-while does) */
static void lower_apply(int b, const double *blk, const double *xb, double *yb) {
  for (int r = 0; r < b; r++) {
    double acc = 0.0;
    for (int c = 0; c < b; c++) acc = acc + blk[r + b * c] * xb[c];
    yb[r] = yb[r] - acc;
  }
}

static int diag_index(int i, const int *row, const int *col) {
  for (int j = row[i - 1]; j <= row[i] - 1; j++)
    if (col[j - 1] == i) return j;
  return -1;
}

/* bcsr_lusgs: (L + I) y = b, then (L + D) x = y, D^-1 by lu_solve_noswp */
void om_bcsr_lusgs(int m, int b, const int *row, const int *col, const double *v, double *x, const double *rhs) {
  long n = (long)b * m;
  double *y = (double *)malloc(sizeof(double) * n);
  double *t = (double *)malloc(sizeof(double) * b);
  for (int i = 1; i <= m; i++) {
    double *yb = &y[(long)b * (i - 1)];
    for (int r = 0; r < b; r++) yb[r] = rhs[(long)b * (i - 1) + r];
    for (int j = row[i - 1]; j <= row[i] - 1 && col[j - 1] < i; j++)
      lower_apply(b, &v[(long)(j - 1) * b * b], &y[(long)b * (col[j - 1] - 1)], yb);
  }
  for (int i = 1; i <= m; i++) {
    double *xb = &x[(long)b * (i - 1)];
    for (int r = 0; r < b; r++) xb[r] = y[(long)b * (i - 1) + r];
    for (int j = row[i - 1]; j <= row[i] - 1 && col[j - 1] < i; j++)
      lower_apply(b, &v[(long)(j - 1) * b * b], &x[(long)b * (col[j - 1] - 1)], xb);
    for (int r = 0; r < b; r++) t[r] = xb[r];
    om_lu_solve_noswp(b, &v[(long)(diag_index(i, row, col) - 1) * b * b], xb, t);
  }
  free(y);
  free(t);
}

/* bcsr_parallel_lusgs with colour sets: set cs holds rows set_rows[set_off[cs-1] .. set_off[cs]-1]
 * (1-based rows), color(i) in 1..nsets */
void om_bcsr_parallel_lusgs(int m, int b, const int *row, const int *col, const double *v, double *x,
                            const double *rhs, int nsets, const int *set_off, const int *set_rows,
                            const int *color) {
  long n = (long)b * m;
  double *y = (double *)malloc(sizeof(double) * n);
  double *t = (double *)malloc(sizeof(double) * b);
  for (int cs = 1; cs <= nsets; cs++)
    for (int k = set_off[cs - 1]; k < set_off[cs]; k++) {
      const int i = set_rows[k];
      double *yb = &y[(long)b * (i - 1)];
      for (int r = 0; r < b; r++) yb[r] = rhs[(long)b * (i - 1) + r];
      for (int j = row[i - 1]; j <= row[i] - 1; j++)
        if (color[col[j - 1] - 1] < cs) lower_apply(b, &v[(long)(j - 1) * b * b], &y[(long)b * (col[j - 1] - 1)], yb);
    }
  for (int cs = nsets; cs >= 1; cs--)
    for (int k = set_off[cs - 1]; k < set_off[cs]; k++) {
      const int i = set_rows[k];
      double *xb = &x[(long)b * (i - 1)];
      for (int r = 0; r < b; r++) xb[r] = y[(long)b * (i - 1) + r];
      for (int j = row[i - 1]; j <= row[i] - 1; j++)
        if (color[col[j - 1] - 1] > cs) lower_apply(b, &v[(long)(j - 1) * b * b], &x[(long)b * (col[j - 1] - 1)], xb);
      for (int r = 0; r < b; r++) t[r] = xb[r];
      om_lu_solve_noswp(b, &v[(long)(diag_index(i, row, col) - 1) * b * b], xb, t);
    }
  free(y);
  free(t);
}

/* add_block_to_bcsr: v(:,:,k) += M at the first k of row i with col(k) == j */
void om_bcsr_add_block(int m, int b, const int *row, const int *col, double *v, int i, int j, const double *M) {
  (void)m;
  for (int k = row[i - 1]; k <= row[i] - 1; k++)
    if (col[k - 1] == j) {
      for (int q = 0; q < b * b; q++) v[(long)(k - 1) * b * b + q] = v[(long)(k - 1) * b * b + q] + M[q];
      break;
    }
}

/* the preconditioner of a GMRES variant: 0 none, 1 block-Jacobi, 2 LU-SGS,
 * 3 coloured LU-SGS */
typedef struct {
  int kind, nsets;
  const int *set_off, *set_rows, *color;
} om_prec;

static void prec_apply(const om_prec *P, int m, int b, const int *row, const int *col, const double *v, double *x,
                       const double *r) {
  if (P->kind == 1)
    om_bcsr_diag(m, b, row, col, v, x, r);
  else if (P->kind == 2)
    om_bcsr_lusgs(m, b, row, col, v, x, r);
  else
    om_bcsr_parallel_lusgs(m, b, row, col, v, x, r, P->nsets, P->set_off, P->set_rows, P->color);
}

/* bcsr_gmres (kind 0), bcsr_gmres_diag (1), bcsr_gmres_lusgs (2),
 * bcsr_gmres_parallel_lusgs (3).  V is n x (m+1) column-major, H is
 * (m+1) x m.  Returns 1 when converged; iters[0] = restarts, iters[1] = rho. */
static int gmres_kind(const om_prec *P, int mrows, int b, const int *row, const int *col, const double *v,
                      double *x, const double *rhs, double tol, int m, double *iters_out) {
  const int maxit = P->kind == 1 ? 40 : 20;
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
        if (P->kind == 0) {
          om_bcsr_matmul(mrows, b, row, col, v, VV(j - 1), zp);
        } else {
          prec_apply(P, mrows, b, row, col, v, VV(j), VV(j - 1));
          om_bcsr_matmul(mrows, b, row, col, v, VV(j), zp);
        }
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
      if (P->kind == 0) {
        for (int ii = 1; ii <= nr; ii++) {
          double zi = z[ii - 1];
          double *vi = VV(ii);
          for (long q = 0; q < n; q++) x[q] = x[q] + zi * vi[q];
        }
      } else {
        for (long q = 0; q < n; q++) r[q] = 0.0;
        for (int ii = 1; ii <= nr; ii++) {
          double zi = z[ii - 1];
          double *vi = VV(ii);
          for (long q = 0; q < n; q++) r[q] = r[q] + zi * vi[q];
        }
        prec_apply(P, mrows, b, row, col, v, zp, r);
        for (long q = 0; q < n; q++) x[q] = x[q] + zp[q];
      }
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

int om_bcsr_gmres_diag(int mrows, int b, const int *row, const int *col, const double *v, double *x,
                       const double *rhs, double tol, int m, double *iters_out) {
  om_prec P = {1, 0, NULL, NULL, NULL};
  return gmres_kind(&P, mrows, b, row, col, v, x, rhs, tol, m, iters_out);
}

/* kind 0 plain, 2 LU-SGS, 3 coloured LU-SGS (colour sets as in om_bcsr_parallel_lusgs) */
int om_bcsr_gmres_kind(int kind, int mrows, int b, const int *row, const int *col, const double *v, double *x,
                       const double *rhs, double tol, int m, int nsets, const int *set_off, const int *set_rows,
                       const int *color, double *iters_out) {
  om_prec P = {kind, nsets, set_off, set_rows, color};
  return gmres_kind(&P, mrows, b, row, col, v, x, rhs, tol, m, iters_out);
}

/* bcsr_pcgm: preconditioned CG from the given x; the preconditioner solves
 * each diagonal block with lu_solve(3, ...) (the reference hard-codes 3).
 * Returns 1 when sqrt(mr . r) <= tol at the end (the reference stops with
 * an error otherwise), -1 when a diagonal block is singular (lu_solve's stop);
 * iters[0] = iterations, iters[1] = sqrt(mr . r).  D holds the first nine
 * values of the b x b block, as the 3 x 3 dummy argument sees them. */
int om_bcsr_pcgm(int mrows, int b, const int *row, const int *col, const double *v, double *x, const double *rhs,
                 double tol, double *iters_out) {
  long n = (long)b * mrows;
  double *r = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n);
  double *Ap = (double *)malloc(sizeof(double) * n);
  double *mr = (double *)malloc(sizeof(double) * n);
  double D[9];
  double rpdrp = 0.0, rdr;
  int k = 0, conv;
  int sing = 0;
#define PREC()                                                                                      \
  for (int i = 1; i <= mrows; i++) {                                                                \
    memcpy(D, &v[(long)(diag_index(i, row, col) - 1) * b * b], sizeof(D));                          \
    sing |= om_lu_solve_plain(3, D, &mr[(long)b * (i - 1)], &r[(long)b * (i - 1)]);                 \
  }
  om_bcsr_matmul(mrows, b, row, col, v, x, r);
  for (long q = 0; q < n; q++) r[q] = rhs[q] - r[q];
  for (long q = 0; q < n; q++) mr[q] = 0.0;
  PREC();
  rdr = vdot(n, mr, r);
  if (sqrt(rdr) < tol) {
    conv = 1;
    rpdrp = rdr;
  } else {
    memcpy(p, mr, sizeof(double) * n);
    for (k = 1; k <= n; k++) {
      om_bcsr_matmul(mrows, b, row, col, v, p, Ap);
      double alpha = rdr / vdot(n, p, Ap);
      for (long q = 0; q < n; q++) x[q] = x[q] + alpha * p[q];
      for (long q = 0; q < n; q++) r[q] = r[q] - alpha * Ap[q];
      PREC();
      rpdrp = vdot(n, mr, r);
      if (sqrt(rpdrp) <= tol)
        if (vnorm(n, r) <= tol) break;
      double beta = rpdrp / rdr;
      for (long q = 0; q < n; q++) p[q] = mr[q] + beta * p[q];
      rdr = rpdrp;
    }
    conv = !(sqrt(rpdrp) > tol);
  }
  if (sing) conv = -1;
#undef PREC
  if (iters_out) {
    iters_out[0] = (double)(k > n ? n : k);
    iters_out[1] = sqrt(rpdrp);
  }
  free(r);
  free(p);
  free(Ap);
  free(mr);
  return conv;
}
