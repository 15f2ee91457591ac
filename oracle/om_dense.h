/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see om_internal.h).
 *
 * Small dense solves used on the path, restated from published algorithms:
 *  - om_lu_solve: LU with partial pivoting as reference LAPACK dgetrf2 +
 *    dgetrs compute it (pivot = first largest |a|, L column scaled by the
 *    reciprocal of the pivot, sequential rank-1 updates; dtrsm-ordered
 *    triangular solves).  Replaces lu_solve_inplace_lapack,
 *    src/linear_algebra/linear_solver_module.F90:643-657.
 *  - om_pinv3: pseudo-inverse V * S^+ * U^T with singular values below
 *    n*1e-12*max(S) dropped (linear_solver_module.F90:683-715), where the SVD
 *    is computed by one-sided (Hestenes) Jacobi rotations instead of dgesvd.
 *  - om_lu_solve_noswp: linear_solver_module.F90:413-445 verbatim order.
 *  - both pseudo-inverses first try the cofactor inverse (om_inv3_regular):
 *    when the determinant proves every singular value above the cut-off,
 *    the pseudo-inverse is the inverse and no SVD is needed.
 */
#ifndef OM_DENSE_H
#define OM_DENSE_H
#include <math.h>
#include <string.h>

/* A is column-major n x n (A[i + n*j]) and is overwritten by its factors. */
static inline void om_lu_solve(int n, double *A, const double *b, double *x) {
  int ipiv[8];
  for (int k = 0; k < n; k++) {
    int p = k;
    double amax = fabs(A[k + n * k]);
    for (int i = k + 1; i < n; i++)
      if (fabs(A[i + n * k]) > amax) {
        amax = fabs(A[i + n * k]);
        p = i;
      }
    ipiv[k] = p;
    if (A[p + n * k] != 0.0) {
      if (p != k)
        for (int j = 0; j < n; j++) {
          double t = A[k + n * j];
          A[k + n * j] = A[p + n * j];
          A[p + n * j] = t;
        }
      double piv = A[k + n * k];
      if (fabs(piv) >= 2.2250738585072014e-308) {
        double r = 1.0 / piv;
        for (int i = k + 1; i < n; i++) A[i + n * k] = A[i + n * k] * r;
      } else {
        for (int i = k + 1; i < n; i++) A[i + n * k] = A[i + n * k] / piv;
      }
    }
    for (int j = k + 1; j < n; j++)
      for (int i = k + 1; i < n; i++) A[i + n * j] = A[i + n * j] - A[i + n * k] * A[k + n * j];
  }
  double w[8];
  for (int i = 0; i < n; i++) w[i] = b[i];
  for (int k = 0; k < n; k++)
    if (ipiv[k] != k) {
      double t = w[k];
      w[k] = w[ipiv[k]];
      w[ipiv[k]] = t;
    }
  for (int k = 0; k < n; k++)
    if (w[k] != 0.0)
      for (int i = k + 1; i < n; i++) w[i] = w[i] - w[k] * A[i + n * k];
  for (int k = n - 1; k >= 0; k--)
    if (w[k] != 0.0) {
      w[k] = w[k] / A[k + n * k];
      for (int i = 0; i < k; i++) w[i] = w[i] - w[k] * A[i + n * k];
    }
  for (int i = 0; i < n; i++) x[i] = w[i];
}

/* One-sided Jacobi SVD of a column-major 3x3 matrix: a = U diag(s) V^T with
 * s sorted in decreasing order. */
static inline void om_svd3(const double *a_in, double *s, double *u, double *v) {
  double a[9];
  for (int i = 0; i < 9; i++) a[i] = a_in[i];
  for (int i = 0; i < 9; i++) v[i] = 0.0;
  v[0] = v[4] = v[8] = 1.0;
  for (int sweep = 0; sweep < 16; sweep++) {
    int rotated = 0;
    for (int pq = 0; pq < 3; pq++) {
      int p = (pq == 2) ? 1 : 0;
      int q = (pq == 0) ? 1 : 2;
      double al = 0.0, be = 0.0, ga = 0.0;
      for (int i = 0; i < 3; i++) {
        al = al + a[i + 3 * p] * a[i + 3 * p];
        be = be + a[i + 3 * q] * a[i + 3 * q];
        ga = ga + a[i + 3 * p] * a[i + 3 * q];
      }
      if (ga == 0.0) continue;
      if (fabs(ga) <= 1e-15 * (sqrt(al) * sqrt(be))) continue;
      rotated = 1;
      double zeta = (be - al) / (2.0 * ga);
      double t;
      if (fabs(zeta) > 1e100)
        t = 0.5 / zeta;
      else
        t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
      double c = 1.0 / sqrt(1.0 + t * t);
      double sn = c * t;
      for (int i = 0; i < 3; i++) {
        double ap = a[i + 3 * p], aq = a[i + 3 * q];
        a[i + 3 * p] = c * ap - sn * aq;
        a[i + 3 * q] = sn * ap + c * aq;
        double vp = v[i + 3 * p], vq = v[i + 3 * q];
        v[i + 3 * p] = c * vp - sn * vq;
        v[i + 3 * q] = sn * vp + c * vq;
      }
    }
    if (!rotated) break;
  }
  double sv[3];
  for (int k = 0; k < 3; k++) {
    sv[k] = sqrt(a[3 * k] * a[3 * k] + a[3 * k + 1] * a[3 * k + 1] + a[3 * k + 2] * a[3 * k + 2]);
    for (int i = 0; i < 3; i++) u[i + 3 * k] = (sv[k] > 0.0) ? a[i + 3 * k] / sv[k] : 0.0;
  }
  /* stable sort by decreasing singular value */
  int ord[3] = {0, 1, 2};
  for (int i = 1; i < 3; i++) {
    int key = ord[i], j = i - 1;
    while (j >= 0 && sv[ord[j]] < sv[key]) {
      ord[j + 1] = ord[j];
      j--;
    }
    ord[j + 1] = key;
  }
  double uu[9], vv[9];
  for (int k = 0; k < 3; k++) {
    s[k] = sv[ord[k]];
    for (int i = 0; i < 3; i++) {
      uu[i + 3 * k] = u[i + 3 * ord[k]];
      vv[i + 3 * k] = v[i + 3 * ord[k]];
    }
  }
  for (int i = 0; i < 9; i++) {
    u[i] = uu[i];
    v[i] = vv[i];
  }
}

/* Cofactor inverse of the 3x3 matrix with rows (a b c), (d e f), (g h i),
 * written column-major into P, when it is provably far from the
 * pseudo-inverse cut-off; returns 0 (P untouched) otherwise.  With F the
 * Frobenius norm, s_max <= F and s_min = |det|/(s_max s_mid) >= 2|det|/F^2,
 * so |det| > 2e-12 F^3 gives s_min > 3.9e-12 s_max: all three singular values
 * pass the 3e-12 * max cut-off of linear_solver_module.F90:703-711 with a
 * margin far above the rounding of det (< 3e-15 F^3), and the pseudo-inverse
 * equals the inverse. */
static inline int om_inv3_regular(double a, double b, double c, double d, double e, double f, double g, double h,
                                  double i, double F2, double *P) {
  const double c00 = e * i - f * h, c01 = f * g - d * i, c02 = d * h - e * g;
  const double det = a * c00 + b * c01 + c * c02;
  if (!(fabs(det) > 2e-12 * (F2 * sqrt(F2)))) return 0;
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
  return 1;
}

/* In-place pseudo-inverse of a column-major 3x3 matrix. */
static inline void om_pinv3(double *A) {
  {
    double F2 = 0.0, P[9];
    for (int k = 0; k < 9; k++) F2 = F2 + A[k] * A[k];
    if (om_inv3_regular(A[0], A[3], A[6], A[1], A[4], A[7], A[2], A[5], A[8], F2, P)) {
      memcpy(A, P, sizeof(P));
      return;
    }
  }
  double s[3], u[9], v[9];
  om_svd3(A, s, u, v);
  double tol = 3.0 * 1e-12 * s[0];
  double sp[3];
  for (int k = 0; k < 3; k++) sp[k] = (s[k] > tol) ? 1.0 / s[k] : 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double acc = 0.0;
      for (int k = 0; k < 3; k++) acc = acc + v[i + 3 * k] * (sp[k] * u[j + 3 * k]);
      A[i + 3 * j] = acc;
    }
}

/* Pseudo-inverse of a symmetric 3x3 matrix (the nodal matrices of the
 * velocity solvers): the cofactor inverse when om_inv3_regular accepts it,
 * otherwise cyclic Jacobi rotations M = Q L Q^T, singular values
 * |L| sorted decreasingly, eigenvalues with |l| <= 3e-12 max|l| dropped;
 * P = sum_k q_k (1/l_k) q_k^T.  A column-major in, P column-major out. */
static inline void om_pinv3_sym(const double *A, double *P) {
  {
    const double a00 = A[0], a01 = A[3], a02 = A[6], a11 = A[4], a12 = A[7], a22 = A[8];
    const double F2 = ((a00 * a00 + a11 * a11) + a22 * a22) + 2.0 * ((a01 * a01 + a02 * a02) + a12 * a12);
    if (om_inv3_regular(a00, a01, a02, a01, a11, a12, a02, a12, a22, F2, P)) return;
  }
  double a[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = A[i + 3 * j];
  double q[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int sweep = 0; sweep < 16; sweep++) {
    int rotated = 0;
    for (int pq = 0; pq < 3; pq++) {
      int p = (pq == 2) ? 1 : 0, qq = (pq == 0) ? 1 : 2, r = 3 - p - qq;
      double app = a[p][p], aqq = a[qq][qq], apq = a[p][qq], arp = a[r][p], arq = a[r][qq];
      if (apq == 0.0) continue;
      if (fabs(apq) <= 1e-15 * sqrt(fabs(app) * fabs(aqq))) continue;
      rotated = 1;
      double theta = (aqq - app) / (2.0 * apq);
      double t;
      if (fabs(theta) > 1e100)
        t = 0.5 / theta;
      else
        t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(1.0 + theta * theta));
      double c = 1.0 / sqrt(1.0 + t * t);
      double s = t * c;
      for (int i = 0; i < 3; i++) {
        double vp = q[i + 3 * p], vq = q[i + 3 * qq];
        q[i + 3 * p] = c * vp - s * vq;
        q[i + 3 * qq] = s * vp + c * vq;
      }
      a[p][p] = app - t * apq;
      a[qq][qq] = aqq + t * apq;
      a[p][qq] = a[qq][p] = 0.0;
      a[r][p] = a[p][r] = c * arp - s * arq;
      a[r][qq] = a[qq][r] = s * arp + c * arq;
    }
    if (!rotated) break;
  }
  double l[3] = {a[0][0], a[1][1], a[2][2]};
  int ord[3] = {0, 1, 2};
  for (int i = 1; i < 3; i++) {
    int key = ord[i], j = i - 1;
    while (j >= 0 && fabs(l[ord[j]]) < fabs(l[key])) {
      ord[j + 1] = ord[j];
      j--;
    }
    ord[j + 1] = key;
  }
  double tol = 3.0 * 1e-12 * fabs(l[ord[0]]);
  double ip[3];
  for (int k = 0; k < 3; k++) ip[k] = (fabs(l[ord[k]]) > tol) ? 1.0 / l[ord[k]] : 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double acc = 0.0;
      for (int k = 0; k < 3; k++) acc = acc + q[i + 3 * ord[k]] * (ip[k] * q[j + 3 * ord[k]]);
      P[i + 3 * j] = acc;
    }
}

/* lu_solve, linear_solver_module.F90:447-484: partial pivoting by the first
 * largest |C(i,k)| (rows always swapped), multipliers C(i,k) * (1/C(k,k)),
 * right-hand side eliminated with the factorisation, back substitution over
 * j increasing.  Returns 1 where the reference stops (|C(n,n)| < 1e-12). */
static inline int om_lu_solve_plain(int n, const double *A, double *x, const double *b) {
  double C[64];
  for (int q = 0; q < n * n; q++) C[q] = A[q];
  for (int i = 0; i < n; i++) x[i] = b[i];
  for (int k = 0; k < n; k++) {
    int p = k;
    double amax = fabs(C[k + n * k]);
    for (int i = k + 1; i < n; i++)
      if (fabs(C[i + n * k]) > amax) {
        amax = fabs(C[i + n * k]);
        p = i;
      }
    for (int j = 0; j < n; j++) {
      double t = C[k + n * j];
      C[k + n * j] = C[p + n * j];
      C[p + n * j] = t;
    }
    double t = x[k];
    x[k] = x[p];
    x[p] = t;
    double invdiag = 1.0 / C[k + n * k];
    for (int i = k + 1; i < n; i++) {
      double lik = C[i + n * k] * invdiag;
      for (int j = k; j < n; j++) C[i + n * j] = C[i + n * j] - lik * C[k + n * j];
      x[i] = x[i] - lik * x[k];
    }
  }
  if (fabs(C[(n - 1) + n * (n - 1)]) < 1e-12) return 1;
  x[n - 1] = x[n - 1] / C[(n - 1) + n * (n - 1)];
  for (int k = n - 2; k >= 0; k--) {
    for (int j = k + 1; j < n; j++) x[k] = x[k] - C[k + n * j] * x[j];
    x[k] = x[k] / C[k + n * k];
  }
  return 0;
}

/* lu_solve_noswp, linear_solver_module.F90:413-445 (no pivoting) */
static inline void om_lu_solve_noswp(int n, const double *A, double *x, const double *b) {
  double C[64];
  for (int i = 0; i < n * n; i++) C[i] = A[i];
  for (int i = 0; i < n; i++) x[i] = b[i];
  for (int k = 0; k < n; k++) {
    double invdiag = 1.0 / C[k + n * k];
    for (int i = k + 1; i < n; i++) {
      double lik = C[i + n * k] * invdiag;
      for (int j = k; j < n; j++) C[i + n * j] = C[i + n * j] - lik * C[k + n * j];
      x[i] = x[i] - lik * x[k];
    }
  }
  x[n - 1] = x[n - 1] / C[(n - 1) + n * (n - 1)];
  for (int k = n - 2; k >= 0; k--) {
    for (int j = k + 1; j < n; j++) x[k] = x[k] - C[k + n * j] * x[j];
    x[k] = x[k] / C[k + n * k];
  }
}

#endif
