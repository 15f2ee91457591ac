// BCSR kernels for the implicit path: SpMV (bcsr_matmul), block-Jacobi
// (bcsr_diag with lu_solve_noswp) and the vector kernels of GMRES
// (src/linear_algebra/sparse_bcsr_linear_module.F90:13-29, 1414-1436,
// 1568-1691; src/linear_algebra/linear_solver_module.F90:413-445).
//
// Blocks keep the reference's column-major b x b layout; lane r of a group
// of b lanes owns row r of its block row.  Reductions are two-level
// (per-block partials, then the last block), so results are deterministic.
#pragma once
#include "physics.hpp"

namespace sfv {

constexpr int kVecThreads = 256;

// y = A x.  T = storage/arithmetic type of A and of the accumulation.
template <typename T, int B>
__global__ __launch_bounds__(256) void k_bcsr_spmv(int m, const int *row, const int *col, const T *v,
                                                   const double *x, double *y) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long i = t / B;
  int r = (int)(t - i * B);
  if (i >= m) return;
  T acc_row = (T)0;
  for (int j = row[i]; j < row[i + 1]; j++) {
    const T *blk = v + (long)j * B * B;
    const double *xb = x + (long)B * col[j];
    T acc = (T)0;
#pragma unroll
    for (int c = 0; c < B; c++) acc = acc + blk[r + B * c] * (T)xb[c];
    acc_row = acc_row + acc;
  }
  y[(long)B * i + r] = (double)acc_row;
}

// Factors of lu_solve_noswp for every diagonal block, computed once:
// F[i] holds U in its upper triangle and the multipliers lik below it.
template <typename T, int B>
__global__ void k_bcsr_factor(int m, const int *diag_idx, const double *v, T *F) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double Cm[B * B];
  const double *A = v + (long)diag_idx[i] * B * B;
#pragma unroll
  for (int q = 0; q < B * B; q++) Cm[q] = A[q];
  double L[B * B];
#pragma unroll
  for (int q = 0; q < B * B; q++) L[q] = 0.0;
#pragma unroll
  for (int k = 0; k < B; k++) {
    double invdiag = 1.0 / Cm[k + B * k];
#pragma unroll
    for (int ii = k + 1; ii < B; ii++) {
      double lik = Cm[ii + B * k] * invdiag;
      L[ii + B * k] = lik;
#pragma unroll
      for (int j = k; j < B; j++) Cm[ii + B * j] = Cm[ii + B * j] - lik * Cm[k + B * j];
    }
  }
  T *o = F + i * B * B;
#pragma unroll
  for (int ii = 0; ii < B; ii++)
#pragma unroll
    for (int j = 0; j < B; j++) o[ii + B * j] = (T)((j >= ii) ? Cm[ii + B * j] : L[ii + B * j]);
}

// x = D^-1 b with the cached factors, in lu_solve_noswp's operation order.
template <typename T, int B>
__global__ void k_bcsr_bjacobi(int m, const T *F, const double *b, double *x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T *f = F + i * B * B;
  T w[B];
#pragma unroll
  for (int q = 0; q < B; q++) w[q] = (T)b[(long)B * i + q];
#pragma unroll
  for (int k = 0; k < B; k++)
#pragma unroll
    for (int ii = k + 1; ii < B; ii++) w[ii] = w[ii] - f[ii + B * k] * w[k];
  w[B - 1] = w[B - 1] / f[(B - 1) + B * (B - 1)];
#pragma unroll
  for (int k = B - 2; k >= 0; k--) {
#pragma unroll
    for (int j = k + 1; j < B; j++) w[k] = w[k] - f[k + B * j] * w[j];
    w[k] = w[k] / f[k + B * k];
  }
#pragma unroll
  for (int q = 0; q < B; q++) x[(long)B * i + q] = (double)w[q];
}

// -------------------------------------------------------- vector reductions
// Sums in two launches: per-block partials (fixed grid, fixed order), then
// one block adds them in block order; deterministic for a given grid.
SFV_DEV void block_sum_store(double v, double *part) {
  __shared__ double s[kVecThreads];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = kVecThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kVecThreads) void k_sum_finalize(const double *part, int nblk, double *out) {
  __shared__ double s[kVecThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kVecThreads) acc = acc + part[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kVecThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// w = w - (*h) * v_prev (if v_prev), then partials of <v_next, w> (or of
// <w, w> when v_next is null).  One pass fuses an MGS update with the next
// dot product.
__global__ __launch_bounds__(kVecThreads) void k_mgs_step(long n, double *w, const double *v_prev,
                                                          const double *h, const double *v_next, double *part) {
  double acc = 0.0;
  const double hh = v_prev ? *h : 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double x = w[i];
    if (v_prev) {
      x = x - hh * v_prev[i];
      w[i] = x;
    }
    acc = acc + (v_next ? v_next[i] * x : x * x);
  }
  block_sum_store(acc, part);
}

// out = a / *s   (normalisation by a device scalar, sqrt taken if take_sqrt)
__global__ void k_scale_div(long n, const double *a, double *out, const double *s, int take_sqrt, double *s_out) {
  double d = take_sqrt ? sqrt(*s) : *s;
  if (blockIdx.x == 0 && threadIdx.x == 0 && s_out) *s_out = d;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = a[i] / d;
}

// r = b - y
__global__ void k_sub(long n, const double *b, const double *y, double *r) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    r[i] = b[i] - y[i];
}

// x = x + y
__global__ void k_add(long n, double *x, const double *y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = x[i] + y[i];
}

// r = sum_{i < k} z[i] * V[:, i]  (sequential in i for each entry)
__global__ void k_combine(long n, int k, const double *V, long ldv, const double *z, double *r) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int i = 0; i < k; i++) acc = acc + z[i] * V[(long)i * ldv + e];
    r[e] = acc;
  }
}

}  // namespace sfv
