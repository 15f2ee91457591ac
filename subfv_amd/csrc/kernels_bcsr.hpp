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

// w = D^-1 w with the cached factors f, in lu_solve_noswp's operation order.
template <typename T, int B>
SFV_DEV void noswp_apply(const T *f, T *w) {
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
}

// x = D^-1 b with the cached factors, in lu_solve_noswp's operation order.
// TI: the type of b (float when it is a column of an FP32 Krylov basis)
template <typename T, int B, typename TI = double>
__global__ void k_bcsr_bjacobi(int m, const T *F, const TI *b, double *x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T *f = F + i * B * B;
  T w[B];
#pragma unroll
  for (int q = 0; q < B; q++) w[q] = (T)b[(long)B * i + q];
  noswp_apply<T, B>(f, w);
#pragma unroll
  for (int q = 0; q < B; q++) x[(long)B * i + q] = (double)w[q];
}

// w = w - blk x (one block of a sweep: the block's product summed over its
// columns from zero, then subtracted)
template <int B>
SFV_DEV void sub_block(const double *blk, const double *xb, double *w) {
#pragma unroll
  for (int r = 0; r < B; r++) {
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < B; c++) acc = acc + blk[r + B * c] * xb[c];
    w[r] = w[r] - acc;
  }
}

// One dependency level of bcsr_lusgs (sparse_bcsr_linear_module.F90:1268-1335):
// out_i = in_i - sum over j = row(i), ... while col(j) < i of A_ij out_col(j),
// then (second sweep, SOLVE) out_i = D_i^-1 out_i.  Rows of one level depend
// only on rows of earlier levels, so the level's rows run in parallel.
template <int B, bool SOLVE>
__global__ __launch_bounds__(256) void k_lusgs_level(int cnt, const int *rows, const int *row, const int *col,
                                                     const double *v, const double *F, const double *in,
                                                     double *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  double w[B];
#pragma unroll
  for (int r = 0; r < B; r++) w[r] = in[(long)B * i + r];
  for (int j = row[i]; j < row[i + 1] && col[j] < i; j++) sub_block<B>(v + (long)j * B * B, out + (long)B * col[j], w);
  if (SOLVE) noswp_apply<double, B>(F + (long)i * B * B, w);
#pragma unroll
  for (int r = 0; r < B; r++) out[(long)B * i + r] = w[r];
}

// One colour set of bcsr_parallel_lusgs (:1438-1502).  Forward (!BACK):
// out_i = in_i - sum over the row's blocks with color(col) < cs of A_ij out_col;
// backward: the blocks with color(col) > cs, then out_i = D_i^-1 out_i.
template <int B, bool BACK>
__global__ __launch_bounds__(256) void k_plusgs_set(int cnt, const int *rows, int cs, const int *color,
                                                    const int *row, const int *col, const double *v,
                                                    const double *F, const double *in, double *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  double w[B];
#pragma unroll
  for (int r = 0; r < B; r++) w[r] = in[(long)B * i + r];
  for (int j = row[i]; j < row[i + 1]; j++) {
    const int cc = color[col[j]];
    if (BACK ? (cc > cs) : (cc < cs)) sub_block<B>(v + (long)j * B * B, out + (long)B * col[j], w);
  }
  if (BACK) noswp_apply<double, B>(F + (long)i * B * B, w);
#pragma unroll
  for (int r = 0; r < B; r++) out[(long)B * i + r] = w[r];
}

// lu_solve (linear_solver_module.F90:447-484) of a 3 x 3 system: first
// largest pivot, rows always swapped, multipliers C(i,k) * (1/C(k,k)),
// back substitution over j increasing.  False where the reference stops
// (|C(3,3)| < 1e-12).
SFV_DEV bool lu_solve_plain3(const double *A, const double *b, double *x) {
  double C[9];
#pragma unroll
  for (int q = 0; q < 9; q++) C[q] = A[q];
#pragma unroll
  for (int i = 0; i < 3; i++) x[i] = b[i];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int p = k;
    double amax = fabs(C[k + 3 * k]);
#pragma unroll
    for (int i = k + 1; i < 3; i++)
      if (fabs(C[i + 3 * k]) > amax) {
        amax = fabs(C[i + 3 * k]);
        p = i;
      }
#pragma unroll
    for (int i = k + 1; i < 3; i++)
      if (i == p) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
          double t = C[k + 3 * j];
          C[k + 3 * j] = C[i + 3 * j];
          C[i + 3 * j] = t;
        }
        double t = x[k];
        x[k] = x[i];
        x[i] = t;
      }
    double invdiag = 1.0 / C[k + 3 * k];
#pragma unroll
    for (int i = k + 1; i < 3; i++) {
      double lik = C[i + 3 * k] * invdiag;
#pragma unroll
      for (int j = k; j < 3; j++) C[i + 3 * j] = C[i + 3 * j] - lik * C[k + 3 * j];
      x[i] = x[i] - lik * x[k];
    }
  }
  if (fabs(C[8]) < 1e-12) return false;
  x[2] = x[2] / C[8];
#pragma unroll
  for (int k = 1; k >= 0; k--) {
#pragma unroll
    for (int j = k + 1; j < 3; j++) x[k] = x[k] - C[k + 3 * j] * x[j];
    x[k] = x[k] / C[k + 3 * k];
  }
  return true;
}

// preconditioner of bcsr_pcgm (:1110-1169): lu_solve(3, v(:,:,d(i)), mr_i, r_i)
// on the first nine values of each diagonal block (the 3 x 3 dummy argument
// of the reference); the first three entries of mr_i are written.
__global__ __launch_bounds__(256) void k_pcg_prec(int m, int b, const int *diag_idx, const double *v,
                                                  const double *r, double *mr, int *err) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double x[3];
  if (!lu_solve_plain3(v + (long)diag_idx[i] * b * b, r + (long)b * i, x)) atomicOr(err, 1);
#pragma unroll
  for (int q = 0; q < 3; q++) mr[(long)b * i + q] = x[q];
}

// y = y + a * x   (a passed by value)
__global__ void k_axpy(long n, double a, const double *x, double *y) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = y[i] + a * x[i];
}

// p = mr + beta * p
__global__ void k_xpby(long n, const double *mr, double beta, double *p) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = mr[i] + beta * p[i];
}

// x = x + z[0] V[:,0] + z[1] V[:,1] + ... added one column at a time (bcsr_gmres)
__global__ void k_add_combination(long n, int k, const double *V, long ldv, const double *z, double *x) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double acc = x[e];
    for (int i = 0; i < k; i++) acc = acc + z[i] * V[(long)i * ldv + e];
    x[e] = acc;
  }
}

// the FP32 copy of the block values (FP32-inner GMRES)
__global__ void k_to_f32(long n, const double *v, float *v32) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v32[i] = (float)v[i];
}

// v(:,:,k) += M (add_block_to_bcsr), FP64 values and their FP32 copy
__global__ void k_add_block(int bb, long k, const double *M, double *v, float *v32) {
  const int q = threadIdx.x;
  if (q >= bb) return;
  const double nv = v[k * bb + q] + M[q];
  v[k * bb + q] = nv;
  v32[k * bb + q] = (float)nv;
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

// The sum of the previous dot's per-block partials, in the order of
// k_sum_finalize, computed by every block of the kernel that needs it (the
// kernel boundary orders the partials' writes before these reads); block 0
// stores it to *h_out.
SFV_DEV double block_total(const double *part, int nblk, double *h_out) {
  __shared__ double s[kVecThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kVecThreads) acc = acc + part[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kVecThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + o];
    __syncthreads();
  }
  const double t = s[0];
  if (h_out && blockIdx.x == 0 && threadIdx.x == 0) *h_out = t;
  __syncthreads();  // s is reused by the caller's reduction
  return t;
}

// One step of the Arnoldi MGS without a separate finalize: h = the sum of
// part_prev (the dot of the previous step, stored to *h_out), w = w - h *
// v_prev, then per-block partials of <v_next, w> (or <w, w> when v_next is
// null) into part_out (a different buffer from part_prev).
__global__ __launch_bounds__(kVecThreads) void k_mgs_chain(long n, double *w, const double *v_prev,
                                                           const double *part_prev, double *h_out,
                                                           const double *v_next, double *part_out) {
  const double hh = v_prev ? block_total(part_prev, (int)gridDim.x, h_out) : 0.0;
  double acc = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double x = w[i];
    if (v_prev) {
      x = x - hh * v_prev[i];
      w[i] = x;
    }
    acc = acc + (v_next ? v_next[i] * x : x * x);
  }
  block_sum_store(acc, part_out);
}

// ------------------------------------------------ classical Gram-Schmidt, twice
// The Arnoldi orthogonalisation of GMRES as CGS2: h1 = V^T w, w -= V h1,
// h2 = V^T w, w -= V h2, h = h1 + h2, then ||w||.  Three passes over V per
// step (the second fuses the first update with the second projection)
// instead of one pass per basis vector.  Partial sums are per block (fixed
// grid kGsBlocks, fixed reduction trees), so results are deterministic.
constexpr int kGsBlocks = 512;
constexpr int kGsMax = 32;  // basis vectors the CGS2 passes handle (restart <= 32)

// per-block sums of acc[c], c < k, into part[c * gridDim.x + blockIdx.x]
template <int KM>
SFV_DEV void gs_block_store(const double *acc, int k, double *part) {
  __shared__ double s[kVecThreads / 64][KM];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < KM; c++) {
    if (c < k) {
      double v = acc[c];
      for (int o = 32; o > 0; o >>= 1) v = v + __shfl_down(v, o);
      if (lane == 0) s[wv][c] = v;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < k) {
    double t = s[0][threadIdx.x];
    for (int q = 1; q < kVecThreads / 64; q++) t = t + s[q][threadIdx.x];
    part[(long)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// h[c] = sum over the blocks of part[c * nb + b], c < k (eight threads per
// column, each a fixed run of blocks, then the eight runs in order)
template <int KM>
SFV_DEV void gs_coeffs(const double *part, int k, int nb, double *h) {
  __shared__ double s2[KM][8];
  const int c = threadIdx.x >> 3, r = threadIdx.x & 7;
  if (c < k && c < KM) {
    const int per = (nb + 7) / 8, b0 = r * per, b1 = min(nb, b0 + per);
    double t = 0.0;
    for (int b = b0; b < b1; b++) t = t + part[(long)c * nb + b];
    s2[c][r] = t;
  }
  __syncthreads();
  if ((int)threadIdx.x < k) {
    double t = s2[threadIdx.x][0];
    for (int q = 1; q < 8; q++) t = t + s2[threadIdx.x][q];
    h[threadIdx.x] = t;
  }
  __syncthreads();
}

// partials of <V[:, c], w>, c < k, into part[c * gridDim.x + blockIdx.x];
// VT: the basis' storage type (float for an FP32 basis, summed in double)
template <int KM, typename VT = double>
__global__ __launch_bounds__(kVecThreads) void k_gs_dot(long n, int k, const VT *V, long ldv, const double *w,
                                                        double *part) {
  double acc[KM];
#pragma unroll
  for (int c = 0; c < KM; c++) acc[c] = 0.0;
  // columns c >= k (KM - k < 4 of them) re-read column k - 1, so every load
  // of an element is issued without a branch; their sums are not stored
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double vv[KM];
#pragma unroll
    for (int c = 0; c < KM; c++) vv[c] = (double)V[(long)(c < k ? c : k - 1) * ldv + i];
    const double wi = w[i];
#pragma unroll
    for (int c = 0; c < KM; c++) acc[c] = acc[c] + vv[c] * wi;
  }
  gs_block_store<KM>(acc, k, part);
}

// One update pass of CGS2.  h = the sums of part_in's k columns; w = w -
// sum_c h[c] V[:, c]; block 0 stores h to hcol (or adds it, accumulate).
// project: then partials of <V[:, c], w> into part_out (layout of k_gs_dot);
// otherwise partials of <w, w>, one per block.
template <int KM, typename VT = double>
__global__ __launch_bounds__(kVecThreads) void k_gs_update(long n, int k, const VT *V, long ldv, double *w,
                                                           const double *part_in, double *hcol, int accumulate,
                                                           int project, double *part_out) {
  __shared__ double h[KM];
  gs_coeffs<KM>(part_in, k, (int)gridDim.x, h);
  if (blockIdx.x == 0 && (int)threadIdx.x < k)
    hcol[threadIdx.x] = accumulate ? hcol[threadIdx.x] + h[threadIdx.x] : h[threadIdx.x];
  double acc[KM];
#pragma unroll
  for (int c = 0; c < KM; c++) acc[c] = 0.0;
  double nrm = 0.0;
  // as in k_gs_dot, columns c >= k re-read column k - 1 (branch-free loads)
  // and are skipped in the update
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double vv[KM];
#pragma unroll
    for (int c = 0; c < KM; c++) vv[c] = (double)V[(long)(c < k ? c : k - 1) * ldv + i];
    double x = w[i];
#pragma unroll
    for (int c = 0; c < KM; c++) x = (c < k) ? x - h[c] * vv[c] : x;
    w[i] = x;
    if (project) {
#pragma unroll
      for (int c = 0; c < KM; c++) acc[c] = acc[c] + vv[c] * x;
    } else {
      nrm = nrm + x * x;
    }
  }
  if (project)
    gs_block_store<KM>(acc, k, part_out);
  else
    block_sum_store(nrm, part_out);
}

// out = w / sqrt(sum of part) (the norm of the last MGS step; block 0 stores
// it to *h_out); out is w itself, or the next column of an FP32 basis
template <typename VT = double>
__global__ __launch_bounds__(kVecThreads) void k_scale_norm(long n, const double *w, const double *part, double *h_out,
                                                            VT *out) {
  const double d = sqrt(block_total(part, (int)gridDim.x, nullptr));
  if (blockIdx.x == 0 && threadIdx.x == 0) *h_out = d;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (VT)(w[i] / d);
}

// out = a / *s   (normalisation by a device scalar, sqrt taken if take_sqrt)
template <typename VT = double>
__global__ void k_scale_div(long n, const double *a, VT *out, const double *s, int take_sqrt, double *s_out) {
  double d = take_sqrt ? sqrt(*s) : *s;
  if (blockIdx.x == 0 && threadIdx.x == 0 && s_out) *s_out = d;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (VT)(a[i] / d);
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
template <typename VT = double>
__global__ void k_combine(long n, int k, const VT *V, long ldv, const double *z, double *r) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int i = 0; i < k; i++) acc = acc + z[i] * (double)V[(long)i * ldv + e];
    r[e] = acc;
  }
}

}  // namespace sfv
