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

// The right-preconditioned matrix of block-Jacobi GMRES, block by block:
// B(q) = A(q) D_j^-1 for block q in column j (0-based col), with D_j^-1
// formed column by column from the cached factors; FP64 and its FP32
// rounding.  A (M^-1 u) becomes one product B u.
template <int B>
__global__ void k_bcsr_rprec(int nnzb, const int *col, const double *v, const double *F64, double *out,
                             float *out32) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnzb) return;
  const double *f = F64 + (long)col[q] * B * B;
  double di[B * B];
#pragma unroll
  for (int k = 0; k < B; k++) {
    double w[B];
#pragma unroll
    for (int r = 0; r < B; r++) w[r] = r == k ? 1.0 : 0.0;
    noswp_apply<double, B>(f, w);
#pragma unroll
    for (int r = 0; r < B; r++) di[r + B * k] = w[r];
  }
  const double *a = v + q * B * B;
#pragma unroll
  for (int r = 0; r < B; r++)
#pragma unroll
    for (int c = 0; c < B; c++) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < B; k++) t = t + a[r + B * k] * di[k + B * c];
      out[q * B * B + r + B * c] = t;
      out32[q * B * B + r + B * c] = (float)t;
    }
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

// ------------------------------------------------ Gram-Schmidt in block passes
// Partial sums are per block (fixed grid kGsBlocks, fixed reduction trees),
// so results are deterministic.
constexpr int kGsBlocks = 512;
constexpr int kGsMax = 32;  // basis vectors the block passes handle (restart <= 32)
// the DCGS2 scalars of a GMRES handle: the two provisional Hessenberg
// columns (kGsMax + 2 each), beta of the cycle's first column, then the
// step's coefficients (a, b, beta, e: 2 kGsMax + 2)
constexpr int kDchBeta0 = 2 * (kGsMax + 2), kDchCoef = kDchBeta0 + 1, kDchSize = kDchCoef + 2 * kGsMax + 2;

// ---- Arnoldi with delayed reorthogonalisation (DCGS2: Swirydowicz, Langou,
// Ananthan, Yang and Thomas, "Low synchronization Gram-Schmidt and GMRES
// algorithms", Numer. Linear Algebra Appl. 28 (2021); Bielich et al.,
// Parallel Computing 112 (2022)).  Step s holds the orthonormal columns
// V(:, 0..s-1) and a provisional next vector u, already orthogonalised once,
// with A M^-1 V(:, s-1) = V(:, 0..s-1) hprov + u.  With z = A M^-1 u, one
// pass forms a = V^T u, b = V^T z, c = u.u and d = u.z; then
//   beta = sqrt(c - a.a),  V(:, s) = (u - V a) / beta,
//   H(0..s-1, s-1) = hprov + a,  H(s, s-1) = beta          (u's second pass)
//   e = (d - a.b) / beta,  Ha = H(0..s, 0..s-1) a,
//   hprov' = ((b - Ha(0..s-1)) / beta, (e - Ha(s)) / beta)    (the next column, first pass)
//   u' = (z - V b - V(:, s) e) / beta
// because A M^-1 V(:, s) = (z - A M^-1 V a) / beta = (z - V H a) / beta.  The
// orthogonalisation reads the basis twice per step instead of three times.

// The step's coefficients from the partials of k_dc_dot (kGsBlocks blocks),
// one block of kDcCoefThreads: co[0..s) = a, co[s..2s) = b, co[2s] = c,
// co[2s+1] = d; then beta, e, Ha; H's corrected column s-1, the new
// provisional column hprov_out (hprov_in is only read, so the two
// alternate), and coef = (a, b, beta, e) for the update pass.  beta_out
// (step 0): beta, the norm of the cycle's first vector.
constexpr int kDcCoefThreads = 1024, kDcRuns = 16, kDcPer = kGsBlocks / kDcRuns;
static_assert(kGsBlocks % kDcRuns == 0, "partials split evenly into runs");

// sum over the 2^L lanes of each aligned group (xor butterfly, fixed order)
template <int L>
SFV_DEV double group_sum(double v) {
#pragma unroll
  for (int o = 1; o < (1 << L); o <<= 1) v = v + __shfl_xor(v, o);
  return v;
}

template <int KM>
__global__ __launch_bounds__(kDcCoefThreads) void k_dc_coef(int s, const double *part, double *H, int ldh,
                                                           const double *hprov_in, double *hprov_out, double *coef,
                                                           double *beta_out) {
  __shared__ double s2[2 * KM + 2][kDcRuns], co[2 * KM + 2], ha[KM + 1], sc[2];
  __shared__ double hs[KM + 1][KM];  // H(0..s, 0..s-1), upper Hessenberg, column s-1 corrected below
  const int ncol = 2 * s + 2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // H's final columns 0..s-2 into LDS while the partials are summed
  for (int q = threadIdx.x; q < (s + 1) * s; q += kDcCoefThreads) {
    const int i = q % (s + 1), k = q / (s + 1);
    if (k < s - 1) hs[i][k] = i <= k + 1 ? H[i + (long)ldh * k] : 0.0;
  }
  // kDcRuns threads per partial column, each a fixed run of kDcPer blocks
  // (loaded together, then added in order), then the runs in order
  for (int c0 = 0; c0 < ncol; c0 += kDcCoefThreads / kDcRuns) {
    const int c = c0 + (int)(threadIdx.x / kDcRuns), r = threadIdx.x % kDcRuns;
    if (c < ncol) {
      const double *pp = part + (long)c * kGsBlocks + r * kDcPer;
      double v[kDcPer];
#pragma unroll
      for (int b = 0; b < kDcPer; b++) v[b] = pp[b];
      double t = 0.0;
#pragma unroll
      for (int b = 0; b < kDcPer; b++) t = t + v[b];
      s2[c][r] = t;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < ncol) {
    double t = s2[threadIdx.x][0];
    for (int r = 1; r < kDcRuns; r++) t = t + s2[threadIdx.x][r];
    co[threadIdx.x] = t;
  }
  __syncthreads();
  if (wv == 0) {  // a.a and a.b over the first wave, then beta and e
    const double ak = lane < s ? co[lane] : 0.0;
    const double aa = group_sum<6>(ak * ak), ab = group_sum<6>(lane < s ? ak * co[s + lane] : 0.0);
    if (lane == 0) {
      const double beta = sqrt(fmax(co[2 * s] - aa, 0.0));
      sc[0] = beta;
      sc[1] = (co[2 * s + 1] - ab) / beta;
    }
  }
  __syncthreads();
  const double beta = sc[0];
  // Ha(i) = sum over k < s of H(i, k) a(k), i <= s, with column s-1 the
  // corrected one (hprov_in + a, beta below it); H is upper Hessenberg.
  // One 32-lane half-wave per row i, lane k the product of column k.
  if (s >= 1 && (int)threadIdx.x <= s) {
    const int i = threadIdx.x;
    hs[i][s - 1] = i < s ? hprov_in[i] + co[i] : beta;
  }
  __syncthreads();
  for (int i = threadIdx.x >> 5; i <= s; i += kDcCoefThreads / 32) {  // uniform over each half-wave
    const int k = threadIdx.x & 31;
    const double t = group_sum<5>(k < s && k >= i - 1 ? hs[i][k] * co[k] : 0.0);
    if (k == 0) ha[i] = t;
  }
  __syncthreads();
  if ((int)threadIdx.x <= s) {
    const int i = threadIdx.x;
    if (s >= 1) H[i + (long)ldh * (s - 1)] = hs[i][s - 1];
    hprov_out[i] = i < s ? (co[s + i] - ha[i]) / beta : (sc[1] - ha[s]) / beta;
  }
  if ((int)threadIdx.x < 2 * s) coef[threadIdx.x] = co[threadIdx.x];
  if (threadIdx.x == 0) {
    coef[2 * s] = beta;
    coef[2 * s + 1] = sc[1];
    if (beta_out) *beta_out = beta;
  }
}

// Wave sums of the NV values of every lane as a butterfly that halves the
// array at each exchange (reduce-scatter): 63 exchanges for 64 values
// instead of 6 per value.  NV is a power of two >= 64; on return lane L holds
// the wave sums of values L * (NV / 64) .. L * (NV / 64) + NV / 64 - 1 in
// v[0 .. NV / 64).  One exchange step per instantiation, so every index is
// a compile-time constant and v stays in registers.
template <int H, int BIT>
SFV_DEV void wave_reduce_scatter_step(double *v, int lane) {
  const bool up = (lane & BIT) != 0;
#pragma unroll
  for (int c = 0; c < H; c++) {
    const double send = up ? v[c] : v[c + H];
    const double keep = up ? v[c + H] : v[c];
    v[c] = keep + __shfl_xor(send, BIT);
  }
  if constexpr (BIT > 1) wave_reduce_scatter_step<H / 2, BIT / 2>(v, lane);
}
template <int NV>
SFV_DEV void wave_reduce_scatter(double *v, int lane) {
  wave_reduce_scatter_step<NV / 2, 32>(v, lane);
}

constexpr int dc_pow2(int x) { return x <= 64 ? 64 : 128; }

// two consecutive basis entries (ldv and the element index even)
template <typename VT> struct Pair;
template <> struct Pair<double> { using T = double2; };
template <> struct Pair<float> { using T = float2; };

// The step's dot pass: per-block partials of a (columns 0..s-1), b
// (s..2s-1), c (2s) and d (2s+1) into part[q * gridDim.x + blockIdx.x].  Each
// thread takes two consecutive elements per iteration (ldv even); an odd n's
// last element goes to thread 0.
template <int KM, typename VT>
__global__ __launch_bounds__(kVecThreads) void k_dc_dot(long n, int s, const VT *V, long ldv, const double *u,
                                                       const double *z, double *part) {
  using V2 = typename Pair<VT>::T;
  // acc: a at [0, KM), b at [KM, 2 KM), c at 2 KM, d at 2 KM + 1, then zeros
  constexpr int NV = dc_pow2(2 * KM + 2);
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; c++) acc[c] = 0.0;
  // columns c >= s (fewer than 4 of them, or all when s = 0) re-read column
  // max(s - 1, 0), so every load is issued without a branch; their sums are
  // not stored
  const int last = s > 0 ? s - 1 : 0;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
  auto add = [&](double vc, int c, double ui, double zi) {
    acc[c] = acc[c] + vc * ui;
    acc[KM + c] = acc[KM + c] + vc * zi;
  };
  for (long i2 = tid; i2 < (n >> 1); i2 += stride) {
    const long i = 2 * i2;
    V2 vv[KM];
#pragma unroll
    for (int c = 0; c < KM; c++) vv[c] = *reinterpret_cast<const V2 *>(V + (long)(c < s ? c : last) * ldv + i);
    const double2 uu = *reinterpret_cast<const double2 *>(u + i), zz = *reinterpret_cast<const double2 *>(z + i);
#pragma unroll
    for (int c = 0; c < KM; c++) {
      add((double)vv[c].x, c, uu.x, zz.x);
      add((double)vv[c].y, c, uu.y, zz.y);
    }
    acc[2 * KM] = (acc[2 * KM] + uu.x * uu.x) + uu.y * uu.y;
    acc[2 * KM + 1] = (acc[2 * KM + 1] + uu.x * zz.x) + uu.y * zz.y;
  }
  if ((n & 1) && tid == 0) {
    const long i = n - 1;
    const double ui = u[i], zi = z[i];
#pragma unroll
    for (int c = 0; c < KM; c++) add((double)V[(long)(c < s ? c : last) * ldv + i], c, ui, zi);
    acc[2 * KM] = acc[2 * KM] + ui * ui;
    acc[2 * KM + 1] = acc[2 * KM + 1] + ui * zi;
  }
  __shared__ double sh[kVecThreads / 64][NV];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  wave_reduce_scatter<NV>(acc, lane);
#pragma unroll
  for (int c = 0; c < NV / 64; c++) sh[wv][lane * (NV / 64) + c] = acc[c];
  __syncthreads();
  if ((int)threadIdx.x < 2 * s + 2) {
    const int t = threadIdx.x;
    const int src = t < s ? t : (t < 2 * s ? KM + (t - s) : 2 * KM + (t - 2 * s));
    double v = sh[0][src];
    for (int q = 1; q < kVecThreads / 64; q++) v = v + sh[q][src];
    part[(long)t * gridDim.x + blockIdx.x] = v;
  }
}

// the step's update pass: V(:, s) = (u - V a) / beta; u = (z - V b - V(:, s) e) / beta,
// two consecutive elements per thread and iteration as in k_dc_dot
template <int KM, typename VT>
__global__ __launch_bounds__(kVecThreads) void k_dc_update(long n, int s, VT *V, long ldv, double *u, const double *z,
                                                          const double *coef) {
  using V2 = typename Pair<VT>::T;
  __shared__ double co[2 * KM + 2];
  if ((int)threadIdx.x < 2 * s + 2) co[threadIdx.x] = coef[threadIdx.x];
  __syncthreads();
  const double beta = co[2 * s], e = co[2 * s + 1];
  const int last = s > 0 ? s - 1 : 0;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
  for (long i2 = tid; i2 < (n >> 1); i2 += stride) {
    const long i = 2 * i2;
    V2 vv[KM];
#pragma unroll
    for (int c = 0; c < KM; c++) vv[c] = *reinterpret_cast<const V2 *>(V + (long)(c < s ? c : last) * ldv + i);
    const double2 uu = *reinterpret_cast<const double2 *>(u + i), zz = *reinterpret_cast<const double2 *>(z + i);
    double q0 = uu.x, q1 = uu.y, x0 = zz.x, x1 = zz.y;
#pragma unroll
    for (int c = 0; c < KM; c++)
      if (c < s) {
        q0 = q0 - co[c] * (double)vv[c].x;
        q1 = q1 - co[c] * (double)vv[c].y;
        x0 = x0 - co[s + c] * (double)vv[c].x;
        x1 = x1 - co[s + c] * (double)vv[c].y;
      }
    q0 = q0 / beta;
    q1 = q1 / beta;
    V2 qv;
    qv.x = (VT)q0;
    qv.y = (VT)q1;
    *reinterpret_cast<V2 *>(V + (long)s * ldv + i) = qv;
    double2 un;
    un.x = (x0 - q0 * e) / beta;
    un.y = (x1 - q1 * e) / beta;
    *reinterpret_cast<double2 *>(u + i) = un;
  }
  if ((n & 1) && tid == 0) {
    const long i = n - 1;
    double q = u[i], x = z[i];
#pragma unroll
    for (int c = 0; c < KM; c++)
      if (c < s) {
        const double vc = (double)V[(long)c * ldv + i];
        q = q - co[c] * vc;
        x = x - co[s + c] * vc;
      }
    q = q / beta;
    V[(long)s * ldv + i] = (VT)q;
    u[i] = (x - q * e) / beta;
  }
}

// w = w / sqrt(sum of part) (the norm of the last MGS step; block 0 stores
// it to *h_out)
__global__ __launch_bounds__(kVecThreads) void k_scale_norm(long n, double *w, const double *part, double *h_out) {
  const double d = sqrt(block_total(part, (int)gridDim.x, nullptr));
  if (blockIdx.x == 0 && threadIdx.x == 0) *h_out = d;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    w[i] = w[i] / d;
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
template <typename VT = double>
__global__ void k_combine(long n, int k, const VT *V, long ldv, const double *z, double *r) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int i = 0; i < k; i++) acc = acc + z[i] * (double)V[(long)i * ldv + e];
    r[e] = acc;
  }
}

}  // namespace sfv
