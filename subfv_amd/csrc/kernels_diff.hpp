// Viscous terms of compute_rhs_ns (src/ns/ns_main.F90:170-175) with a
// constant viscosity: the nodal momentum system of
// ns_vectorial_diffusion_module.F90:7-320 and the nodal heat system of
// ns_diffusion_module.F90:7-206.
//
// With mu constant in space and time (use_sutherland = .false.) the nodal
// matrices N, S, S~^T and B depend on the mesh only, so the setup kernel
// forms, once per mesh, every product the residual needs.  With Sutherland's
// law (compute_mu, ns_vectorial_diffusion_module.F90:322-359) mu varies per
// cell and step: the same setup kernel then runs before every residual, on
// the cell viscosities and their minimum mu0corr (one wavefront per
// vertex, matrices in LDS, or in a global workspace for vertices beyond 16
// sub-faces or sub-elements, the inverse in the order of LU + unblocked
// dgetri, products summed over k in increasing order):
//   momentum: SNS = (S~^T N^-1) S, SB = S~^T (N^-1 B), the sub-element sums of
//             S~^T, and N^-1 S, N^-1 B (the face velocities Vf = N^-1 (S Ve + B)
//             are formed as N^-1 S Ve + N^-1 B);
//   heat:     SNS_h, SB_h and the row sums of S~^T_h.
// Each step, one wavefront per vertex evaluates the reference's scatter
// terms from the cell states and writes one row of 18 partial sums per
// node-sorted sub-element slot; k_diff_gather adds a cell's rows in
// increasing vertex id into rdiff[C][4] (rhs of rho u, rho v, rho w, E) and
// diag[C][25] (the point-implicit block).
#pragma once
#include "device_mesh.hpp"
#include "kernels_flux.hpp"
#include "physics.hpp"

namespace sfv {

// Two size classes of vertex: up to kDiffMaxSub sub-faces and sub-elements
// (prism, hex and quad meshes; the setup's matrices in LDS, four vertices per
// apply workgroup, the apply in a compact LDS layout when no vertex has more
// than kDiffMaxSubC sub-elements) and up to kDiffMaxSubL (tetrahedral meshes;
// the setup's matrices in a global workspace, one vertex per apply workgroup).
constexpr int kDiffMaxSub = 16;             // sub-faces (and sub-elements) per vertex, small class
constexpr int kDiffMaxF = 3 * kDiffMaxSub;  // 48
constexpr int kDiffMaxSubL = 64;            // large class
constexpr int kDiffSlot = 18;               // partials per slot

struct DiffParams {
  double mu, mu0, kap, cv;
  double pr;                  // Prandtl number
  double suth_mu0, suth_T0;   // Sutherland reference viscosity and temperature
  int suth, pad_;             // use_sutherland: mu per cell from mu_cell, mu0corr from mu_min
  int bc_V_id[11], bc_T_id[11];  // 1 neumann, 2 dirichlet, by BC index (1-based)
  double bc_val_V[11][3], bc_val_T[11];
};

// packed per-vertex block: sizes for nf sub-faces and ns sub-elements
SFV_DEV __host__ long diff_block_size(int nf, int ns) {
  const long F = 3 * nf, E = 3 * ns;
  return E * E + E + 9L * ns + F * E + F + (long)ns * ns + 2L * ns;
}

// The setup's four matrices live in dynamic LDS sized for the mesh's largest
// vertex (fm = 3 * max sub-faces, em = 3 * max sub-elements): N fm x fm, then
// S (F x E), S~^T (E x F) and the products W (E x F), fm * em each, all stored
// with the vertex's own leading dimensions.  Prism meshes need 26 KB instead
// of 74 KB for the 16-sub-face limit, so several vertices share a CU.
SFV_DEV __host__ size_t diff_setup_lds_bytes(int fm, int em) {
  return sizeof(double) * ((size_t)fm * fm + 3ull * fm * em);
}
// the large class keeps the same four matrices in a global workspace, one
// stride of doubles per setup block, kDiffWsBlocks blocks per launch
SFV_DEV __host__ long diff_setup_ws_stride(int fm, int em) { return (long)fm * fm + 3L * fm * em; }
constexpr int kDiffWsBlocks = 2048;

template <int MS>
struct DiffSetupLds {
  double B[3 * MS], x[3 * MS], work[3 * MS];
  int piv[3 * MS];
  int lse[MS], rse[MS];                    // local slot of each sub-face's left/right cell (-1 none)
  // the vertex's mesh data, staged once so the assembly loops read LDS only
  double g[MS][4];                         // unit normal and area of each sub-face
  int le[MS], re[MS];                      // its left and right neighbour codes
  double sevol[MS], semu[MS];              // volume and cell viscosity of each sub-element
  int secell[MS];
  int ssf[MS][kSlotSrc + 1];               // local sub-faces of each sub-element, count first
};

struct DiffArgs {
  DevMesh m;
  DiffParams p;
  const long *doff;   // [V+1] offsets of the packed vertex blocks
  double *pre;        // packed blocks
  const double *U;    // [C][5]
  double *PD;         // [SE][18] node-sorted slot partials
  double *rdiff;      // [C][4]
  double *diag;       // [C][25]
  const double *mu_cell;  // [C] viscosity of each cell (Sutherland), else nullptr
  int set_fm, set_em;     // setup: largest 3 * sub-faces and 3 * sub-elements of a vertex
  const double *mu_min;   // mu0corr, their minimum (Sutherland)
  int vper_xcd, vblocks;  // XCD slices of the vertex blocks (apply)
  int cper_xcd, cblocks;  // and of the cell blocks (gather)
  int parts;              // apply: 1 momentum system, 2 heat system, 3 both
  double *ws;             // setup of the large class: the matrices of each block, ws_stride doubles apart
  long ws_stride;
  int v0;                 // setup: first vertex of the launch
  const uint32_t *slot_ssf;  // [SE] node-sorted slot: its sub-element's vertex-local sub-faces, a byte each, 0xFF ends
  const double *slot_vol;    // [SE] node-sorted slot: its sub-element's volume
};

// slot_vol of every node-sorted slot from the cell-major sub-element volumes
__global__ __launch_bounds__(256) void k_slot_vol(long n, const int *v_seid, const double *se_vol, double *out) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = se_vol[v_seid[q]];
}

SFV_DEV void wsync_d() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A := A^-1 (n x n column-major in LDS, ld n): LU with partial pivoting in
// lu_solve's order, then inv(U) (dtrti2) and inv(A) L = inv(U) (dgetri).
SFV_DEV void wave_inverse(double *A, int n, int *piv, double *work, int lane) {
#define AA(i, j) A[(i) + n * (j)]
  for (int k = 0; k < n; k++) {
    // pivot: the first row of the largest |A(i, k)|, i >= k, as the serial
    // scan with a strict > finds it (NaN never replaces the running maximum,
    // so a NaN diagonal keeps p = k): each lane scans rows k + lane + 64 t in
    // increasing order, then a butterfly over the wave, ties to the lower row
    int p;
    {
      double v = -1.0;
      int idx = k + lane;
      for (int i = k + lane; i < n; i += 64) {
        double a = fabs(AA(i, k));
        if (!(a == a)) a = -1.0;
        if (a > v) {
          v = a;
          idx = i;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(idx, off);
        if (ov > v || (ov == v && oi < idx)) {
          v = ov;
          idx = oi;
        }
      }
      const double d0 = fabs(AA(k, k));
      p = (d0 == d0) ? idx : k;
    }
    if (lane == 0) piv[k] = p;
    if (AA(p, k) != 0.0) {
      if (p != k)
        for (int j = lane; j < n; j += 64) {
          double t = AA(k, j);
          AA(k, j) = AA(p, j);
          AA(p, j) = t;
        }
      wsync_d();
      const double pv = AA(k, k);
      if (fabs(pv) >= 2.2250738585072014e-308) {
        const double r = 1.0 / pv;
        for (int i = k + 1 + lane; i < n; i += 64) AA(i, k) = AA(i, k) * r;
      } else {
        for (int i = k + 1 + lane; i < n; i += 64) AA(i, k) = AA(i, k) / pv;
      }
    }
    wsync_d();
    const int m = n - k - 1;
    for (int q = lane; q < m * m; q += 64) {
      const int i = k + 1 + q % m, j = k + 1 + q / m;
      AA(i, j) = AA(i, j) - AA(i, k) * AA(k, j);
    }
    wsync_d();
  }
  // dtrti2 (upper, non-unit).  Its dtrmv reads x(jj) = A(jj, j) at step jj
  // before any later step changes it (step jj' only updates rows above jj'),
  // so every row of the column is one independent sum in dtrmv's order:
  // x(i) A(i, i) (when x(i) != 0), then + x(jj) A(i, jj) for jj = i+1..j-1
  // with x(jj) != 0; one lane per row, one synchronisation per column.
  for (int j = 0; j < n; j++) {
    for (int i = lane; i < j; i += 64) work[i] = AA(i, j);
    if (lane == 0) AA(j, j) = 1.0 / AA(j, j);
    wsync_d();
    const double ajj = -AA(j, j);
    for (int i = lane; i < j; i += 64) {
      const double xi = work[i];
      double acc = xi != 0.0 ? xi * AA(i, i) : xi;
      for (int jj = i + 1; jj < j; jj++) {
        const double t = work[jj];
        if (t != 0.0) acc = acc + t * AA(i, jj);
      }
      AA(i, j) = ajj * acc;
    }
    wsync_d();
  }
  // inv(A) L = inv(U), columns right to left
  for (int j = n - 1; j >= 0; j--) {
    for (int i = j + 1 + lane; i < n; i += 64) {
      work[i] = AA(i, j);
      AA(i, j) = 0.0;
    }
    wsync_d();
    for (int i = lane; i < n; i += 64) {
      double acc = AA(i, j);
      for (int jj = j + 1; jj < n; jj++) acc = acc + (-1.0 * work[jj]) * AA(i, jj);
      AA(i, j) = acc;
    }
    wsync_d();
  }
  for (int j = n - 2; j >= 0; j--) {
    const int jp = piv[j];
    if (jp != j)
      for (int i = lane; i < n; i += 64) {
        double t = AA(i, j);
        AA(i, j) = AA(i, jp);
        AA(i, jp) = t;
      }
    wsync_d();
  }
#undef AA
}

// C(m,n) = A(m,k) B(k,n), each entry summed over k in order from zero
SFV_DEV void wave_gemm(int m, int n, int k, const double *A, const double *B, double *C, int lane) {
  for (int q = lane; q < m * n; q += 64) {
    const int i = q % m, j = q / m;
    double acc = 0.0;
    for (int l = 0; l < k; l++) acc = acc + B[l + k * j] * A[i + m * l];
    C[i + m * j] = acc;
  }
}

// y(m) = A(m,n) x(n)
SFV_DEV void wave_gemv(int m, int n, const double *A, const double *x, double *y, int lane) {
  for (int i = lane; i < m; i += 64) {
    double acc = 0.0;
    for (int j = 0; j < n; j++) acc = acc + x[j] * A[i + m * j];
    y[i] = acc;
  }
}

SFV_DEV void sf_normal(const DevMesh &m, long s, double sign, double *n) {
  n[0] = sign * m.sfn[s];
  n[1] = sign * m.sfn[m.SF + s];
  n[2] = sign * m.sfn[2 * m.SF + s];
}

// blc_3 of the constant-viscosity form (ns_vectorial_diffusion_module.F90:92-96),
// entry (i, j): s (mu (nl.np) d_ij + (mu - mu0) np_i nl_j + (mu0 - 2/3 mu) np_j nl_i)
SFV_DEV double blc_mu_ij(double s, double mu, double mu0, const double *nl, const double *np, int i, int j) {
  const double d = nl[0] * np[0] + nl[1] * np[1] + nl[2] * np[2];
  return s * (mu * d * (i == j ? 1.0 : 0.0) + (mu - mu0) * (np[i] * nl[j]) + (mu0 - (2.0 / 3.0) * mu) * (np[j] * nl[i]));
}

// One 64-thread workgroup per vertex (vertex v0 + blockIdx.x).  MS: the size
// class; GWS: the four matrices in the global workspace (else dynamic LDS).
template <int MS, bool GWS>
__global__ __launch_bounds__(64) void k_diff_setup(DiffArgs a) {
  __shared__ DiffSetupLds<MS> L;
  extern __shared__ double dyn_lds[];
  const int fm = a.set_fm, em = a.set_em;
  double *const Nm = GWS ? a.ws + a.ws_stride * blockIdx.x : dyn_lds;
  double *const Sm = Nm + fm * fm, *const Tm = Sm + fm * em, *const Wm = Tm + fm * em;
  const DevMesh &m = a.m;
  const DiffParams &P = a.p;
  const long v = (long)a.v0 + blockIdx.x;
  if (v >= m.V) return;
  const int lane = threadIdx.x;
  const int sf0 = m.sf_off[v], nf = m.sf_off[v + 1] - sf0;
  const int se0 = m.se_off[v], ns = m.se_off[v + 1] - se0;
  const int F = 3 * nf, E = 3 * ns;
  double *out = a.pre + a.doff[v];
  for (int l = lane; l < nf; l += 64) {
    const long s = sf0 + l;
    for (int d = 0; d < 4; d++) L.g[l][d] = m.sfn[d * m.SF + s];
    L.le[l] = m.sf_le[s];
    L.re[l] = m.sf_re[s];
  }
  for (int q = lane; q < ns; q += 64) {
    const long se = m.v_seid[se0 + q];
    const int cell = m.se_cell[se0 + q];
    L.secell[q] = cell;
    L.sevol[q] = m.se_vol[se];
    L.semu[q] = a.mu_cell ? a.mu_cell[cell] : P.mu;
    const int pb = m.se_sfoff[se], pe = m.se_sfoff[se + 1];
    L.ssf[q][0] = pe - pb;  // at most kSlotSrc (checked when the groups are built)
    for (int pp = pb; pp < pe; pp++) L.ssf[q][1 + pp - pb] = (int)(m.se_sf[pp] - sf0);
  }
  wsync_d();
  // local slots of the left / right cells of each sub-face
  for (int l = lane; l < nf; l += 64) {
    const int le = L.le[l], re = L.re[l];
    int ls = -1, rs = -1;
    for (int q = 0; q < ns; q++) {
      const int c = L.secell[q];
      if (c == le) ls = q;
      if (re >= 0 && c == re) rs = q;
    }
    L.lse[l] = ls;
    L.rse[l] = rs;
  }
  for (int q = lane; q < fm * fm; q += 64) Nm[q] = 0.0;
  for (int q = lane; q < fm * em; q += 64) Sm[q] = Tm[q] = 0.0;
  for (int q = lane; q < 3 * MS; q += 64) L.B[q] = 0.0;
  wsync_d();

  // ---------------- heat (scalar entries, reference loop order).  Sub-face l
  // only writes row l of N and S, column l of S~^T and B[l], so one lane per
  // sub-face adds every entry's terms in the serial loop's order.
  for (int l0 = lane; l0 < nf; l0 += 64) {
    {
      const int l = l0;
      const double al = L.g[l][3];
      for (int side = 0; side < 2; side++) {
        const int so = side == 0 ? L.lse[l] : L.rse[l];
        if (side == 1 && so < 0) {
          const int re = L.re[l];
          if (re <= -2 && P.bc_T_id[-re - 1] == 2) {
            L.B[l] = al * P.bc_val_T[-re - 1];
            for (int c = 0; c < ns; c++) Sm[l + nf * c] = 0.0;
            for (int c = 0; c < nf; c++) Nm[l + nf * c] = 0.0;
            Nm[l + nf * l] = al;
          }
          continue;
        }
        const double sgl = side == 0 ? 1.0 : -1.0;
        const double nl[3] = {sgl * L.g[l][0], sgl * L.g[l][1], sgl * L.g[l][2]};
        const int cell = L.secell[so];
        const double vol = L.sevol[so];
        const double kap = a.mu_cell ? L.semu[so] * P.cv * kGamma / P.pr : P.kap;  // kappa (:186-194)
        for (int e = 0; e < L.ssf[so][0]; e++) {
          const int po = L.ssf[so][1 + e];
          const double sgp = L.le[po] == cell ? 1.0 : -1.0;
          const double np[3] = {sgp * L.g[po][0], sgp * L.g[po][1], sgp * L.g[po][2]};
          const double blc = kap * al * L.g[po][3] * (1.0 / vol) * (np[0] * nl[0] + np[1] * nl[1] + np[2] * nl[2]);
          Nm[l + nf * po] = Nm[l + nf * po] + blc;
          Sm[l + nf * so] = Sm[l + nf * so] + blc;
          Tm[so + ns * l] = Tm[so + ns * l] + blc;
        }
      }
    }
  }
  wsync_d();
  wave_inverse(Nm, nf, L.piv, L.work, lane);
  // SNS_h = (T N^-1) S ; SB_h = T (N^-1 B) ; b1 = row sums of T
  double *o_snsh = out + (long)E * E + E + 9L * ns + (long)F * E + F;
  double *o_sbh = o_snsh + (long)ns * ns;
  double *o_b1 = o_sbh + ns;
  wave_gemm(ns, nf, nf, Tm, Nm, Wm, lane);
  wsync_d();
  wave_gemm(ns, ns, nf, Wm, Sm, o_snsh, lane);
  wave_gemv(nf, nf, Nm, L.B, L.x, lane);
  wsync_d();
  wave_gemv(ns, nf, Tm, L.x, o_sbh, lane);
  for (int l = lane; l < ns; l += 64) {
    double acc = 0.0;
    for (int c = 0; c < nf; c++) acc = acc + Tm[l + ns * c];
    o_b1[l] = acc;
  }
  wsync_d();

  // ---------------- momentum (3x3 blocks, lane (i, j) owns entry (i, j) of a block)
  for (int q = lane; q < fm * fm; q += 64) Nm[q] = 0.0;
  for (int q = lane; q < fm * em; q += 64) Sm[q] = Tm[q] = 0.0;
  for (int q = lane; q < 3 * MS; q += 64) L.B[q] = 0.0;
  wsync_d();
  // sub-face l only writes its rows of N and S and its columns of S~^T:
  // lanes (l mod 7, i, j), l = lane / 9, lane / 9 + 7, ...
  if (lane < 63) {
    const int i = (lane % 9) % 3, j = (lane % 9) / 3;
    const double mu0 = a.mu_cell ? *a.mu_min : P.mu0;
    for (int l = lane / 9; l < nf; l += 7) {
      const int re = L.re[l];
      const double al = L.g[l][3];
      bool neumann, dirichlet;
      if (re >= 0) {
        neumann = dirichlet = false;
      } else if (re <= -2) {
        dirichlet = P.bc_V_id[-re - 1] == 2;
        neumann = P.bc_V_id[-re - 1] == 1;
      } else {
        neumann = true;
        dirichlet = false;
      }
      const int fl0 = 3 * l;
      const int nsides = (!neumann && !dirichlet) ? 2 : 1;
      for (int side = 0; side < nsides; side++) {
        const int so = side == 0 ? L.lse[l] : L.rse[l];
        const int se0l = 3 * so;
        const double sgl = side == 0 ? 1.0 : -1.0;
        const double nl[3] = {sgl * L.g[l][0], sgl * L.g[l][1], sgl * L.g[l][2]};
        const int cell = L.secell[so];
        const double mu = L.semu[so];  // mu_arr of the side's cell
        const double vol = L.sevol[so];
        for (int e = 0; e < L.ssf[so][0]; e++) {
          const int po = L.ssf[so][1 + e];
          const int fp0 = 3 * po;
          const double sgp = L.le[po] == cell ? 1.0 : -1.0;
          const double np[3] = {sgp * L.g[po][0], sgp * L.g[po][1], sgp * L.g[po][2]};
          const double sc = al * L.g[po][3] * (1.0 / vol);
          const double bmu = blc_mu_ij(sc, mu, mu0, nl, np, i, j);
          const double bmu_t = blc_mu_ij(sc, mu, mu0, nl, np, j, i);
          if (!neumann && !dirichlet) {
            Tm[(se0l + i) + E * (fl0 + j)] = Tm[(se0l + i) + E * (fl0 + j)] + bmu_t;
            Nm[(fl0 + i) + F * (fp0 + j)] = Nm[(fl0 + i) + F * (fp0 + j)] + bmu;
            Sm[(fl0 + i) + F * (se0l + j)] = Sm[(fl0 + i) + F * (se0l + j)] + bmu;
          } else if (neumann) {
            const double dd = nl[0] * np[0] + nl[1] * np[1] + nl[2] * np[2];
            const double bn = sc * (dd * (i == j ? 1.0 : 0.0) + (np[j] * nl[i]) / 3.0);
            Nm[(fl0 + i) + F * (fp0 + j)] = Nm[(fl0 + i) + F * (fp0 + j)] + bn;
            Sm[(fl0 + i) + F * (se0l + j)] = Sm[(fl0 + i) + F * (se0l + j)] + bn;
            Tm[(se0l + i) + E * (fl0 + j)] = Tm[(se0l + i) + E * (fl0 + j)] + bmu_t;
          } else {
            Tm[(se0l + i) + E * (fl0 + j)] = Tm[(se0l + i) + E * (fl0 + j)] + bmu_t;
          }
        }
      }
    }
  }
  wsync_d();
  // Dirichlet rows (each sub-face's rows only receive its own terms above)
  if (lane == 0) {
    for (int l = 0; l < nf; l++) {
      const int re = L.re[l];
      if (re <= -2 && P.bc_V_id[-re - 1] == 2) {
        const double al = L.g[l][3];
        for (int d = 0; d < 3; d++) {
          const int r = 3 * l + d;
          L.B[r] = al * P.bc_val_V[-re - 1][d];
          for (int c = 0; c < E; c++) Sm[r + F * c] = 0.0;
          for (int c = 0; c < F; c++) Nm[r + F * c] = 0.0;
        }
        for (int d = 0; d < 3; d++) Nm[(3 * l + d) + F * (3 * l + d)] = al * 1.0;
      }
    }
  }
  wsync_d();
  double *o_sns = out;
  double *o_sb = o_sns + (long)E * E;
  double *o_blcs = o_sb + E;
  double *o_ninvs = o_blcs + 9L * ns;
  double *o_ninvb = o_ninvs + (long)F * E;
  // sub-element sums of S~^T blocks (before anything overwrites T)
  for (int q = lane; q < 9 * ns; q += 64) {
    const int p = q / 9, ij = q % 9, i = ij % 3, j = ij / 3;
    double acc = 0.0;
    for (int l = 0; l < nf; l++) acc = acc + Tm[(3 * p + i) + E * (3 * l + j)];
    o_blcs[q] = acc;
  }
  wave_inverse(Nm, F, L.piv, L.work, lane);
  wave_gemm(E, F, F, Tm, Nm, Wm, lane);
  wsync_d();
  wave_gemm(E, E, F, Wm, Sm, o_sns, lane);
  wave_gemv(F, F, Nm, L.B, L.x, lane);
  wsync_d();
  wave_gemv(E, F, Tm, L.x, o_sb, lane);
  for (int q = lane; q < F; q += 64) o_ninvb[q] = L.x[q];
  wave_gemm(F, E, F, Nm, Sm, o_ninvs, lane);
}

// temp_u and dtheta_dU (ns_euler_primitives_module.F90:52-63, ns_diffusion_module.F90:196-206)
SFV_DEV double temp_u(const double *u, double cv) {
  const double tau = 1.0 / u[0];
  return (u[4] * tau - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * (tau * tau)) / cv;
}
SFV_DEV void dtheta_du(const double *u, double cv, double *P) {
  const double v1 = u[1] / u[0], v2 = u[2] / u[0], v3 = u[3] / u[0];
  const double nrm = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
  P[0] = (nrm * nrm - u[4] / u[0]) / (u[0] * cv);
  for (int k = 1; k < 4; k++) P[k] = -u[k] / ((u[0] * u[0]) * cv);
  P[4] = 1.0 / (u[0] * cv);
}

constexpr int kDiffWaves = 4;
// slot-term data staged in LDS per vertex: sb and blcs (12 ns), snsh, sbh and
// b1 (ns^2 + 2 ns), the diagonal 3x3 blocks of S~^T N^-1 S (9 ns); vertices
// with more sub-elements read them from the block in HBM instead
constexpr int kDiffStage = 240;
// compact class of the apply: up to kDiffMaxSub sub-faces and kDiffMaxSubC
// sub-elements (prism and hex meshes), 4 KB of LDS per vertex, so eight
// workgroups of four vertices (eight wavefronts per SIMD, the most a SIMD
// holds) share a CU instead of seven
constexpr int kDiffMaxSubC = 8;
constexpr int kDiffStageC = kDiffMaxSubC * kDiffMaxSubC + 23 * kDiffMaxSubC;  // 248

// MF: sub-faces, ME: sub-elements per vertex; STAGE: doubles of staging
template <int MF, int ME, int STAGE>
struct DiffApplyLds {
  static constexpr int kStage = STAGE;
  static constexpr int kPc = STAGE > 13 * ME ? STAGE : 13 * ME;
  double Ve[3 * ME], Vf[3 * MF], SVe[3 * ME];
  double th[ME], rho[ME], P[ME][5], vol[ME], pu[ME];
  double nrm[MF][4];           // unit normal and area of each sub-face of the vertex
  int le[MF];                  // left cell of each sub-face
  int cell[ME];                // cell of each slot
  int ssf[ME][kSlotSrc + 1];   // local sub-faces of each slot's sub-element, count first
  double pc[kPc];              // the slot terms' block data, staged while the row pass streams;
                               // then the stresses (9 ns) and work terms (4 ns)
};

// (c) of k_diff_apply: the slot rows of one vertex from the row pass's Vf and
// SVe.  STAGED: sb, blcs, snsh, sbh, b1 and the diagonal 3x3 blocks of
// S~^T N^-1 S were staged in L.pc; otherwise they are read from the block.
// The viscous work is spread over (slot, tensor entry) and (slot, sub-face)
// lanes; every value is formed with the same operations in the same order
// as one lane per slot would.
template <bool STAGED, class LdsT>
SFV_DEV void diff_slot_terms(const DiffArgs &a, LdsT &L, int lane, int ns, int se0, int E, const double *sns,
                             const double *sb, const double *blcs, const double *snsh) {
  const int o_tail = 12 * ns, o_diag = 12 * ns + ns * ns + 2 * ns;
  const double *c_sb = STAGED ? L.pc : sb, *c_blcs = STAGED ? L.pc + 3 * ns : blcs;
  const double *c_snsh = STAGED ? L.pc + o_tail : snsh;
  const double *c_sbh = c_snsh + (long)ns * ns, *c_b1 = c_sbh + ns;
  // diagonal block entry (i, j) of slot l
  auto sdiag = [&](int l, int i, int j) -> double {
    return STAGED ? L.pc[o_diag + 9 * l + i + 3 * j] : sns[(3 * l + i) + (long)E * (3 * l + j)];
  };
  double *PD = a.PD + (long)kDiffSlot * se0;
  // the momentum system's terms (build_momentum_diff_mat_into_global_mat_diag)
  // and the heat system's (build_diff_mat_into_global_mat_diag), or both
  const bool mom = a.parts & 1, heat = a.parts & 2;
  // diagonal block of the momentum rows, 9 entries per slot
  for (int t = lane; t < 9 * ns; t += 64) {
    const int l = t / 9, ij = t - 9 * l, i = ij % 3, j = ij / 3;
    const double rho = L.rho[l];
    PD[(long)kDiffSlot * l + 4 + ij] = mom ? -sdiag(l, i, j) / rho + c_blcs[9 * l + j + 3 * i] / rho : 0.0;
  }
  // momentum rhs, 3 per slot
  for (int t = lane; t < 3 * ns; t += 64) {
    const int l = t / 3, i = t - 3 * l;
    double mv = 0.0;
    for (int j = 0; j < 3; j++) mv = mv + sdiag(l, i, j) * L.Ve[3 * l + j];
    PD[(long)kDiffSlot * l + i] = mom ? (L.SVe[3 * l + i] - mv) + c_sb[3 * l + i] : 0.0;
  }
  // heat row of the diagonal, 5 per slot
  for (int t = lane; t < 5 * ns; t += 64) {
    const int l = t / 5, k = t - 5 * l;
    const double sll = c_snsh[l + (long)ns * l];
    PD[(long)kDiffSlot * l + 13 + k] = heat ? (c_b1[l] - sll) * L.P[l][k] : 0.0;
  }
  wsync_d();
  // the viscous stress of each slot's sub-element, entry q = i + 3 j per
  // lane: sum over its sub-faces of area (vm_i nl_j + vm_j nl_i - 2/3 dv
  // delta_ij), times mu / vol.  The staged sb / blcs and diagonal blocks
  // are dead now: spc takes the first, the work terms the second region
  // (or the free staging area when nothing was staged).
  double *spc = L.pc;
  double *wt = L.pc + (STAGED ? o_diag : 9 * ns);
  for (int t = lane; t < 9 * ns; t += 64) {
    const int l = t / 9, q = t - 9 * l, i = q % 3, j = q / 3;
    const int cell = L.cell[l], cnt = L.ssf[l][0];
    double sp = 0.0;
    for (int e = 0; e < cnt; e++) {
      const int lo = L.ssf[l][1 + e];
      const double sg = L.le[lo] == cell ? 1.0 : -1.0;
      double nl[3], vm[3];
      for (int d = 0; d < 3; d++) nl[d] = sg * L.nrm[lo][d];
      for (int d = 0; d < 3; d++) vm[d] = L.Vf[3 * lo + d] - L.Ve[3 * l + d];
      const double dv = vm[0] * nl[0] + vm[1] * nl[1] + vm[2] * nl[2];
      const double ar = L.nrm[lo][3];
      const double vmi = L.Vf[3 * lo + i] - L.Ve[3 * l + i], vmj = L.Vf[3 * lo + j] - L.Ve[3 * l + j];
      const double nli = sg * L.nrm[lo][i], nlj = sg * L.nrm[lo][j];
      sp = sp + ar * (vmi * nlj + vmj * nli - (2.0 / 3.0) * dv * (i == j ? 1.0 : 0.0));
    }
    const double mu = a.mu_cell ? a.mu_cell[cell] : a.p.mu;
    spc[t] = sp * mu / L.vol[l];
  }
  wsync_d();
  // area * (Spc Vf) . nl of each (slot, sub-face)
  for (int t = lane; t < kSlotSrc * ns; t += 64) {
    const int l = t / kSlotSrc, e = t - kSlotSrc * l;
    if (e < L.ssf[l][0]) {
      const int lo = L.ssf[l][1 + e];
      const double sg = L.le[lo] == L.cell[l] ? 1.0 : -1.0;
      const double *S = spc + 9 * l;
      double nl[3];
      for (int d = 0; d < 3; d++) nl[d] = sg * L.nrm[lo][d];
      double sv[3];
      for (int i = 0; i < 3; i++)
        sv[i] = S[i] * L.Vf[3 * lo] + S[i + 3] * L.Vf[3 * lo + 1] + S[i + 6] * L.Vf[3 * lo + 2];
      wt[t] = L.nrm[lo][3] * (sv[0] * nl[0] + sv[1] * nl[1] + sv[2] * nl[2]);
    }
  }
  wsync_d();
  // viscous work (its terms in sub-face order) and heat rhs, one lane per slot
  for (int l = lane; l < ns; l += 64) {
    const int cnt = L.ssf[l][0];
    double work = 0.0;
    for (int e = 0; e < cnt; e++) work = work + wt[kSlotSrc * l + e];
    double h = 0.0;
    for (int p = 0; p < ns; p++) h = h + c_snsh[l + (long)ns * p] * L.th[p];
    const double sll = c_snsh[l + (long)ns * l];
    const double pu = L.pu[l];
    h = h - sll * pu + (c_b1[l] - sll) * pu - c_b1[l] * L.th[l] + c_sbh[l];
    PD[(long)kDiffSlot * l + 3] = mom && heat ? work + h : (mom ? work : h);
  }
}

// One wavefront per vertex: the scatter terms of
// build_momentum_diff_mat_into_global_mat_diag (:231-317) and
// build_diff_mat_into_global_mat_diag (:139-182) per sub-element slot.
// Row of PD: [0..2] rhs(2:4), [3] rhs(5), [4..12] diag(2:4, 2:4) column-major,
// [13..17] diag(5, 1:5).
//
// (a) lanes < ns load their slot's cell state and sub-element sub-face list,
//     lanes < nf their sub-face's geometry, all into LDS, in parallel;
// (b) lane r < F + E owns one row of the two column-major streams N^-1 S
//     (F x E) and S~^T N^-1 S (E x E): Vf = N^-1 S Ve + N^-1 B and
//     SVe = S~^T N^-1 S Ve, coalesced;
// (c) the slot terms, spread over the lanes by (slot, entry), read only LDS
//     and the vertex's block the wavefront has just streamed.
// MF, ME, STAGE: the size class (DiffApplyLds); WAVES: vertices
// (wavefronts) per workgroup; MINW: wavefronts per SIMD the registers allow.
template <int MF, int ME, int STAGE, int WAVES, int MINW>
__global__ __launch_bounds__(64 * WAVES, MINW) void k_diff_apply(DiffArgs a) {
  using Lds = DiffApplyLds<MF, ME, STAGE>;
  __shared__ Lds lds_all[WAVES];
  const DevMesh &m = a.m;
  // the wavefront's index is uniform: the vertex's offsets come through the
  // scalar cache
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int lb = xcd_block(a.vper_xcd);  // neighbouring vertices share cells
  if (lb >= a.vblocks) return;
  const long v = (long)lb * WAVES + wib;
  if (v >= m.V) return;  // whole wavefront
  Lds &L = lds_all[wib];
  const int sf0 = m.sf_off[v], nf = m.sf_off[v + 1] - sf0;
  const int se0 = m.se_off[v], ns = m.se_off[v + 1] - se0;
  const int F = 3 * nf, E = 3 * ns;
  const double cv = a.p.cv;
  const double *pre = a.pre + a.doff[v];
  const double *sns = pre, *sb = sns + (long)E * E, *blcs = sb + E, *ninvs = blcs + 9L * ns;
  const double *ninvb = ninvs + (long)F * E, *snsh = ninvb + F, *sbh = snsh + (long)ns * ns, *b1 = sbh + ns;
  // ---- (a)
  for (int l = lane; l < ns; l += 64) {
    const int c = m.se_cell[se0 + l];
    const uint32_t code = a.slot_ssf[se0 + l];
    L.vol[l] = a.slot_vol[se0 + l];
    double u[5];
    for (int k = 0; k < 5; k++) u[k] = a.U[5L * c + k];
    int cnt = 0;
    for (int e = 0; e < kSlotSrc; e++) {
      const int j = (code >> (8 * e)) & 0xFFu;
      if (j == 0xFF) break;
      L.ssf[l][1 + e] = j;
      cnt++;
    }
    L.ssf[l][0] = cnt;
    for (int k = 0; k < 3; k++) L.Ve[3 * l + k] = u[k + 1] / u[0];
    L.rho[l] = u[0];
    L.cell[l] = c;
    L.th[l] = temp_u(u, cv);
    double P[5];
    dtheta_du(u, cv, P);
    double pu = 0.0;
    for (int k = 0; k < 5; k++) pu = pu + P[k] * u[k];
    for (int k = 0; k < 5; k++) L.P[l][k] = P[k];
    L.pu[l] = pu;
  }
  for (int l = lane; l < nf; l += 64) {
    const long s = sf0 + l;
    for (int k = 0; k < 4; k++) L.nrm[l][k] = m.sfn[k * m.SF + s];
    L.le[l] = m.sf_le[s];
  }
  // stage the slot terms' block data: the loads overlap the row pass below
  const bool staged = ns * ns + 23 * ns <= Lds::kStage;
  const int o_tail = 12 * ns, o_diag = 12 * ns + ns * ns + 2 * ns;
  if (staged) {
    for (int t = lane; t < 12 * ns; t += 64) L.pc[t] = sb[t];  // sb then blcs, contiguous
    for (int t = lane; t < ns * ns + 2 * ns; t += 64) L.pc[o_tail + t] = snsh[t];  // snsh, sbh, b1
    for (int t = lane; t < 9 * ns; t += 64) {
      const int l = t / 9, ij = t - 9 * l, i = ij % 3, j = ij / 3;
      L.pc[o_diag + t] = sns[(3 * l + i) + (long)E * (3 * l + j)];
    }
  }
  wsync_d();
  // ---- (b)
  for (int r = lane; r < F + E; r += 64) {
    const bool vf = r < F;
    const double *col = vf ? ninvs + r : sns + (r - F);
    const long ld = vf ? F : E;
    double acc = 0.0;
#pragma unroll 6
    for (int j = 0; j < E; j++) acc = acc + L.Ve[j] * col[ld * j];
    if (vf)
      L.Vf[r] = acc + ninvb[r];
    else
      L.SVe[r - F] = acc;
  }
  wsync_d();
  // ---- (c), with the slot data in LDS or read from the block in HBM
  if (staged)
    diff_slot_terms<true>(a, L, lane, ns, se0, E, sns, sb, blcs, snsh);
  else
    diff_slot_terms<false>(a, L, lane, ns, se0, E, sns, sb, blcs, snsh);
}

// rdiff and diag of each cell: its slots' rows in increasing vertex id.  One
// lane per (cell, row entry), kDgCells cells per workgroup, so the 18 entries
// of a row are read by consecutive lanes; the sums are then assembled from
// LDS into rdiff (4 per cell) and diag (25 per cell), written by consecutive
// lanes too.
constexpr int kDgCells = 14;  // 14 x 18 = 252 of 256 lanes
constexpr int kDgBatch = 8;   // rows loaded together per lane
__global__ __launch_bounds__(256) void k_diff_gather(DiffArgs a) {
  __shared__ double s[kDgCells][kDiffSlot];
  const DevMesh &m = a.m;
  const int lb = xcd_block(a.cper_xcd);
  if (lb >= a.cblocks) return;  // uniform over the workgroup
  const int t = threadIdx.x, cl = t / kDiffSlot, k = t - cl * kDiffSlot;
  const long c0 = (long)lb * kDgCells;
  if (cl < kDgCells && c0 + cl < m.C) {
    const long c = c0 + cl;
    const int b = m.e_voff[c], nv = m.e_voff[c + 1] - b;
    // kDgBatch slot ids, then their rows' entries (row 0 stands in past the
    // cell's last vertex), so the loads of a batch are in flight together;
    // the sum takes the cell's rows in vertex order
    double acc = 0.0;
    for (int j0 = 0; j0 < nv; j0 += kDgBatch) {
      int sl[kDgBatch];
#pragma unroll
      for (int u = 0; u < kDgBatch; u++) sl[u] = j0 + u < nv ? m.c_slot[b + j0 + u] : 0;
      double x[kDgBatch];
#pragma unroll
      for (int u = 0; u < kDgBatch; u++) x[u] = a.PD[(long)kDiffSlot * sl[u] + k];
#pragma unroll
      for (int u = 0; u < kDgBatch; u++)
        if (j0 + u < nv) acc = acc + x[u];
    }
    s[cl][k] = acc;
  }
  __syncthreads();
  // [0..3] rhs of rho u, rho v, rho w, E; [4..12] diag(2:4, 2:4) column-major;
  // [13..17] diag(5, 1:5), added to zero as the 5x5 block is formed
  for (int q = t; q < kDgCells * 29; q += 256) {
    const int cc = q / 29, o = q - cc * 29;
    const long c = c0 + cc;
    if (c >= m.C) continue;
    const double *d = s[cc];
    if (o < 4) {
      a.rdiff[4 * c + o] = d[o];
    } else {
      const int e = o - 4, row = e % 5, col = e / 5;
      double x = 0.0;
      if (row >= 1 && row <= 3 && col >= 1 && col <= 3) x = d[4 + (row - 1) + 3 * (col - 1)];
      if (row == 4) x = 0.0 + d[13 + col];
      a.diag[25 * c + e] = x;
    }
  }
}

}  // namespace sfv
