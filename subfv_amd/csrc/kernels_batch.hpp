// Multi-instance launches for served solves: one launch advances every
// instance of a batch by one stage of the reference's forward-Euler loop
// (first order).  Each instance keeps its own mesh, state buffers, partials,
// step state and reductions; a block finds its instance and its block within
// the instance in a table built when the batch membership changes, so the
// per-instance arithmetic (and hence every result) is exactly that of the
// single-instance kernels.
//
// State buffers rotate: at phase p an instance reads U[(cur0 + p) % 3] and
// writes U[(cur0 + p + 1) % 3], so three phases form a closed cycle that one
// captured hipGraph can replay.
#pragma once
#include <hip/hip_cooperative_groups.h>

#include "kernels_cell.hpp"
#include "kernels_flux.hpp"

namespace sfv {

struct InstDesc {
  FluxArgs fa;    // fa.U unused (state from U[])
  GatherArgs ga;
  UpdateArgs ua;  // ua.U / ua.Uout unused
  double *U[3];
  int cur0;       // state index at phase 0
  int ngroups;    // vertex groups of the node sweep
  int ncblk;      // cell blocks (kCellThreads cells each)
  int pad;
};

typedef const InstDesc __attribute__((address_space(4))) CInstDesc;

constexpr int kTabShift = 20;  // table entry = instance << 20 | block within the instance
constexpr int kTabMask = (1 << kTabShift) - 1;

template <int SCHEME>
__global__ __launch_bounds__(kGrpThreads, 3) void k_node_flux_multi(const InstDesc *d, const int *gtab, int total,
                                                                   int per_xcd, int phase) {
  __shared__ FluxLds L;
  const int lb = xcd_block(per_xcd);
  if (lb >= total) return;
  const int e = gtab[lb];
  // the descriptors do not change during a launch: reading them through the
  // constant address space lets the compiler keep them across the sweep's
  // barriers instead of re-reading them
  const InstDesc &D = *(const InstDesc *)((const CInstDesc *)d + (e >> kTabShift));
  // a table entry past the instance's groups would read foreign headers
  if ((e & kTabMask) >= D.ngroups) return;
  node_flux_group<SCHEME>(D.fa, D.U[(D.cur0 + phase) % 3], e & kTabMask, L);
}

__global__ __launch_bounds__(kCellThreads) void k_gather_multi(const InstDesc *d, const int *ctab, int total) {
  if ((int)blockIdx.x >= total) return;
  const int e = ctab[blockIdx.x];
  const InstDesc &D = *(const InstDesc *)((const CInstDesc *)d + (e >> kTabShift));
  if ((e & kTabMask) >= D.ncblk) return;
  gather_block(D.ga, e & kTabMask);
}

__global__ __launch_bounds__(kCellThreads) void k_finalize_dt_multi(const InstDesc *d) {
  const InstDesc &D = d[blockIdx.x];
  finalize_dt(D.ga.blk, D.ncblk, D.ga.st, 0);
}

__global__ __launch_bounds__(kCellThreads) void k_update_multi(const InstDesc *d, const int *ctab, int total,
                                                               int phase) {
  if ((int)blockIdx.x >= total) return;
  const int e = ctab[blockIdx.x];
  const InstDesc &D = *(const InstDesc *)((const CInstDesc *)d + (e >> kTabShift));
  if ((e & kTabMask) >= D.ncblk) return;
  update_block<false>(D.ua, D.U[(D.cur0 + phase) % 3], D.U[(D.cur0 + phase + 1) % 3], e & kTabMask);
}

__global__ __launch_bounds__(kCellThreads) void k_finalize_omega_multi(const InstDesc *d) {
  const InstDesc &D = d[blockIdx.x];
  finalize_omega(D.ua.blk, D.ncblk, D.ua.st, 1);
}

__global__ __launch_bounds__(kCellThreads) void k_update_fixup_multi(const InstDesc *d, const int *ctab, int total,
                                                                     int phase) {
  if ((int)blockIdx.x >= total) return;
  const int e = ctab[blockIdx.x];
  const InstDesc &D = *(const InstDesc *)((const CInstDesc *)d + (e >> kTabShift));
  if ((e & kTabMask) >= D.ncblk) return;
  update_fixup_block<false>(D.ua, D.U[(D.cur0 + phase) % 3], D.U[(D.cur0 + phase + 1) % 3], e & kTabMask);
}

// ------------------------------------------------------- persistent form
// One cooperative launch advances the batch by n_steps steps: the grid stays
// resident and walks the four phases of every step (node sweep, gather,
// update, omega fixup) with a grid-wide barrier after each.  Within a phase
// a workgroup takes its next item (a vertex group or a cell block of some
// instance) from a device-side queue: one counter per XCD slice of the
// phase's table, so neighbouring items stay on one XCD.  The reductions of
// an instance (dt after its gather, omega after its update) are finished by
// the instance's last block to arrive, as the single-instance kernels
// finish them, from the same block minima in the same order.
struct PersistArgs {
  const InstDesc *d;
  const int *gtab, *ctab;
  int total_g, total_c, n_inst, n_steps;
  unsigned *qctr;  // [n_steps][4][8] queue counters, zero at launch
  unsigned *done;  // [n_steps][2][n_inst] blocks finished per instance, zero at launch
};

// the next item of this workgroup's XCD slice [lo, hi) of a phase table, or
// -1 when the slice is drained (uniform over the workgroup)
SFV_DEV int next_item(unsigned *ctr, int lo, int hi, int *s_item) {
  if (threadIdx.x == 0) *s_item = lo + (int)atomicAdd(ctr, 1u);
  __syncthreads();
  const int k = *s_item;
  __syncthreads();
  return k < hi ? k : -1;
}

// this block's arrival at an instance's reduction; true for the last one
SFV_DEV bool last_arrival(unsigned *cnt, int nblk, int *s_flag) {
  if (threadIdx.x == 0) {
    __threadfence();  // the block minimum written before it is visible device-wide
    *s_flag = atomicAdd(cnt, 1u) == (unsigned)(nblk - 1);
    __threadfence();  // and the last one reads every block's
  }
  __syncthreads();
  const bool last = *s_flag != 0;
  __syncthreads();
  return last;
}

template <int SCHEME>
__global__ __launch_bounds__(kGrpThreads, 3) void k_batch_persistent(PersistArgs a) {
  __shared__ FluxLds L;
  __shared__ int s_item, s_flag;
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int x = blockIdx.x & 7;
  const int gper = (a.total_g + 7) / 8, cper = (a.total_c + 7) / 8;
  const int glo = x * gper, ghi = glo + gper < a.total_g ? glo + gper : a.total_g;
  const int clo = x * cper, chi = clo + cper < a.total_c ? clo + cper : a.total_c;
  for (int s = 0; s < a.n_steps; s++) {
    const int phase = s % 3;
    unsigned *q = a.qctr + 32L * s;
    unsigned *dn = a.done + 2L * a.n_inst * s;
    // node sweep
    for (int k; (k = next_item(q + x, glo, ghi, &s_item)) >= 0;) {
      const int e = a.gtab[k];
      const InstDesc &D = *(const InstDesc *)((const CInstDesc *)a.d + (e >> kTabShift));
      if ((e & kTabMask) < D.ngroups) node_flux_group<SCHEME>(D.fa, D.U[(D.cur0 + phase) % 3], e & kTabMask, L);
      __syncthreads();
    }
    grid.sync();
    // gather, then dt of each instance by its last block
    for (int k; (k = next_item(q + 8 + x, clo, chi, &s_item)) >= 0;) {
      const int e = a.ctab[k], i = e >> kTabShift;
      const InstDesc &D = *(const InstDesc *)((const CInstDesc *)a.d + i);
      if ((e & kTabMask) >= D.ncblk) continue;
      gather_block(D.ga, e & kTabMask);
      if (last_arrival(dn + i, D.ncblk, &s_flag)) finalize_dt(D.ga.blk, D.ncblk, D.ga.st, 0);
      __syncthreads();
    }
    grid.sync();
    // update, then omega and the time advance by the last block
    for (int k; (k = next_item(q + 16 + x, clo, chi, &s_item)) >= 0;) {
      const int e = a.ctab[k], i = e >> kTabShift;
      const InstDesc &D = *(const InstDesc *)((const CInstDesc *)a.d + i);
      if ((e & kTabMask) >= D.ncblk) continue;
      update_block<false>(D.ua, D.U[(D.cur0 + phase) % 3], D.U[(D.cur0 + phase + 1) % 3], e & kTabMask);
      if (last_arrival(dn + a.n_inst + i, D.ncblk, &s_flag)) finalize_omega(D.ua.blk, D.ncblk, D.ua.st, 1);
      __syncthreads();
    }
    grid.sync();
    for (int k; (k = next_item(q + 24 + x, clo, chi, &s_item)) >= 0;) {
      const int e = a.ctab[k];
      const InstDesc &D = *(const InstDesc *)((const CInstDesc *)a.d + (e >> kTabShift));
      if ((e & kTabMask) >= D.ncblk) continue;
      update_fixup_block<false>(D.ua, D.U[(D.cur0 + phase) % 3], D.U[(D.cur0 + phase + 1) % 3], e & kTabMask);
      __syncthreads();
    }
    grid.sync();
  }
}

}  // namespace sfv
