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

}  // namespace sfv
