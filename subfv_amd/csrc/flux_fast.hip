// The node sweep of kernels_flux.hpp in its own translation unit, compiled
// with floating-point contraction and reciprocal math (build.py: FAST_FLAGS):
// the 1e-12 arithmetic mode (subfv_set_arith).  A division becomes a product
// with the divisor's reciprocal (computed once where the divisor repeats, as
// across the nodal-solver iterations) and a product feeding a sum one fused
// multiply-add; results then differ from the reference's in the last bits
// (held to 1e-12 relative after 200 steps by the tests).  The device code of
// each translation unit is a code object of its own, so the exact kernels
// keep their arithmetic.
#include "kernels_flux.hpp"

namespace sfv {

template <int SCHEME>
__global__ __launch_bounds__(kGrpThreads, 3) void k_node_flux_fast(FluxArgs a) {
  __shared__ FluxLds L;
  const int lb = xcd_block(a.per_xcd);
  if (lb >= a.nblocks) return;  // uniform over the workgroup
  node_flux_group<SCHEME>(a, a.U, lb, L);
}

void launch_node_flux_fast(int scheme, unsigned grid, hipStream_t s, const FluxArgs &a) {
  switch (scheme) {
    case kMultiPoint: hipLaunchKernelGGL(k_node_flux_fast<kMultiPoint>, dim3(grid), dim3(kGrpThreads), 0, s, a); break;
    case kMultiPointIso: hipLaunchKernelGGL(k_node_flux_fast<kMultiPointIso>, dim3(grid), dim3(kGrpThreads), 0, s, a); break;
    case kThreeWave: hipLaunchKernelGGL(k_node_flux_fast<kThreeWave>, dim3(grid), dim3(kGrpThreads), 0, s, a); break;
    case kTwoWave: hipLaunchKernelGGL(k_node_flux_fast<kTwoWave>, dim3(grid), dim3(kGrpThreads), 0, s, a); break;
    case kModifiedThreeWave:
      hipLaunchKernelGGL(k_node_flux_fast<kModifiedThreeWave>, dim3(grid), dim3(kGrpThreads), 0, s, a);
      break;
    case kMultiPointPressure:
      hipLaunchKernelGGL(k_node_flux_fast<kMultiPointPressure>, dim3(grid), dim3(kGrpThreads), 0, s, a);
      break;
    default: hipLaunchKernelGGL(k_node_flux_fast<kMultiPointPressurePh>, dim3(grid), dim3(kGrpThreads), 0, s, a); break;
  }
}

}  // namespace sfv
