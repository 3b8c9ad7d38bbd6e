// step_f64.hip -- FP64 instantiations of the fused step.  This translation
// unit is compiled with -ffp-contract=off: no multiply-add is fused, so every
// operation rounds as in hippoLBM's CPU build.
#include "launch.hpp"

namespace hlbm {

hipError_t launch_step_f64(int coll, int mode, int npt, dim3 grid, dim3 block, const StepParams& p, hipStream_t s) {
  return step_dispatch<double>(coll, mode, npt, grid, block, p, s);
}

hipError_t launch_batch_f64(int coll, unsigned blocks, const BatchParams& bp, hipStream_t s) {
  if (coll == COLL_BGK)
    hipLaunchKernelGGL((batch_kernel<double, COLL_BGK>), dim3(blocks), dim3(256), 0, s, bp);
  else
    hipLaunchKernelGGL((batch_kernel<double, COLL_MRT>), dim3(blocks), dim3(256), 0, s, bp);
  return hipGetLastError();
}

}  // namespace hlbm
