// step_f64.hip -- FP64 instantiations of the fused step.  This translation
// unit is compiled with -ffp-contract=off: no multiply-add is fused, so every
// operation rounds as in hippoLBM's CPU build.
#include "launch.hpp"

namespace hlbm {

template <int COLL, int MODE>
static hipError_t go(dim3 grid, dim3 block, const StepParams& p, hipStream_t s) {
  hipLaunchKernelGGL((step_kernel<double, COLL, MODE>), grid, block, 0, s, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || MODE == MODE_COLLIDE || !p.prog_planes) return e;
  hipLaunchKernelGGL((boundary_kernel<double, COLL, MODE>), dim3(boundary_blocks(p)), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_step_f64(int coll, int mode, dim3 grid, dim3 block, const StepParams& p, hipStream_t s) {
  if (coll == COLL_BGK) {
    if (mode == MODE_FUSED) return go<COLL_BGK, MODE_FUSED>(grid, block, p, s);
    if (mode == MODE_COLLIDE) return go<COLL_BGK, MODE_COLLIDE>(grid, block, p, s);
    return go<COLL_BGK, MODE_STREAM>(grid, block, p, s);
  }
  if (mode == MODE_FUSED) return go<COLL_MRT, MODE_FUSED>(grid, block, p, s);
  if (mode == MODE_COLLIDE) return go<COLL_MRT, MODE_COLLIDE>(grid, block, p, s);
  return go<COLL_BGK, MODE_STREAM>(grid, block, p, s);  // streaming does not collide
}

}  // namespace hlbm
