// step_lowp.hip -- FP32-compute instantiations of the fused step over shifted
// populations g = f - w stored as FP32, FP16 or BF16.
#include "launch.hpp"

namespace hlbm {

hipError_t launch_step_lowp(int storage, int coll, int mode, int npt, dim3 grid, dim3 block, const StepParams& p,
                            hipStream_t s) {
  switch (storage) {
    case 1:
      return step_dispatch<float>(coll, mode, npt, grid, block, p, s);
    case 2:
      return step_dispatch<_Float16>(coll, mode, npt, grid, block, p, s);
    case 3:
      return step_dispatch<bf16s>(coll, mode, npt, grid, block, p, s);
    default:
      return hipErrorInvalidValue;
  }
}

template <class S>
static hipError_t batch_one(int coll, unsigned blocks, const BatchParams& bp, hipStream_t s) {
  if (coll == COLL_BGK)
    hipLaunchKernelGGL((batch_kernel<S, COLL_BGK>), dim3(blocks), dim3(256), 0, s, bp);
  else
    hipLaunchKernelGGL((batch_kernel<S, COLL_MRT>), dim3(blocks), dim3(256), 0, s, bp);
  return hipGetLastError();
}

hipError_t launch_batch_lowp(int storage, int coll, unsigned blocks, const BatchParams& bp, hipStream_t s) {
  switch (storage) {
    case 1:
      return batch_one<float>(coll, blocks, bp, s);
    case 2:
      return batch_one<_Float16>(coll, blocks, bp, s);
    case 3:
      return batch_one<bf16s>(coll, blocks, bp, s);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace hlbm
