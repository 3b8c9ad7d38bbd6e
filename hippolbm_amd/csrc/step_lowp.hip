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
static hipError_t batch_pair(const unsigned blocks[2], const BatchParams& bp, hipStream_t s) {
  hipLaunchKernelGGL((batch_kernel<S, COLL_BGK>), dim3(blocks[0]), dim3(256), 0, s, bp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((batch_kernel<S, COLL_MRT>), dim3(blocks[1]), dim3(256), 0, s, bp);
  return hipGetLastError();
}

hipError_t launch_batch_lowp(int storage, const unsigned blocks[2], const BatchParams& bp, hipStream_t s) {
  switch (storage) {
    case 1:
      return batch_pair<float>(blocks, bp, s);
    case 2:
      return batch_pair<_Float16>(blocks, bp, s);
    case 3:
      return batch_pair<bf16s>(blocks, bp, s);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace hlbm
