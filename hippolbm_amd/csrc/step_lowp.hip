// step_lowp.hip -- FP32-compute instantiations of the fused step over shifted
// populations g = f - w stored as FP32, FP16 or BF16.
#include "launch.hpp"

namespace hlbm {

template <class S, int COLL, int MODE, int NPT, bool WIDE>
static hipError_t go(dim3 grid, dim3 block, const StepParams& p, hipStream_t s) {
  hipError_t e = hipSuccess;
  if constexpr (MODE == MODE_AA_EVEN || MODE == MODE_AA_ODD) {
    if (p.oflow && !p.skip_program) {  // outflow values first: the step below updates nodes in place
      hipLaunchKernelGGL((outflow_capture_kernel<S, MODE>), dim3((p.oflow_n + 255) / 256), dim3(256), 0, s, p);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  // the program-plane nodes and the bulk touch disjoint nodes (A/B) or disjoint
  // slots (AA), so either kernel may run first
  const bool planes = MODE != MODE_COLLIDE && p.prog_planes;
  if (planes && p.boundary_first) {
    hipLaunchKernelGGL((boundary_kernel<S, COLL, MODE>), dim3(boundary_blocks(p)), dim3(BOUNDARY_BLOCK), 0, s, p);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (MODE != MODE_AA_FIX) {  // AA_FIX only touches program-plane nodes
    hipLaunchKernelGGL((step_kernel<S, COLL, MODE, NPT, WIDE>), grid, block, 0, s, p);
    e = hipGetLastError();
  }
  if (e != hipSuccess || !planes || p.boundary_first) return e;
  hipLaunchKernelGGL((boundary_kernel<S, COLL, MODE>), dim3(boundary_blocks(p)), dim3(BOUNDARY_BLOCK), 0, s, p);
  return hipGetLastError();
}

template <class S, int COLL, int MODE>
static hipError_t go_npt(int npt, dim3 grid, dim3 block, const StepParams& p, hipStream_t s) {
  if (wide_offsets(p, sizeof(S))) return go<S, COLL, MODE, 1, true>(grid, block, p, s);
  return npt == 2 ? go<S, COLL, MODE, 2, false>(grid, block, p, s) : go<S, COLL, MODE, 1, false>(grid, block, p, s);
}

template <class S>
static hipError_t dispatch(int coll, int mode, int npt, dim3 grid, dim3 block, const StepParams& p, hipStream_t s) {
  if (mode == MODE_STREAM) return go_npt<S, COLL_BGK, MODE_STREAM>(npt, grid, block, p, s);
  if (mode == MODE_AA_STREAM) return go_npt<S, COLL_BGK, MODE_AA_STREAM>(npt, grid, block, p, s);
  if (mode == MODE_AA_FIX) return go_npt<S, COLL_BGK, MODE_AA_FIX>(npt, grid, block, p, s);
  if (mode == MODE_AA_EVEN)
    return coll == COLL_BGK ? go_npt<S, COLL_BGK, MODE_AA_EVEN>(npt, grid, block, p, s)
                            : go_npt<S, COLL_MRT, MODE_AA_EVEN>(npt, grid, block, p, s);
  if (mode == MODE_AA_ODD)
    return coll == COLL_BGK ? go_npt<S, COLL_BGK, MODE_AA_ODD>(npt, grid, block, p, s)
                            : go_npt<S, COLL_MRT, MODE_AA_ODD>(npt, grid, block, p, s);
  if (coll == COLL_BGK) {
    if (mode == MODE_FUSED) return go_npt<S, COLL_BGK, MODE_FUSED>(npt, grid, block, p, s);
    return go_npt<S, COLL_BGK, MODE_COLLIDE>(npt, grid, block, p, s);
  }
  if (mode == MODE_FUSED) return go_npt<S, COLL_MRT, MODE_FUSED>(npt, grid, block, p, s);
  return go_npt<S, COLL_MRT, MODE_COLLIDE>(npt, grid, block, p, s);
}

hipError_t launch_step_lowp(int storage, int coll, int mode, int npt, dim3 grid, dim3 block, const StepParams& p,
                            hipStream_t s) {
  switch (storage) {
    case 1:
      return dispatch<float>(coll, mode, npt, grid, block, p, s);
    case 2:
      return dispatch<_Float16>(coll, mode, npt, grid, block, p, s);
    case 3:
      return dispatch<bf16s>(coll, mode, npt, grid, block, p, s);
    default:
      return hipErrorInvalidValue;
  }
}

template <class S>
static hipError_t batch_pair(unsigned blocks, const BatchParams& bp, hipStream_t s) {
  hipLaunchKernelGGL((batch_kernel<S, COLL_BGK>), dim3(blocks), dim3(256), 0, s, bp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((batch_kernel<S, COLL_MRT>), dim3(blocks), dim3(256), 0, s, bp);
  return hipGetLastError();
}

hipError_t launch_batch_lowp(int storage, unsigned blocks, const BatchParams& bp, hipStream_t s) {
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
