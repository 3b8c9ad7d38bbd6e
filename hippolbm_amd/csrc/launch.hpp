// launch.hpp -- host-side launchers for the step kernels (one translation unit
// per floating-point contract mode).
#pragma once
#include <cmath>

#include <hip/hip_runtime.h>

#include "fp32_constants.hpp"
#include "step_kernel.hpp"

namespace hlbm {

// Blocks of 256 threads covering every node of the program planes.
// The bulk kernel's fast variant addresses a plane with 32-bit byte offsets
// and reads obstacle bounce-back from the precomputed per-node mask.  The
// general variant (64-bit addresses, neighbour flags read directly) runs when
// some offset, (N-1)*element size, needs more than 32 bits, when wall
// bounce-back is active and no mask was supplied, or when start_iteration
// boxes are set (the 32-bit bulk path leaves them out).
inline bool wide_offsets(const StepParams& p, size_t esize) {
  if (p.force_wide) return true;
  if (p.nset) return true;  // start_iteration boxes are applied by finish_node, which the wide variant uses
  if (p.any_wall && p.wall_bb && !p.bbm) return true;
  return (uint64_t)(p.N - 1u) * esize > 0xFFFFFFFFull;
}

// Uniform constants of the FP32-compute forms, derived from tau and c_i . F
// (bgk_fp32_constants, mrt_fp32_rates: fp32_constants.hpp).
// c_1, c_3, c_5 are the unit vectors +x, +y, +z, so ef[1], ef[3], ef[5] are F.
static_assert(CX(1) == 1 && CY(3) == 1 && CZ(5) == 1, "axis directions");
inline void set_fp32_constants(StepParams& p) {
  p.has_force = 0;
  for (int i = 0; i < Q; i++)
    if (p.ef[i] != 0.0) p.has_force = 1;
  p.f3[0] = 3.0f * (float)p.ef[1];
  p.f3[1] = 3.0f * (float)p.ef[3];
  p.f3[2] = 3.0f * (float)p.ef[5];
  bgk_fp32_constants(p.tau, p.k1, p.kw);
  mrt_fp32_rates(p.mrt_custom, p.mrt, p.inv_tau, p.mrtf);
  for (int side = 0; side < 2; side++)
    for (int j = 0; j < 5; j++) p.cavf[side][j] = (float)p.cav[side][j];
}

// Program-plane launches use small workgroups: a few planes of nodes are too
// few 256-thread workgroups to reach every CU, and each lane's accesses are
// latency-bound (on an x-normal plane every lane touches its own cache lines).
constexpr unsigned BOUNDARY_BLOCK = 64;

inline unsigned boundary_blocks(const StepParams& p) {
  uint64_t n = 0;
  for (int pl = 0; pl < 6; pl++)
    if ((p.prog_planes >> pl) & 1) {
      const int dim = plane_dim(pl);
      n += dim == 0 ? (uint64_t)p.ny * p.nz : (dim == 1 ? (uint64_t)p.nx * p.nz : (uint64_t)p.nx * p.ny);
    }
  return (unsigned)((n + BOUNDARY_BLOCK - 1) / BOUNDARY_BLOCK);
}

// One step: outflow capture (in-place AA), program-plane kernel, bulk kernel,
// special_kernel (nodes next to a wall across a periodic face).  The
// program-plane nodes, the special nodes and the bulk touch disjoint nodes
// (A/B) or disjoint slots (AA), so their order is free.
template <class S, int COLL, int MODE, int NPT, bool WIDE, bool PERM>
hipError_t step_go(dim3 grid, dim3 block, const StepParams& p, hipStream_t s) {
  hipError_t e = hipSuccess;
  if constexpr (MODE == MODE_AA_EVEN || MODE == MODE_AA_ODD) {
    // outflow values first: the step below updates nodes in place
    if ((p.parts & PART_CAPTURE) && p.oflow && !p.skip_program) {
      hipLaunchKernelGGL((outflow_capture_kernel<S, MODE>), dim3((p.oflow_n + 255) / 256), dim3(256), 0, s, p);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  const bool planes = MODE != MODE_COLLIDE && p.prog_planes && (p.parts & PART_BOUNDARY);
  if (planes && p.boundary_first) {
    hipLaunchKernelGGL((boundary_kernel<S, COLL, MODE>), dim3(boundary_blocks(p)), dim3(BOUNDARY_BLOCK), 0, s, p);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (p.parts & PART_BULK) {
    if (MODE != MODE_AA_FIX) {  // AA_FIX only touches program-plane nodes
      hipLaunchKernelGGL((step_kernel<S, COLL, MODE, NPT, WIDE, PERM>), grid, block, 0, s, p);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    if (p.nspec) {  // special nodes on program planes are left out of the program-plane kernel too
      hipLaunchKernelGGL((special_kernel<S, COLL, MODE>), dim3((p.nspec + 63) / 64), dim3(64), 0, s, p);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  if (!planes || p.boundary_first) return e;
  hipLaunchKernelGGL((boundary_kernel<S, COLL, MODE>), dim3(boundary_blocks(p)), dim3(BOUNDARY_BLOCK), 0, s, p);
  return hipGetLastError();
}

// A layout other than the default runs one node per thread (npt 1).
template <class S, int COLL, int MODE>
hipError_t step_go_npt(int npt, dim3 grid, dim3 block, const StepParams& p, hipStream_t s) {
  const bool perm = !(p.lay.ord[0] == 0 && p.lay.ord[1] == 1);
  if (perm)
    return wide_offsets(p, sizeof(S)) ? step_go<S, COLL, MODE, 1, true, true>(grid, block, p, s)
                                      : step_go<S, COLL, MODE, 1, false, true>(grid, block, p, s);
  if (wide_offsets(p, sizeof(S))) return step_go<S, COLL, MODE, 1, true, false>(grid, block, p, s);
  return npt == 2 ? step_go<S, COLL, MODE, 2, false, false>(grid, block, p, s)
                  : step_go<S, COLL, MODE, 1, false, false>(grid, block, p, s);
}

template <class S>
hipError_t step_dispatch(int coll, int mode, int npt, dim3 grid, dim3 block, const StepParams& p, hipStream_t s) {
  if (mode == MODE_STREAM) return step_go_npt<S, COLL_BGK, MODE_STREAM>(npt, grid, block, p, s);  // no collision
  if (mode == MODE_AA_STREAM) return step_go_npt<S, COLL_BGK, MODE_AA_STREAM>(npt, grid, block, p, s);
  if (mode == MODE_AA_FIX) return step_go_npt<S, COLL_BGK, MODE_AA_FIX>(npt, grid, block, p, s);
  if (mode == MODE_AA_EVEN)
    return coll == COLL_BGK ? step_go_npt<S, COLL_BGK, MODE_AA_EVEN>(npt, grid, block, p, s)
                            : step_go_npt<S, COLL_MRT, MODE_AA_EVEN>(npt, grid, block, p, s);
  if (mode == MODE_AA_ODD)
    return coll == COLL_BGK ? step_go_npt<S, COLL_BGK, MODE_AA_ODD>(npt, grid, block, p, s)
                            : step_go_npt<S, COLL_MRT, MODE_AA_ODD>(npt, grid, block, p, s);
  if (coll == COLL_BGK) {
    if (mode == MODE_FUSED) return step_go_npt<S, COLL_BGK, MODE_FUSED>(npt, grid, block, p, s);
    return step_go_npt<S, COLL_BGK, MODE_COLLIDE>(npt, grid, block, p, s);
  }
  if (mode == MODE_FUSED) return step_go_npt<S, COLL_MRT, MODE_FUSED>(npt, grid, block, p, s);
  return step_go_npt<S, COLL_MRT, MODE_COLLIDE>(npt, grid, block, p, s);
}

// FP64 storage, FP64 compute, no contraction (step_f64.hip)
hipError_t launch_step_f64(int coll, int mode, int npt, dim3 grid, dim3 block, const StepParams& p, hipStream_t s);
// FP32 compute on shifted FP32/FP16/BF16 storage (step_lowp.hip); storage = hlbm_storage
hipError_t launch_step_lowp(int storage, int coll, int mode, int npt, dim3 grid, dim3 block, const StepParams& p,
                            hipStream_t s);

// batched domains: one launch of the BGK or the MRT kernel for one step with
// `blocks` workgroups (any number is correct: a workgroup walks units
// grid-stride)
hipError_t launch_batch_f64(int coll, unsigned blocks, const BatchParams& bp, hipStream_t s);
hipError_t launch_batch_lowp(int storage, int coll, unsigned blocks, const BatchParams& bp, hipStream_t s);

}  // namespace hlbm
