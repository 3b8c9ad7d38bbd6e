// util_kernels.hpp -- launchers for the initialisation / conversion / macro /
// reduction kernels of util_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/hlbm.h"

namespace hlbm {

struct ObstacleDev {
  int32_t kind;   // 0 ball, 1 wall, 2 quadric
  int32_t lo[3];  // candidate node box (inclusive), already clipped to the grid
  int32_t hi[3];
  int32_t glo[3]; // the same box clipped to the grid plus one ghost layer on each
  int32_t ghi[3]; // periodic axis (coordinates -1 and n), for the ghost flags
  int32_t pad;
  double p0[3];   // ball center / wall bmin
  double p1[3];   // wall bmax
  double r2;      // ball radius squared
  double q[16];   // quadric matrix, row-major
};

// The device form of one obstacle on an nn[0] x nn[1] x nn[2] node grid of
// spacing dx; false when its candidate box misses the grid (and, with per, the
// ghost layers of the periodic axes).  Balls and walls take their box from
// int(bound/dx) (update_obstacles.cu:58-63); a quadric covers everything
// (Quadric::covered, obstacle.hpp:121-125).  The reference restricts the box to
// the Extend traversal, which includes the ghost layers of periodic axes
// (grid.hpp:406-428, make_domain.hpp:181-188): glo/ghi keep the first of them.
inline bool obstacle_to_dev(const hlbm_obstacle& o, const int nn[3], double dx, ObstacleDev& od,
                            const int* per = nullptr) {
  od = ObstacleDev{};
  od.kind = o.kind;
  bool inside = true, ghost = per != nullptr;
  for (int d = 0; d < 3; d++) {
    const int g = per && per[d] ? 1 : 0;
    if (o.kind == 2) {
      od.lo[d] = 0;
      od.hi[d] = nn[d] - 1;
      od.glo[d] = -g;
      od.ghi[d] = nn[d] - 1 + g;
      continue;
    }
    const double bmin = o.kind == 0 ? o.p0[d] - o.radius : o.p0[d];
    const double bmax = o.kind == 0 ? o.p0[d] + o.radius : o.p1[d];
    // int(b/dx) truncation; the global node index is the node index
    const int a = int(bmin / dx), b = int(bmax / dx);
    if (b < 0 || a > nn[d] - 1) inside = false;
    if (b < -g || a > nn[d] - 1 + g) ghost = false;
    od.lo[d] = a > 0 ? a : 0;
    od.hi[d] = b < nn[d] - 1 ? b : nn[d] - 1;
    od.glo[d] = a > -g ? a : -g;
    od.ghi[d] = b < nn[d] - 1 + g ? b : nn[d] - 1 + g;
    od.p0[d] = o.p0[d];
    od.p1[d] = o.p1[d];
  }
  od.r2 = o.radius * o.radius;
  for (int k = 0; k < 16; k++) od.q[k] = o.q[k];
  return inside || ghost;
}

// per-plane statistics of plane_velocity_profile
struct ProfileStat {
  double sum, vmin, vmax;
  int32_t count, pad;
};

// Kernels over the node grid take the domain's NodeLayout (d3q19.hpp).
hipError_t launch_rasterize(uint8_t* flags, int nx, int ny, int nz, double dx, const ObstacleDev* obs, int nobs,
                            NodeLayout L, hipStream_t s);
hipError_t launch_build_bbm(const uint8_t* flags, int nx, int ny, int nz, const int* per, uint32_t* bbm,
                            NodeLayout L, hipStream_t s);
// Ghost flags: a byte per node of the grid padded by one layer on every axis,
// index ((z+1)*(ny+2) + (y+1))*(nx+2) + (x+1); only the entries outside the
// grid on periodic axes are used.  rasterize_ghost writes the obstacles' walls
// there as the reference rasterizes its ghost nodes (at their own positions);
// ghost_images fills them with the flags of their periodic images.
inline size_t ghost_count(int nx, int ny, int nz) { return (size_t)(nx + 2) * (ny + 2) * (nz + 2); }
hipError_t launch_rasterize_ghost(uint8_t* gflags, int nx, int ny, int nz, const int* per, double dx,
                                  const ObstacleDev* obs, int nobs, hipStream_t s);
hipError_t launch_ghost_images(uint8_t* gflags, const uint8_t* flags, int nx, int ny, int nz, const int* per,
                               NodeLayout L, hipStream_t s);
// Nodes next to a wall across a periodic face (see find_special in
// util_kernels.hip); appends (node, direction mask) pairs through *count.
hipError_t launch_find_special(const uint8_t* flags, const uint8_t* gflags, int nx, int ny, int nz, const int* per,
                               uint32_t* bbm, uint32_t* out_idx, uint32_t* out_mask, uint32_t* count, uint32_t cap,
                               NodeLayout L, hipStream_t s);
// The initial-state and conversion kernels write and read shifted storage
// relative to the rest state sh (RestShift, d3q19.hpp); FP64 storage ignores it.
hipError_t launch_set_distribution(int storage, void* f, uint64_t stride, NodeLayout L, const int* lo,
                                   const int* hi, double value, RestShift sh, hipStream_t s);
hipError_t launch_double_couette(int storage, void* f, uint64_t stride, int nx, int ny, int nz, int dim,
                                 const double* dU, const double* U, RestShift sh, NodeLayout L, hipStream_t s);
hipError_t launch_convert(int storage, bool to_store, void* fplane, double* tmp, uint32_t n, int q, RestShift sh,
                          hipStream_t s);
// a per-node array of esize (1, 2, 4, 8) bytes from layout `from` into `to`
hipError_t launch_permute(void* dst, const void* src, size_t esize, const int n[3], NodeLayout from, NodeLayout to,
                          hipStream_t s);
// shifted storage: g_i += delta[class(i)] over N nodes (a change of rest state)
hipError_t launch_reshift(int storage, void* f, uint64_t stride, uint32_t N, RestShift delta, hipStream_t s);
hipError_t launch_macro(int storage, const void* f, uint64_t stride, uint32_t N, const uint8_t* flags, double* m0,
                        double* m1, hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s);
// 19 planes of n (a multiple of 256) two-byte elements at a and at b (38 n bytes each)
hipError_t launch_placement_probe(const void* a, void* b, uint32_t n, hipStream_t s);
int reduce_parts(uint32_t N);
hipError_t launch_reduce(const double* m0, const double* m1, const uint8_t* flags, uint32_t N, double* part,
                         double* out, hipStream_t s);

// column partials needed by launch_plane_profile (ProfileStat entries)
size_t profile_parts(int nx, int ny, int nz, int dim);
hipError_t launch_plane_profile(const double* m1, const uint8_t* flags, uint32_t N, int nx, int ny, int nz, int dim,
                                double scale, ProfileStat* part, ProfileStat* out, NodeLayout L, hipStream_t s);
// The library-owned column partials of hlbm_plane_profile_noalloc on the
// current device (HLBM_PROFILE_SCRATCH_COLUMNS records); no allocation.
hipError_t profile_scratch(ProfileStat** out);
static_assert(sizeof(ProfileStat) == sizeof(hlbm_plane_stat), "hlbm_plane_stat is ProfileStat's ABI form");
static_assert(offsetof(ProfileStat, vmax) == offsetof(hlbm_plane_stat, vmax) &&
                  offsetof(ProfileStat, count) == offsetof(hlbm_plane_stat, count),
              "hlbm_plane_stat is ProfileStat's ABI form");
hipError_t launch_line_velocity(const double* m1, uint32_t N, int nx, int ny, int nz, const int* start, int axis,
                                int n, double scale, double* out, NodeLayout L, hipStream_t s);

}  // namespace hlbm
