// util_kernels.hpp -- launchers for the initialisation / conversion / macro /
// reduction kernels of util_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hlbm.h"

namespace hlbm {

struct ObstacleDev {
  int32_t kind;   // 0 ball, 1 wall, 2 quadric
  int32_t lo[3];  // candidate node box (inclusive), already clipped to the grid
  int32_t hi[3];
  int32_t pad;
  double p0[3];   // ball center / wall bmin
  double p1[3];   // wall bmax
  double r2;      // ball radius squared
  double q[16];   // quadric matrix, row-major
};

// The device form of one obstacle on an nn[0] x nn[1] x nn[2] node grid of
// spacing dx; false when its candidate box misses the grid.  Balls and walls
// take their box from int(bound/dx) (update_obstacles.cu:58-63); a quadric
// covers everything (Quadric::covered, obstacle.hpp:121-125).
inline bool obstacle_to_dev(const hlbm_obstacle& o, const int nn[3], double dx, ObstacleDev& od) {
  od = ObstacleDev{};
  od.kind = o.kind;
  bool inside = true;
  for (int d = 0; d < 3; d++) {
    if (o.kind == 2) {
      od.lo[d] = 0;
      od.hi[d] = nn[d] - 1;
      continue;
    }
    const double bmin = o.kind == 0 ? o.p0[d] - o.radius : o.p0[d];
    const double bmax = o.kind == 0 ? o.p0[d] + o.radius : o.p1[d];
    // int(b/dx) truncation; the global node index is the node index
    const int a = int(bmin / dx), b = int(bmax / dx);
    if (b < 0 || a > nn[d] - 1) inside = false;
    od.lo[d] = a > 0 ? a : 0;
    od.hi[d] = b < nn[d] - 1 ? b : nn[d] - 1;
    od.p0[d] = o.p0[d];
    od.p1[d] = o.p1[d];
  }
  od.r2 = o.radius * o.radius;
  for (int k = 0; k < 16; k++) od.q[k] = o.q[k];
  return inside;
}

// per-plane statistics of plane_velocity_profile
struct ProfileStat {
  double sum, vmin, vmax;
  int32_t count, pad;
};

hipError_t launch_rasterize(uint8_t* flags, int nx, int ny, int nz, double dx, const ObstacleDev* obs, int nobs,
                            hipStream_t s);
hipError_t launch_build_bbm(const uint8_t* flags, int nx, int ny, int nz, const int* per, uint32_t* bbm,
                            hipStream_t s);
hipError_t launch_set_distribution(int storage, void* f, uint64_t stride, int nx, int ny, const int* lo,
                                   const int* hi, double value, hipStream_t s);
hipError_t launch_double_couette(int storage, void* f, uint64_t stride, int nx, int ny, int nz, int dim,
                                 const double* dU, const double* U, hipStream_t s);
hipError_t launch_convert(int storage, bool to_store, void* fplane, double* tmp, uint32_t n, int q, hipStream_t s);
hipError_t launch_macro(int storage, const void* f, uint64_t stride, uint32_t N, const uint8_t* flags, double* m0,
                        double* m1, hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s);
int reduce_parts(uint32_t N);
hipError_t launch_reduce(const double* m0, const double* m1, const uint8_t* flags, uint32_t N, double* part,
                         double* out, hipStream_t s);

// column partials needed by launch_plane_profile (ProfileStat entries)
size_t profile_parts(int nx, int ny, int nz, int dim);
hipError_t launch_plane_profile(const double* m1, const uint8_t* flags, uint32_t N, int nx, int ny, int nz, int dim,
                                double scale, ProfileStat* part, ProfileStat* out, hipStream_t s);
hipError_t launch_line_velocity(const double* m1, uint32_t N, int nx, int ny, int nz, const int* start, int axis,
                                int n, double scale, double* out, hipStream_t s);

}  // namespace hlbm
