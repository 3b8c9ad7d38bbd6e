// util_kernels.hpp -- launchers for the initialisation / conversion / macro /
// reduction kernels of util_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hlbm {

struct ObstacleDev {
  int32_t kind;   // 0 ball, 1 wall
  int32_t lo[3];  // candidate node box (inclusive), already clipped to the grid
  int32_t hi[3];
  int32_t pad;
  double p0[3];   // ball center / wall bmin
  double p1[3];   // wall bmax
  double r2;      // ball radius squared
};

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
