// util_kernels.hip -- initialisation, conversion, macro and reduction kernels.
// Compiled with -ffp-contract=off (obstacle rasterization and the FP64 macro
// must round exactly as hippoLBM's CPU build does).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "step_kernel.hpp"
#include "util_kernels.hpp"

namespace hlbm {

// ---------------------------------------------------------------------------
// update_obstacles (src/obstacle/update_obstacles.cu:32-71): for each obstacle
// in id order, nodes inside its truncated bounding box (host-computed, already
// clipped to the grid) are tested with Ball::solid (|p-c|^2 <= r^2),
// Wall::solid (strict AABB) or Quadric::solid (quadric form <= 0),
// obstacle.hpp:13-17,61-64,97,136-138.  Position of a node
// is (local + offset) * dx = node * dx (grid.hpp:355-361).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool obstacle_solid(const ObstacleDev& o, double px, double py, double pz) {
  if (o.kind == 0) {
    const double rx = px - o.p0[0], ry = py - o.p0[1], rz = pz - o.p0[2];
    return (rx * rx + ry * ry + rz * rz) <= o.r2;
  }
  if (o.kind == 2) {
    // Quadric::solid: [p,1] Q [p,1]^T <= 0, as p . (Q [p,1]) with the rows of Q first
    double r[4];
    for (int i = 0; i < 4; i++) r[i] = o.q[4 * i] * px + o.q[4 * i + 1] * py + o.q[4 * i + 2] * pz + o.q[4 * i + 3];
    return px * r[0] + py * r[1] + pz * r[2] + r[3] <= 0.0;
  }
  return o.p0[0] < px && px < o.p1[0] && o.p0[1] < py && py < o.p1[1] && o.p0[2] < pz && pz < o.p1[2];
}

__global__ void k_rasterize(uint8_t* flags, int nx, int ny, int nz, double dx, const ObstacleDev* obs, int nobs,
                            NodeLayout L) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  const size_t idx = node_at(L, x, y, z);
  // local index i = node + 2, offset = -2: (double(i) + double(-2)) * dx
  const double px = (double(x + 2) + double(-2)) * dx;
  const double py = (double(y + 2) + double(-2)) * dx;
  const double pz = (double(z + 2) + double(-2)) * dx;
  bool solid = false;
  for (int k = 0; k < nobs; k++) {
    const ObstacleDev o = obs[k];
    if (x < o.lo[0] || x > o.hi[0] || y < o.lo[1] || y > o.hi[1] || z < o.lo[2] || z > o.hi[2]) continue;
    if (obstacle_solid(o, px, py, pz)) solid = true;
  }
  if (solid) flags[idx] = 1;
}

// Ghost nodes of the periodic axes, first layer (coordinate -1 or n on the
// axis): the reference rasterizes them at their own positions, node * dx with
// a negative node index below the grid (update_obstacles.cu:58-67 over the
// Extend box), not as periodic images.
__device__ __forceinline__ size_t ghost_index(int x, int y, int z, int nx, int ny) {
  return ((size_t)(z + 1) * (size_t)(ny + 2) + (size_t)(y + 1)) * (size_t)(nx + 2) + (size_t)(x + 1);
}

// Coordinates of padded-grid thread (tx, ty, tz) when it is a ghost node of the
// periodic axes' first layer; false for Real nodes and for positions outside a
// non-periodic face (not in the Extend box).
__device__ __forceinline__ bool ghost_node(int tx, int ty, int tz, int nx, int ny, int nz, int px, int py, int pz,
                                           int& x, int& y, int& z) {
  x = tx - 1;
  y = ty - 1;
  z = tz - 1;
  if (x > nx || y > ny || z > nz) return false;
  const bool ox = x < 0 || x >= nx, oy = y < 0 || y >= ny, oz = z < 0 || z >= nz;
  if (!(ox || oy || oz)) return false;
  return !((ox && !px) || (oy && !py) || (oz && !pz));
}

__global__ void k_rasterize_ghost(uint8_t* gflags, int nx, int ny, int nz, int pxx, int pyy, int pzz, double dx,
                                  const ObstacleDev* obs, int nobs) {
  int x, y, z;
  if (!ghost_node(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y, blockIdx.z, nx, ny,
                  nz, pxx, pyy, pzz, x, y, z))
    return;
  const double px = (double(x + 2) + double(-2)) * dx;
  const double py = (double(y + 2) + double(-2)) * dx;
  const double pz = (double(z + 2) + double(-2)) * dx;
  bool solid = false;
  for (int k = 0; k < nobs; k++) {
    const ObstacleDev o = obs[k];
    if (x < o.glo[0] || x > o.ghi[0] || y < o.glo[1] || y > o.ghi[1] || z < o.glo[2] || z > o.ghi[2]) continue;
    if (obstacle_solid(o, px, py, pz)) solid = true;
  }
  if (solid) gflags[ghost_index(x, y, z, nx, ny)] = 1;
}

__global__ void k_ghost_images(uint8_t* gflags, const uint8_t* flags, int nx, int ny, int nz, int px, int py, int pz,
                               NodeLayout L) {
  int x, y, z;
  if (!ghost_node(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y, blockIdx.z, nx, ny,
                  nz, px, py, pz, x, y, z))
    return;
  const int ix = x < 0 ? nx - 1 : (x >= nx ? 0 : x);
  const int iy = y < 0 ? ny - 1 : (y >= ny ? 0 : y);
  const int iz = z < 0 ? nz - 1 : (z >= nz ? 0 : z);
  gflags[ghost_index(x, y, z, nx, ny)] = flags[node_at(L, ix, iy, iz)];
}

// ---------------------------------------------------------------------------
// Walls across a periodic face (wall_bounce_back, bounce_back.hpp:163-185).
// The reference bounces a wall node w off its neighbour w + c_i before
// streaming, over the Extend box.  When w lies on the face of a periodic axis
// and w + c_i across it, that neighbour is a ghost node: its flag is the
// ghost's own (rasterized at its position), and its populations are those it
// received in the previous step's streaming, since ghosts are refreshed only
// inside streaming (streaming.cu:107-111).  Those are the populations of the
// ghost's periodic image x at the start of the step (x's pre-collision state)
// as long as no boundary operator changes x.  So for a fluid node x whose
// upstream node x - c_i lies across periodic faces only and wraps to a wall:
//   - ghost of x (x shifted across the face) fluid: f_i(x) receives f_opp(i)(x)
//     of x's previous pre-collision state: bit i of the node's mask;
//   - ghost of x a wall: f_i(x) receives the wall's own streamed value, which
//     the engine does not track (walls hold no populations it follows): bit 31,
//     the host rejects the configuration.
// Nodes with a marked direction are appended to (out_idx, out_mask) and get
// bit 31 in the bulk mask, so the bulk kernels leave them to special_kernel.
// ---------------------------------------------------------------------------
__global__ void k_find_special(const uint8_t* flags, const uint8_t* gflags, int nx, int ny, int nz, int px, int py,
                               int pz, uint32_t* bbm, uint32_t* out_idx, uint32_t* out_mask, uint32_t* count,
                               uint32_t cap, NodeLayout L) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  const bool face = (px && (x == 0 || x == nx - 1)) || (py && (y == 0 || y == ny - 1)) ||
                    (pz && (z == 0 || z == nz - 1));
  if (!face) return;
  const size_t idx = node_at(L, x, y, z);
  if (flags[idx]) return;
  uint32_t m = 0;
  static_for<1, Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    int sx = x - CX(i), sy = y - CY(i), sz = z - CZ(i);
    int gx = x, gy = y, gz = z;  // the ghost of x across the crossed faces
    bool crossed = false;
    if (sx < 0 || sx >= nx) {
      if (!px) return;
      gx = sx < 0 ? x + nx : x - nx;
      sx = sx < 0 ? nx - 1 : 0;
      crossed = true;
    }
    if (sy < 0 || sy >= ny) {
      if (!py) return;
      gy = sy < 0 ? y + ny : y - ny;
      sy = sy < 0 ? ny - 1 : 0;
      crossed = true;
    }
    if (sz < 0 || sz >= nz) {
      if (!pz) return;
      gz = sz < 0 ? z + nz : z - nz;
      sz = sz < 0 ? nz - 1 : 0;
      crossed = true;
    }
    if (!crossed || !flags[node_at(L, sx, sy, sz)]) return;
    m |= gflags[ghost_index(gx, gy, gz, nx, ny)] ? (1u << 31) : (1u << i);
  });
  if (!m) return;
  const uint32_t k = atomicAdd(count, 1u);
  if (k < cap) {
    out_idx[k] = (uint32_t)idx;
    out_mask[k] = m;
  }
  bbm[idx] |= 1u << 31;
}

// set_distribution (set_distribution.hpp:38-41): f_i = value * w_i on a node box.
template <class S>
__global__ void k_set_dist(void* fv, uint64_t stride, NodeLayout L, int lo0, int lo1, int lo2, int hi0, int hi1,
                           int hi2, double value, RestShift sh) {
  const int x = lo0 + blockIdx.x * blockDim.x + threadIdx.x;
  const int y = lo1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int z = lo2 + blockIdx.z;
  if (x > hi0 || y > hi1 || z > hi2) return;
  const size_t idx = node_at(L, x, y, z);
  S* f = (S*)fv;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const double v = value * WD(i);
    if constexpr (StoreOps<S>::shifted)
      StoreOps<S>::st(f, (size_t)i * stride + idx, (float)(v - WD(i) - sh.s[dir_class(i)]));
    else
      StoreOps<S>::st(f, (size_t)i * stride + idx, v);
  }
}

// init_double_couette (src/prepro/include/hippoLBM/prepro/double_couette.hpp:26-45):
// on every node, u = U - dU * (node coordinate along dim), f_i = w_i * feq(1, u)
// in the reference's evaluation order.
template <class S>
__global__ void k_double_couette(void* fv, uint64_t stride, int nx, int ny, int nz, int dim, double dUx, double dUy,
                                 double dUz, double Ux, double Uy, double Uz, RestShift sh, NodeLayout L) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= nx || y >= ny || z >= nz) return;
  const size_t idx = node_at(L, x, y, z);
  const double value = dim == 0 ? x : (dim == 1 ? y : z);
  const double ux = Ux - dUx * value, uy = Uy - dUy * value, uz = Uz - dUz * value;
  const double u_squ = ux * ux + uy * uy + uz * uz;
  S* f = (S*)fv;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const double eu = ux * CX(i) + uy * CY(i) + uz * CZ(i);
    const double v = 1. * WD(i) * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * u_squ);
    if constexpr (StoreOps<S>::shifted)
      StoreOps<S>::st(f, (size_t)i * stride + idx, (float)(v - WD(i) - sh.s[dir_class(i)]));
    else
      StoreOps<S>::st(f, (size_t)i * stride + idx, v);
  }
}

template <class S>
static hipError_t double_couette_t(void* f, uint64_t stride, int nx, int ny, int nz, int dim, const double* dU,
                                   const double* U, RestShift sh, NodeLayout L, hipStream_t s) {
  dim3 b(64, 4, 1), g((nx + 63) / 64, (ny + 3) / 4, nz);
  hipLaunchKernelGGL((k_double_couette<S>), g, b, 0, s, f, stride, nx, ny, nz, dim, dU[0], dU[1], dU[2], U[0], U[1],
                     U[2], sh, L);
  return hipGetLastError();
}

hipError_t launch_double_couette(int storage, void* f, uint64_t stride, int nx, int ny, int nz, int dim,
                                 const double* dU, const double* U, RestShift sh, NodeLayout L, hipStream_t s) {
  switch (storage) {
    case 0:
      return double_couette_t<double>(f, stride, nx, ny, nz, dim, dU, U, sh, L, s);
    case 1:
      return double_couette_t<float>(f, stride, nx, ny, nz, dim, dU, U, sh, L, s);
    case 2:
      return double_couette_t<_Float16>(f, stride, nx, ny, nz, dim, dU, U, sh, L, s);
    case 3:
      return double_couette_t<bf16s>(f, stride, nx, ny, nz, dim, dU, U, sh, L, s);
  }
  return hipErrorInvalidValue;
}

// FP64 plane -> storage plane (adds the shift)
template <class S>
__global__ void k_from_f64(void* fv, const double* src, uint32_t n, int q, double sig) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  S* f = (S*)fv;
  if constexpr (StoreOps<S>::shifted)
    StoreOps<S>::st(f, i, (float)(src[i] - WD(q) - sig));
  else
    StoreOps<S>::st(f, i, src[i]);
}

template <class S>
__global__ void k_to_f64(double* dst, const void* fv, uint32_t n, int q, double sig) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const S* f = (const S*)fv;
  if constexpr (StoreOps<S>::shifted)
    dst[i] = (double)StoreOps<S>::ld(f, i) + (WD(q) + sig);
  else
    dst[i] = StoreOps<S>::ld(f, i);
}

// macro_variables on the current (pre-collision) state, macro_variables.hpp:48-79.
template <class S>
__global__ void k_macro(const void* fv, uint64_t stride, uint32_t N, const uint8_t* flags, double* m0, double* m1) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const S* f = (const S*)fv;
  if (flags[idx]) {
    m1[idx] = 0.0;
    m1[(size_t)N + idx] = 0.0;
    m1[2 * (size_t)N + idx] = 0.0;
    return;
  }
  if constexpr (!StoreOps<S>::shifted) {
    double fl[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) fl[i] = StoreOps<S>::ld(f, (size_t)i * stride + idx);
    double rho, ux, uy, uz;
    macro_f64(fl, rho, ux, uy, uz);
    m0[idx] = rho;
    m1[idx] = ux;
    m1[(size_t)N + idx] = uy;
    m1[2 * (size_t)N + idx] = uz;
  } else {
    float dr = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
#pragma unroll
    for (int i = 0; i < Q; i++) {
      const float g = StoreOps<S>::ld(f, (size_t)i * stride + idx);
      dr += g;
      jx += float(CX(i)) * g;
      jy += float(CY(i)) * g;
      jz += float(CZ(i)) * g;
    }
    const float rho = 1.0f + dr;
    m0[idx] = rho;
    m1[idx] = jx / rho;
    m1[(size_t)N + idx] = jy / rho;
    m1[2 * (size_t)N + idx] = jz / rho;
  }
}

// simulation_state (simulation_state.hpp:47-79): per-block partial sums in
// FP64, then one block folds the partials in a fixed order (deterministic).
constexpr int RED_BLOCK = 256;

__global__ void __launch_bounds__(RED_BLOCK)
    k_reduce_partial(const double* m0, const double* m1, const uint8_t* flags, uint32_t N, double* part) {
  __shared__ double s_sum[RED_BLOCK], s_min[RED_BLOCK], s_max[RED_BLOCK], s_jx[RED_BLOCK], s_jy[RED_BLOCK],
      s_jz[RED_BLOCK], s_cnt[RED_BLOCK];
  double sum = 0.0, vmin = 1.7976931348623157e308, vmax = -1.7976931348623157e308, jx = 0.0, jy = 0.0, jz = 0.0,
         cnt = 0.0;
  const uint32_t stride = gridDim.x * RED_BLOCK;
  for (uint32_t i = blockIdx.x * RED_BLOCK + threadIdx.x; i < N; i += stride) {
    const double r = m0[i];
    const double ux = m1[i], uy = m1[(size_t)N + i], uz = m1[2 * (size_t)N + i];
    const double v = sqrt(ux * ux + uy * uy + uz * uz);
    sum += r;
    vmin = fmin(vmin, v);
    vmax = fmax(vmax, v);
    jx += r * ux;
    jy += r * uy;
    jz += r * uz;
    cnt += flags[i] ? 0.0 : 1.0;
  }
  const int t = threadIdx.x;
  s_sum[t] = sum;
  s_min[t] = vmin;
  s_max[t] = vmax;
  s_jx[t] = jx;
  s_jy[t] = jy;
  s_jz[t] = jz;
  s_cnt[t] = cnt;
  __syncthreads();
  for (int w = RED_BLOCK / 2; w > 0; w >>= 1) {
    if (t < w) {
      s_sum[t] += s_sum[t + w];
      s_min[t] = fmin(s_min[t], s_min[t + w]);
      s_max[t] = fmax(s_max[t], s_max[t + w]);
      s_jx[t] += s_jx[t + w];
      s_jy[t] += s_jy[t + w];
      s_jz[t] += s_jz[t + w];
      s_cnt[t] += s_cnt[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    double* o = part + 7 * (size_t)blockIdx.x;
    o[0] = s_sum[0];
    o[1] = s_min[0];
    o[2] = s_max[0];
    o[3] = s_jx[0];
    o[4] = s_jy[0];
    o[5] = s_jz[0];
    o[6] = s_cnt[0];
  }
}

__global__ void k_reduce_final(const double* part, int nparts, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double r[7] = {0.0, 1.7976931348623157e308, -1.7976931348623157e308, 0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < nparts; b++) {
    const double* o = part + 7 * (size_t)b;
    r[0] += o[0];
    r[1] = fmin(r[1], o[1]);
    r[2] = fmax(r[2], o[2]);
    r[3] += o[3];
    r[4] += o[4];
    r[5] += o[5];
    r[6] += o[6];
  }
  for (int k = 0; k < 7; k++) out[k] = r[k];
}

// plane_velocity_profile (src/check/plane_velocity_profile.cu:81-101): per
// plane perpendicular to `dim`, over the fluid nodes, |scale*u| summed, counted,
// min and max.  Pass 1: one thread per (x, a) column walks the other axis b,
// where (a, b) = (z, y) for dim x or z and (y, z) for dim y, so every thread
// feeds a single plane and every row of a wave is one coalesced read; it
// writes the column's statistics to part[a*nx + x].  Pass 2: one block per
// plane folds its columns in a fixed order (deterministic sums).
__global__ void __launch_bounds__(256)
    k_profile_cols(const double* m1, const uint8_t* flags, uint32_t N, int nx, int ny, int nz, int dim,
                   double scale, ProfileStat* part, NodeLayout L) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (x >= nx) return;
  size_t idx, step;
  int nb;
  if (dim == 1) {  // a = y, b = z
    idx = node_at(L, x, a, 0);
    step = L.st[2];
    nb = nz;
  } else {  // a = z, b = y
    idx = node_at(L, x, 0, a);
    step = L.st[1];
    nb = ny;
  }
  double sum = 0.0, vmin = 1.7976931348623157e308, vmax = -1.7976931348623157e308;
  int cnt = 0;
  // every load is issued (walls are selected out afterwards), so the unrolled
  // iterations keep their loads in flight together
#pragma unroll 8
  for (int b = 0; b < nb; b++, idx += step) {
    const bool fluid = flags[idx] == 0;
    const double ux = scale * m1[idx], uy = scale * m1[(size_t)N + idx], uz = scale * m1[2 * (size_t)N + idx];
    const double v = sqrt(ux * ux + uy * uy + uz * uz);
    sum += fluid ? v : 0.0;
    cnt += fluid ? 1 : 0;
    vmin = fluid ? fmin(vmin, v) : vmin;
    vmax = fluid ? fmax(vmax, v) : vmax;
  }
  ProfileStat o;
  o.sum = sum;
  o.vmin = vmin;
  o.vmax = vmax;
  o.count = cnt;
  o.pad = 0;
  part[(size_t)a * nx + x] = o;
}

__global__ void __launch_bounds__(256)
    k_profile_planes(const ProfileStat* part, int nx, int na, int dim, ProfileStat* out) {
  __shared__ double s_sum[256], s_min[256], s_max[256];
  __shared__ int s_cnt[256];
  const int p = blockIdx.x, t = threadIdx.x;
  // dim x: plane p = column x of every a; dims y/z: plane p = row a of every x
  const int n = dim == 0 ? na : nx;
  const size_t first = dim == 0 ? (size_t)p : (size_t)p * nx;
  const size_t step = dim == 0 ? (size_t)nx : 1;
  double sum = 0.0, vmin = 1.7976931348623157e308, vmax = -1.7976931348623157e308;
  int cnt = 0;
  for (int k = t; k < n; k += 256) {
    const ProfileStat c = part[first + (size_t)k * step];
    sum += c.sum;
    cnt += c.count;
    vmin = fmin(vmin, c.vmin);
    vmax = fmax(vmax, c.vmax);
  }
  s_sum[t] = sum;
  s_min[t] = vmin;
  s_max[t] = vmax;
  s_cnt[t] = cnt;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s_sum[t] += s_sum[t + w];
      s_min[t] = fmin(s_min[t], s_min[t + w]);
      s_max[t] = fmax(s_max[t], s_max[t + w]);
      s_cnt[t] += s_cnt[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    ProfileStat o;
    o.sum = s_sum[0];
    o.vmin = s_min[0];
    o.vmax = s_max[0];
    o.count = s_cnt[0];
    o.pad = 0;
    out[p] = o;
  }
}

// plot_line_velocity (src/check/plot_over_line.cu:77-90): scale*u at the
// nodes start + k*e_axis; nodes outside the domain give 0.
__global__ void k_line_velocity(const double* m1, uint32_t N, int nx, int ny, int nz, int x0, int y0, int z0,
                                int axis, int n, double scale, double* out, NodeLayout L) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int x = x0 + (axis == 0 ? k : 0), y = y0 + (axis == 1 ? k : 0), z = z0 + (axis == 2 ? k : 0);
  double u[3] = {0.0, 0.0, 0.0};
  if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz) {
    const size_t idx = node_at(L, x, y, z);
    for (int c = 0; c < 3; c++) u[c] = scale * m1[(size_t)c * N + idx];
  }
  for (int c = 0; c < 3; c++) out[3 * (size_t)k + c] = u[c];
}

// Streaming copy, the HBM reference point for the roofline and for scaling
// the FETCH_SIZE / WRITE_SIZE counters: one 16-byte element per thread, one
// workgroup per 4 KiB, non-temporal stores.  Of the forms measured on MI355X
// (grid-stride over 256K workgroups 5.4-5.8 TB/s; 2, 4 or 8 elements per
// thread 5.6-5.8; non-temporal loads; 4- and 8-byte elements) this one is the
// fastest: 6.3 TB/s for 4 and 16 GiB.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) hbm_copy_kernel(u32x4* __restrict__ dst, const u32x4* __restrict__ src,
                                                       uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(src[i], dst + i);
}

// Placement probe: the memory traffic of a 16-bit D3Q19 pull step without its
// arithmetic.  19 planes of n two-byte elements lie back to back in `a` and in
// `b`; element i of plane q in b takes element i + o_q of plane q in a, with
// o_q the neighbour offsets of a 512 x 512 x n/2^18 lattice (x fastest),
// clamped to the plane.
__constant__ int32_t k_probe_off[19] = {0,      1,       -1,      512,     -512,    262144,  -262144,
                                        513,    -511,    511,     -513,    262145,  -262143, 262143,
                                        -262145, 262656, -261632, 261632,  -262656};
__global__ void __launch_bounds__(256) k_placement_probe(const uint16_t* __restrict__ a, uint16_t* __restrict__ b,
                                                         uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int q = 0; q < 19; q++) {
    const int64_t j = min<int64_t>(max<int64_t>((int64_t)i + k_probe_off[q], 0), (int64_t)n - 1);
    b[(uint64_t)q * n + i] = a[(uint64_t)q * n + (uint64_t)j];
  }
}

// ---------------------------------------------------------------------------
// Bounce-back mask of the bulk kernel (wall_bounce_back, bounce_back.hpp:163-185):
// bit 0 = the node is a wall; for a fluid node, bit i (1..18) = the upstream
// node x - c_i (periodic wrap applied) is a wall.  An upstream node outside a
// non-periodic face sets no bit: the face bounce-back covers it.
// ---------------------------------------------------------------------------
__global__ void k_build_bbm(const uint8_t* flags, int nx, int ny, int nz, int px, int py, int pz, uint32_t* bbm,
                            NodeLayout L) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  const size_t idx = node_at(L, x, y, z);
  uint32_t m = 0;
  if (flags[idx]) {
    m = 1u;
  } else {
    static_for<1, Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      int sx = x - CX(i), sy = y - CY(i), sz = z - CZ(i);
      if (sx < 0 || sx >= nx) {
        if (!px) return;
        sx = sx < 0 ? nx - 1 : 0;
      }
      if (sy < 0 || sy >= ny) {
        if (!py) return;
        sy = sy < 0 ? ny - 1 : 0;
      }
      if (sz < 0 || sz >= nz) {
        if (!pz) return;
        sz = sz < 0 ? nz - 1 : 0;
      }
      if (flags[node_at(L, sx, sy, sz)]) m |= 1u << i;
    });
  }
  bbm[idx] = m;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
hipError_t launch_build_bbm(const uint8_t* flags, int nx, int ny, int nz, const int* per, uint32_t* bbm,
                            NodeLayout L, hipStream_t s) {
  dim3 b(64, 4, 1), g((nx + 63) / 64, (ny + 3) / 4, nz);
  hipLaunchKernelGGL(k_build_bbm, g, b, 0, s, flags, nx, ny, nz, per[0], per[1], per[2], bbm, L);
  return hipGetLastError();
}

hipError_t launch_rasterize(uint8_t* flags, int nx, int ny, int nz, double dx, const ObstacleDev* obs, int nobs,
                            NodeLayout L, hipStream_t s) {
  dim3 b(64, 4, 1), g((nx + 63) / 64, (ny + 3) / 4, nz);
  hipLaunchKernelGGL(k_rasterize, g, b, 0, s, flags, nx, ny, nz, dx, obs, nobs, L);
  return hipGetLastError();
}

static dim3 padded_grid(int nx, int ny, int nz) { return dim3((nx + 2 + 63) / 64, (ny + 2 + 3) / 4, nz + 2); }

hipError_t launch_rasterize_ghost(uint8_t* gflags, int nx, int ny, int nz, const int* per, double dx,
                                  const ObstacleDev* obs, int nobs, hipStream_t s) {
  hipLaunchKernelGGL(k_rasterize_ghost, padded_grid(nx, ny, nz), dim3(64, 4, 1), 0, s, gflags, nx, ny, nz, per[0],
                     per[1], per[2], dx, obs, nobs);
  return hipGetLastError();
}

hipError_t launch_ghost_images(uint8_t* gflags, const uint8_t* flags, int nx, int ny, int nz, const int* per,
                               NodeLayout L, hipStream_t s) {
  hipLaunchKernelGGL(k_ghost_images, padded_grid(nx, ny, nz), dim3(64, 4, 1), 0, s, gflags, flags, nx, ny, nz, per[0],
                     per[1], per[2], L);
  return hipGetLastError();
}

hipError_t launch_find_special(const uint8_t* flags, const uint8_t* gflags, int nx, int ny, int nz, const int* per,
                               uint32_t* bbm, uint32_t* out_idx, uint32_t* out_mask, uint32_t* count, uint32_t cap,
                               NodeLayout L, hipStream_t s) {
  dim3 b(64, 4, 1), g((nx + 63) / 64, (ny + 3) / 4, nz);
  hipLaunchKernelGGL(k_find_special, g, b, 0, s, flags, gflags, nx, ny, nz, per[0], per[1], per[2], bbm, out_idx,
                     out_mask, count, cap, L);
  return hipGetLastError();
}

template <class S>
static hipError_t set_dist_t(void* f, uint64_t stride, NodeLayout L, const int* lo, const int* hi, double value,
                             RestShift sh, hipStream_t s) {
  const int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
  dim3 b(64, 4, 1), g((ex + 63) / 64, (ey + 3) / 4, ez);
  hipLaunchKernelGGL((k_set_dist<S>), g, b, 0, s, f, stride, L, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2],
                     value, sh);
  return hipGetLastError();
}

hipError_t launch_set_distribution(int storage, void* f, uint64_t stride, NodeLayout L, const int* lo,
                                   const int* hi, double value, RestShift sh, hipStream_t s) {
  switch (storage) {
    case 0:
      return set_dist_t<double>(f, stride, L, lo, hi, value, sh, s);
    case 1:
      return set_dist_t<float>(f, stride, L, lo, hi, value, sh, s);
    case 2:
      return set_dist_t<_Float16>(f, stride, L, lo, hi, value, sh, s);
    case 3:
      return set_dist_t<bf16s>(f, stride, L, lo, hi, value, sh, s);
  }
  return hipErrorInvalidValue;
}

template <class S>
static hipError_t conv_t(bool to_store, void* f, double* tmp, uint32_t n, int q, double sig, hipStream_t s) {
  dim3 b(256), g((n + 255) / 256);
  if (to_store)
    hipLaunchKernelGGL((k_from_f64<S>), g, b, 0, s, f, (const double*)tmp, n, q, sig);
  else
    hipLaunchKernelGGL((k_to_f64<S>), g, b, 0, s, tmp, (const void*)f, n, q, sig);
  return hipGetLastError();
}

hipError_t launch_convert(int storage, bool to_store, void* fplane, double* tmp, uint32_t n, int q, RestShift sh,
                          hipStream_t s) {
  const double sig = sh.s[dir_class(q)];
  switch (storage) {
    case 0:
      return conv_t<double>(to_store, fplane, tmp, n, q, sig, s);
    case 1:
      return conv_t<float>(to_store, fplane, tmp, n, q, sig, s);
    case 2:
      return conv_t<_Float16>(to_store, fplane, tmp, n, q, sig, s);
    case 3:
      return conv_t<bf16s>(to_store, fplane, tmp, n, q, sig, s);
  }
  return hipErrorInvalidValue;
}

// One per-node array (esize bytes per node) from layout `from` into layout
// `to`; threads walk the source order (coalesced reads).
template <class T>
__global__ void k_permute(T* dst, const T* src, uint32_t N, int n0, int n1, NodeLayout from, NodeLayout to) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t n01 = (uint32_t)n0 * (uint32_t)n1;
  const uint32_t c2 = i / n01, r = i - c2 * n01, c1 = r / (uint32_t)n0, c0 = r - c1 * (uint32_t)n0;
  int c[3];
  c[0] = from.ord[0] == 0 ? c0 : (from.ord[1] == 0 ? c1 : c2);
  c[1] = from.ord[0] == 1 ? c0 : (from.ord[1] == 1 ? c1 : c2);
  c[2] = from.ord[0] == 2 ? c0 : (from.ord[1] == 2 ? c1 : c2);
  dst[node_at(to, c[0], c[1], c[2])] = src[i];
}

hipError_t launch_permute(void* dst, const void* src, size_t esize, const int n[3], NodeLayout from, NodeLayout to,
                          hipStream_t s) {
  const uint32_t N = (uint32_t)n[0] * (uint32_t)n[1] * (uint32_t)n[2];
  const int n0 = n[from.ord[0]], n1 = n[from.ord[1]];
  dim3 b(256), g((N + 255) / 256);
  switch (esize) {
    case 1:
      hipLaunchKernelGGL((k_permute<uint8_t>), g, b, 0, s, (uint8_t*)dst, (const uint8_t*)src, N, n0, n1, from, to);
      break;
    case 2:
      hipLaunchKernelGGL((k_permute<uint16_t>), g, b, 0, s, (uint16_t*)dst, (const uint16_t*)src, N, n0, n1, from, to);
      break;
    case 4:
      hipLaunchKernelGGL((k_permute<uint32_t>), g, b, 0, s, (uint32_t*)dst, (const uint32_t*)src, N, n0, n1, from, to);
      break;
    case 8:
      hipLaunchKernelGGL((k_permute<uint64_t>), g, b, 0, s, (uint64_t*)dst, (const uint64_t*)src, N, n0, n1, from, to);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Shifted storage moving to another rest state: g += delta[class(i)].
template <class S>
__global__ void k_reshift(void* fv, uint64_t stride, uint32_t N, RestShift delta) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  S* f = (S*)fv;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const size_t k = (size_t)i * stride + idx;
    StoreOps<S>::st(f, k, StoreOps<S>::ld(f, k) + (float)delta.s[dir_class(i)]);
  }
}

hipError_t launch_reshift(int storage, void* f, uint64_t stride, uint32_t N, RestShift delta, hipStream_t s) {
  dim3 b(256), g((N + 255) / 256);
  switch (storage) {
    case 1:
      hipLaunchKernelGGL((k_reshift<float>), g, b, 0, s, f, stride, N, delta);
      break;
    case 2:
      hipLaunchKernelGGL((k_reshift<_Float16>), g, b, 0, s, f, stride, N, delta);
      break;
    case 3:
      hipLaunchKernelGGL((k_reshift<bf16s>), g, b, 0, s, f, stride, N, delta);
      break;
    default:
      return hipSuccess;  // FP64 storage is not shifted
  }
  return hipGetLastError();
}

hipError_t launch_macro(int storage, const void* f, uint64_t stride, uint32_t N, const uint8_t* flags, double* m0,
                        double* m1, hipStream_t s) {
  dim3 b(256), g((N + 255) / 256);
  switch (storage) {
    case 0:
      hipLaunchKernelGGL((k_macro<double>), g, b, 0, s, f, stride, N, flags, m0, m1);
      break;
    case 1:
      hipLaunchKernelGGL((k_macro<float>), g, b, 0, s, f, stride, N, flags, m0, m1);
      break;
    case 2:
      hipLaunchKernelGGL((k_macro<_Float16>), g, b, 0, s, f, stride, N, flags, m0, m1);
      break;
    case 3:
      hipLaunchKernelGGL((k_macro<bf16s>), g, b, 0, s, f, stride, N, flags, m0, m1);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s) {
  const uint64_t n = bytes / 16;
  const uint64_t blocks = (n + 255) / 256;
  if (blocks == 0 || blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(hbm_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (u32x4*)dst, (const u32x4*)src, n);
  return hipGetLastError();
}

hipError_t launch_placement_probe(const void* a, void* b, uint32_t n, hipStream_t s) {
  if (n == 0 || n % 256 != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_placement_probe, dim3(n / 256), dim3(256), 0, s, (const uint16_t*)a, (uint16_t*)b, n);
  return hipGetLastError();
}

int reduce_parts(uint32_t N) {
  int parts = (int)((N + RED_BLOCK - 1) / RED_BLOCK);
  return parts > 1024 ? 1024 : (parts < 1 ? 1 : parts);
}

size_t profile_parts(int nx, int ny, int nz, int dim) { return (size_t)nx * (dim == 1 ? ny : nz); }

hipError_t launch_plane_profile(const double* m1, const uint8_t* flags, uint32_t N, int nx, int ny, int nz, int dim,
                                double scale, ProfileStat* part, ProfileStat* out, NodeLayout L, hipStream_t s) {
  const int na = dim == 1 ? ny : nz;
  hipLaunchKernelGGL(k_profile_cols, dim3((unsigned)((nx + 255) / 256), (unsigned)na), dim3(256), 0, s, m1, flags, N,
                     nx, ny, nz, dim, scale, part, L);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int planes = dim == 0 ? nx : na;
  hipLaunchKernelGGL(k_profile_planes, dim3((unsigned)planes), dim3(256), 0, s, (const ProfileStat*)part, nx, na, dim,
                     out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Column partials of plane_velocity_profile for hlbm_plane_profile_noalloc: a
// file-scope device array, one copy per GPU for the life of the process.  In
// host code its name is only a stand-in; kernels get the address that
// hipGetSymbolAddress returns on the current device, looked up once per device.
// k_profile_cols writes every record it covers before k_profile_planes reads
// it, so nothing has to start at zero.
// ---------------------------------------------------------------------------
static __device__ __align__(256) ProfileStat scratch[HLBM_PROFILE_SCRATCH_COLUMNS];

hipError_t profile_scratch(ProfileStat** out) {
  constexpr int MAX_DEV = 64;
  static std::mutex mu;
  static ProfileStat* ptr[MAX_DEV] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= MAX_DEV) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    ptr[dev] = (ProfileStat*)p;
  }
  *out = ptr[dev];
  return hipSuccess;
}

hipError_t launch_line_velocity(const double* m1, uint32_t N, int nx, int ny, int nz, const int* start, int axis,
                                int n, double scale, double* out, NodeLayout L, hipStream_t s) {
  hipLaunchKernelGGL(k_line_velocity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m1, N, nx, ny, nz, start[0],
                     start[1], start[2], axis, n, scale, out, L);
  return hipGetLastError();
}

hipError_t launch_reduce(const double* m0, const double* m1, const uint8_t* flags, uint32_t N, double* part,
                         double* out, hipStream_t s) {
  const int parts = reduce_parts(N);
  hipLaunchKernelGGL(k_reduce_partial, dim3(parts), dim3(RED_BLOCK), 0, s, m0, m1, flags, N, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, s, (const double*)part, parts, out);
  return hipGetLastError();
}

}  // namespace hlbm
