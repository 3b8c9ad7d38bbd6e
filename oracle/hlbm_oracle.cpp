// hlbm_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of hippoLBM's D3Q19 step chain, used as the checker for
// the HIP engine.  It is never linked into, loaded by or called from the
// product library (libhlbm.so); only tests/, the smoke check and the CPU
// baseline leg of bench.py load it.
//
// The restatement is deliberately literal: it keeps the reference's storage
// (one lattice, SoA f[N_all*i + idx] over the "All" box with two ghost layers
// per side, x fastest), its in-place swap streaming, its single-rank periodic
// ghost copy, its index lists (all/real/planes) and its per-operator loop and
// floating-point evaluation order.  Build it with -ffp-contract=off so that no
// multiply-add is fused (the reference's CPU build, x86-64 without -mfma,
// never fuses).
//
// Each exported function names the reference file:line it follows.  Paths are
// relative to the hippoLBM source tree.
//
// Pinning (how this restatement is checked, see tests/test_oracle_*.py):
//   * the 7 regression decks (regression-test/*.msp): sum(rho), min/max |u|;
//   * closed forms: weights/moments of the D3Q19 set, BGK fixed point and
//     conservation, MRT == BGK under the LBGK parametrisation
//     (src/collision/include/hippoLBM/collision/mrt.hpp:48-50), all-periodic
//     streaming == cyclic shift, Zou/He reproducing the prescribed velocity,
//     Poiseuille/Couette analytic profiles, wall-count rasterization pins.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int Q = 19;
// D3Q19 ordering of src/grid/include/hippoLBM/grid/scheme.hpp:52-70.
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int WDEN[Q] = {3, 18, 18, 18, 18, 18, 18, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
// scheme.hpp:32 -- w = 1.0 / WDen
inline double W(int i) { return 1.0 / WDEN[i]; }

constexpr int FLUID = -1;  // include/hippoLBM/core/enum.hpp:59
constexpr int WALL = -2;   // include/hippoLBM/core/enum.hpp:60
constexpr int GHOST = 2;   // src/grid/include/hippoLBM/grid/make_domain.hpp:32

// bounce-back save/restore direction sets, src/bcs/include/hippoLBM/bcs/bounce_back.hpp:65-140
// index = dim*2 + side (side 0 = left/"_0", 1 = right/"_l")
constexpr int PRE_FID[6][5] = {{2, 10, 8, 12, 14}, {1, 9, 7, 11, 13}, {4, 8, 9, 16, 18},
                               {3, 7, 10, 15, 17}, {6, 13, 12, 17, 16}, {5, 14, 11, 18, 15}};
constexpr int POST_FID[6][5] = {{1, 9, 7, 11, 13}, {2, 10, 8, 12, 14}, {3, 7, 10, 15, 17},
                                {4, 8, 9, 16, 18}, {5, 14, 11, 18, 15}, {6, 13, 12, 17, 16}};

struct Grid {
  int n[3];    // nodes per axis (cell_dims + 1), single rank
  int L[3];    // "All" box lengths, n + 2*GHOST
  int per[3];  // periodicity
  int ext_lo[3], ext_hi[3];  // inclusive extend box, make_domain.hpp:181-188
  double dx;
  size_t N;
  std::vector<double> f, m0, m1;
  std::vector<int> obst, level;
  std::vector<int> all_, real_;
  // plane lists in LBMGridRegion order: 0 xy_0, 1 xy_l, 2 xz_0, 3 xz_l, 4 yz_0, 5 yz_l
  std::vector<int> plane[6];
  // bounce_back_manager<19>::_data_ : 0 x left, 1 x right, 2 y left, 3 y right, 4 z bottom, 5 z top
  std::vector<double> bb[6];

  inline int idx(int x, int y, int z) const { return (z * L[1] + y) * L[0] + x; }
  inline double& F(size_t i, int q) { return f[N * (size_t)q + i]; }
  inline bool defined(int x, int y, int z) const {  // grid.hpp:198-203
    return !(x < ext_lo[0] || x > ext_hi[0] || y < ext_lo[1] || y > ext_hi[1] || z < ext_lo[2] ||
             z > ext_hi[2]);
  }
};

// bounce_back.hpp:33-56: plane traversal for (dim, side)
inline int plane_of(int dim, int side) {
  if (dim == 0) return side ? 5 : 4;
  if (dim == 1) return side ? 3 : 2;
  return side ? 1 : 0;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// domain + build_traversal + define_grid_3dq19
//   src/grid/domain_lbm.cpp:66-110, make_domain.hpp:157-264 (single rank),
//   grid_region.hpp:89-197, define_lbm_fields.cpp:58-86
// ---------------------------------------------------------------------------
void* orc_create(int nx, int ny, int nz, int px, int py, int pz, double dx) {
  Grid* g = new Grid();
  g->n[0] = nx;
  g->n[1] = ny;
  g->n[2] = nz;
  g->per[0] = px;
  g->per[1] = py;
  g->per[2] = pz;
  g->dx = dx;
  for (int d = 0; d < 3; d++) {
    g->L[d] = g->n[d] + 2 * GHOST;
    g->ext_lo[d] = 0;
    g->ext_hi[d] = g->L[d] - 1;
    if (!g->per[d]) {  // single rank: coord 0 == ndims-1
      g->ext_lo[d] += GHOST;
      g->ext_hi[d] -= GHOST;
    }
  }
  g->N = (size_t)g->L[0] * g->L[1] * g->L[2];
  g->f.assign(g->N * Q, 0.0);
  g->m0.assign(g->N, 0.0);
  g->m1.assign(g->N * 3, 0.0);
  g->obst.assign(g->N, 0);
  g->level.assign(g->N, 3);
  // build_traversal, grid_region.hpp:107-132
  for (int z = 0; z < g->L[2]; z++)
    for (int y = 0; y < g->L[1]; y++)
      for (int x = 0; x < g->L[0]; x++) {
        int id = g->idx(x, y, z);
        g->all_.push_back(id);
        int p[3] = {x, y, z};
        bool in_ext = true, in_real = true, in_inside = true;
        for (int d = 0; d < 3; d++) {
          if (p[d] < g->ext_lo[d] || p[d] > g->ext_hi[d]) in_ext = false;
          if (p[d] < GHOST || p[d] > g->L[d] - 1 - GHOST) in_real = false;
          if (p[d] < GHOST + 1 || p[d] > g->L[d] - 2 - GHOST) in_inside = false;
        }
        g->level[id] = 3;
        if (in_ext) {
          g->level[id] = 2;
          if (in_real) {
            g->real_.push_back(id);
            g->level[id] = 1;
            if (in_inside) g->level[id] = 0;
          }
        }
      }
  // plane lists, grid_region.hpp:148-193 (single rank: all six exist)
  const int r0x = GHOST, r0y = GHOST, r0z = GHOST;
  const int rlx = g->L[0] - 1 - GHOST, rly = g->L[1] - 1 - GHOST, rlz = g->L[2] - 1 - GHOST;
  for (int y = 0; y < g->L[1]; y++)
    for (int x = 0; x < g->L[0]; x++) {
      g->plane[0].push_back(g->idx(x, y, r0z));
      g->plane[1].push_back(g->idx(x, y, rlz));
    }
  for (int z = 0; z < g->L[2]; z++)
    for (int x = 0; x < g->L[0]; x++) {
      g->plane[2].push_back(g->idx(x, r0y, z));
      g->plane[3].push_back(g->idx(x, rly, z));
    }
  for (int z = 0; z < g->L[2]; z++)
    for (int y = 0; y < g->L[1]; y++) {
      g->plane[4].push_back(g->idx(r0x, y, z));
      g->plane[5].push_back(g->idx(rlx, y, z));
    }
  return g;
}

void orc_destroy(void* h) { delete (Grid*)h; }

void orc_dims(void* h, int* out) {  // [Lx, Ly, Lz, nx, ny, nz]
  Grid* g = (Grid*)h;
  for (int d = 0; d < 3; d++) {
    out[d] = g->L[d];
    out[3 + d] = g->n[d];
  }
}
double* orc_f(void* h) { return ((Grid*)h)->f.data(); }
double* orc_m0(void* h) { return ((Grid*)h)->m0.data(); }
double* orc_m1(void* h) { return ((Grid*)h)->m1.data(); }
int* orc_obst(void* h) { return ((Grid*)h)->obst.data(); }
int* orc_level(void* h) { return ((Grid*)h)->level.data(); }
int orc_plane_size(void* h, int p) { return (int)((Grid*)h)->plane[p].size(); }
const int* orc_plane(void* h, int p) { return ((Grid*)h)->plane[p].data(); }

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// ---------------------------------------------------------------------------
// lbm_parameters, src/grid/lbm_parameters.cpp:73-99 and
// src/grid/include/hippoLBM/grid/lbm_parameters.hpp:61-82,132-136
// out = [tau, nu, dtLB, celerity, Fx_lbm, Fy_lbm, Fz_lbm]
// ---------------------------------------------------------------------------
void orc_lbm_parameters(double dx, double nuth, double celerity, int has_tau, double tau, double dt,
                        const double* Fext, double* out) {
  double tau_, nu_, dtLB_, c_;
  if (dt > 0.0) {  // define_by_dt
    dtLB_ = dt;
    tau_ = 3 * nuth * dtLB_ / (dx * dx) + 0.5;
    nu_ = (tau_ - 0.5) / 3.0;
    c_ = dx / dtLB_;
  } else if (has_tau) {  // define_by_tau
    tau_ = tau;
    dtLB_ = dx * dx * (tau_ - 0.5) / (3.0 * nuth);
    nu_ = nuth * dtLB_ / (dx * dx);
    c_ = dx / dtLB_;
  } else {  // define_by_c
    c_ = celerity;
    dtLB_ = dx / c_;
    nu_ = nuth * dtLB_ / (dx * dx);
    tau_ = 3. * nu_ + 0.5;
  }
  // convert_force<LBM_UNITS>: value * dtLB * dtLB / dx with dx = celerity*dtLB
  const double dxc = c_ * dtLB_;
  out[0] = tau_;
  out[1] = nu_;
  out[2] = dtLB_;
  out[3] = c_;
  for (int d = 0; d < 3; d++) out[4 + d] = Fext[d] * dtLB_ * dtLB_ / dxc;
}

// ---------------------------------------------------------------------------
// init_obst, src/grid/init_obst.cu:54-60 (All points := FLUIDE_)
// ---------------------------------------------------------------------------
void orc_init_obst(void* h) {
  Grid* g = (Grid*)h;
  std::fill(g->obst.begin(), g->obst.end(), FLUID);
}

// ---------------------------------------------------------------------------
// set_distribution, src/grid/set_distribution.cu:69-109 +
// set_distribution.hpp:38-41 (f_i = value * w_i)
// ---------------------------------------------------------------------------
static void restrict_to_extend(Grid* g, const int gmin[3], const int gmax[3], int lo[3], int hi[3], bool* ok) {
  // grid.hpp:406-428 with InputBoxArea = Local: p_local = p - offset = p + GHOST
  *ok = true;
  for (int d = 0; d < 3; d++) {
    int a = gmin[d] + GHOST, b = gmax[d] + GHOST;
    if (g->ext_hi[d] < a || b < g->ext_lo[d]) *ok = false;  // box3d.hpp:201-210
    lo[d] = std::max(a, g->ext_lo[d]);
    hi[d] = std::min(b, g->ext_hi[d]);
  }
}

void orc_set_distribution(void* h, double value, int has_bounds, const double* bmin, const double* bmax,
                          int do_update);

void orc_update_ghost(void* h);

void orc_set_distribution(void* h, double value, int has_bounds, const double* bmin, const double* bmax,
                          int do_update) {
  Grid* g = (Grid*)h;
  if (has_bounds) {
    const double Dx = g->dx;
    int gmin[3] = {int(bmin[0] / Dx), int(bmin[1] / Dx), int(bmin[2] / Dx)};
    int gmax[3] = {int(bmax[0] / Dx), int(bmax[1] / Dx), int(bmax[2] / Dx)};
    int lo[3], hi[3];
    bool ok;
    restrict_to_extend(g, gmin, gmax, lo, hi, &ok);
    if (!ok) return;
    for (int z = lo[2]; z <= hi[2]; z++)
      for (int y = lo[1]; y <= hi[1]; y++)
        for (int x = lo[0]; x <= hi[0]; x++) {
          size_t id = g->idx(x, y, z);
          for (int i = 0; i < Q; i++) g->F(id, i) = value * W(i);
        }
  } else if (do_update) {
    for (int id : g->real_)
      for (int i = 0; i < Q; i++) g->F(id, i) = value * W(i);
    orc_update_ghost(h);
  } else {
    for (int id : g->all_)
      for (int i = 0; i < Q; i++) g->F(id, i) = value * W(i);
  }
}

// Direct per-node population write (seeded initial states for parity tests).
void orc_set_node_populations(void* h, const double* fnode) {
  // fnode: SoA over the Real nodes, fnode[i*nreal + (z*ny + y)*nx + x]
  Grid* g = (Grid*)h;
  const size_t nr = (size_t)g->n[0] * g->n[1] * g->n[2];
  for (int z = 0; z < g->n[2]; z++)
    for (int y = 0; y < g->n[1]; y++)
      for (int x = 0; x < g->n[0]; x++) {
        size_t id = g->idx(x + GHOST, y + GHOST, z + GHOST);
        size_t k = ((size_t)z * g->n[1] + y) * g->n[0] + x;
        for (int i = 0; i < Q; i++) g->F(id, i) = fnode[i * nr + k];
      }
}

void orc_get_node_populations(void* h, double* fnode) {
  Grid* g = (Grid*)h;
  const size_t nr = (size_t)g->n[0] * g->n[1] * g->n[2];
  for (int z = 0; z < g->n[2]; z++)
    for (int y = 0; y < g->n[1]; y++)
      for (int x = 0; x < g->n[0]; x++) {
        size_t id = g->idx(x + GHOST, y + GHOST, z + GHOST);
        size_t k = ((size_t)z * g->n[1] + y) * g->n[0] + x;
        for (int i = 0; i < Q; i++) fnode[i * nr + k] = g->F(id, i);
      }
}

// ---------------------------------------------------------------------------
// update_obstacles, src/obstacle/update_obstacles.cu:32-71 with
// Ball::solid / Wall::solid, src/obstacle/include/hippoLBM/obstacle/obstacle.hpp:13-17,46-64,86-97
// kind 0 = ball (p0 = center, r), kind 1 = wall (p0 = bmin, p1 = bmax)
// ---------------------------------------------------------------------------
void orc_update_obstacle(void* h, int kind, const double* p0, const double* p1, double r) {
  Grid* g = (Grid*)h;
  double bmin[3], bmax[3];
  if (kind == 0) {
    for (int d = 0; d < 3; d++) {
      bmin[d] = p0[d] - r;
      bmax[d] = p0[d] + r;
    }
  } else {
    for (int d = 0; d < 3; d++) {
      bmin[d] = p0[d];
      bmax[d] = p1[d];
    }
  }
  const double dx = g->dx;
  int gmin[3] = {int(bmin[0] / dx), int(bmin[1] / dx), int(bmin[2] / dx)};
  int gmax[3] = {int(bmax[0] / dx), int(bmax[1] / dx), int(bmax[2] / dx)};
  int lo[3], hi[3];
  bool ok;
  restrict_to_extend(g, gmin, gmax, lo, hi, &ok);
  if (!ok) return;
  const double r2 = r * r;
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++) {
        // compute_position<Global>, grid.hpp:355-361: (local + offset) * dx, offset = -GHOST
        double px = (double(i) + double(-GHOST)) * dx;
        double py = (double(j) + double(-GHOST)) * dx;
        double pz = (double(k) + double(-GHOST)) * dx;
        bool solid;
        if (kind == 0) {
          double rx = px - p0[0], ry = py - p0[1], rz = pz - p0[2];
          solid = (rx * rx + ry * ry + rz * rz) <= r2;
        } else {
          solid = p0[0] < px && px < p1[0] && p0[1] < py && py < p1[1] && p0[2] < pz && pz < p1[2];
        }
        if (solid) g->obst[g->idx(i, j, k)] = WALL;
      }
}

// init_double_couette, src/prepro/init_double_couette.cu:80-120 and
// src/prepro/include/hippoLBM/prepro/double_couette.hpp:26-45: on the Real
// box, u = U - dU * (global coordinate along dim), dU = U / (0.5 * (n - 1)),
// f_i = 1. * w_i * (1. + 3. eu + 4.5 eu eu - 1.5 u.u).  U is in LBM units.
void orc_init_double_couette(void* h, int dim, const double* U) {
  Grid* g = (Grid*)h;
  const int n = g->n[dim];
  double dU[3];
  for (int d = 0; d < 3; d++) dU[d] = U[d] / (0.5 * (n - 1));
  for (int k = GHOST; k < GHOST + g->n[2]; k++)
    for (int j = GHOST; j < GHOST + g->n[1]; j++)
      for (int i = GHOST; i < GHOST + g->n[0]; i++) {
        const int c[3] = {i, j, k};
        const double value = c[dim] - GHOST;
        const double u[3] = {U[0] - dU[0] * value, U[1] - dU[1] * value, U[2] - dU[2] * value};
        const double u_squ = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const size_t idx = g->idx(i, j, k);
        for (int q = 0; q < Q; q++) {
          const double eu = u[0] * EX[q] + u[1] * EY[q] + u[2] * EZ[q];
          g->f[(size_t)q * g->N + idx] = 1. * W(q) * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * u_squ);
        }
      }
}

// Quadric obstacle (register_quadrics.cpp:72-84, obstacle.hpp:107-138): the
// candidate box is Quadric::covered() = [-1e6, 1e6]^3 truncated by int(b/dx)
// (clamped to the int range) and restricted to the Extend box; solid where
// [p,1] Q [p,1]^T <= 0.  onika::math::quadric_eval is not in the reference
// tree: the form is evaluated as p . (Q [p,1]) with the rows of Q first, left
// to right -- parity unpinned at the rounding level of that order.
void orc_update_quadric(void* h, const double* q) {
  Grid* g = (Grid*)h;
  const double dx = g->dx;
  auto trunc = [](double v) {
    if (v > 2147483647.0) return 2147483647;
    if (v < -2147483647.0) return -2147483647;
    return int(v);
  };
  int gmin[3], gmax[3];
  for (int d = 0; d < 3; d++) {
    gmin[d] = trunc(-1e6 / dx);
    gmax[d] = trunc(1e6 / dx);
  }
  int lo[3], hi[3];
  bool ok;
  restrict_to_extend(g, gmin, gmax, lo, hi, &ok);
  if (!ok) return;
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++) {
        const double x[3] = {(double(i) + double(-GHOST)) * dx, (double(j) + double(-GHOST)) * dx,
                             (double(k) + double(-GHOST)) * dx};
        double r[4];
        for (int a = 0; a < 4; a++) r[a] = q[4 * a] * x[0] + q[4 * a + 1] * x[1] + q[4 * a + 2] * x[2] + q[4 * a + 3];
        if (x[0] * r[0] + x[1] * r[1] + x[2] * r[2] + r[3] <= 0.0) g->obst[g->idx(i, j, k)] = WALL;
      }
}

// ---------------------------------------------------------------------------
// macro_variables, src/collision/include/hippoLBM/collision/macro_variables.hpp:48-79
// (All points)
// ---------------------------------------------------------------------------
void orc_macro_variables(void* h) {
  Grid* g = (Grid*)h;
  const long N = (long)g->N;
#pragma omp parallel for schedule(static)
  for (long id = 0; id < N; id++) {
    if (g->obst[id] >= FLUID) {
      double rho = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
      for (int i = 0; i < Q; i++) {
        const double s = g->f[(size_t)N * i + id];
        ux += s * EX[i];
        uy += s * EY[i];
        uz += s * EZ[i];
        rho += s;
      }
      if (rho > 1.0e-14) {
        ux /= rho;
        uy /= rho;
        uz /= rho;
      }
      g->m0[id] = rho;
      g->m1[id] = ux;
      g->m1[N + id] = uy;
      g->m1[2 * N + id] = uz;
    } else {
      g->m1[id] = 0;
      g->m1[N + id] = 0;
      g->m1[2 * N + id] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// bgk<19, Real>, src/collision/include/hippoLBM/collision/bgk.hpp:47-62,
// operator src/collision/bgk.cu:77-94 (all points, level mask)
// ---------------------------------------------------------------------------
void orc_bgk(void* h, double tau, double Fx, double Fy, double Fz) {
  Grid* g = (Grid*)h;
  const long N = (long)g->N;
#pragma omp parallel for schedule(static)
  for (long id = 0; id < N; id++) {
    const bool update = (g->level[id] <= 1) && (g->obst[id] == FLUID);
    const double rho = g->m0[id];
    const double ux = g->m1[id];
    const double uy = g->m1[N + id];
    const double uz = g->m1[2 * N + id];
    const double u_squ = (ux * ux + uy * uy + uz * uz);
    for (int i = 0; i < Q; i++) {
      double& fi = g->f[(size_t)N * i + id];
      const double w = W(i);
      double ef = EX[i] * Fx + EY[i] * Fy + EZ[i] * Fz;
      double eu = EX[i] * ux + EY[i] * uy + EZ[i] * uz;
      double feq = w * rho * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * u_squ);
      fi += update * ((feq - fi) / tau + 3. * rho * w * ef);
    }
  }
}

// ---------------------------------------------------------------------------
// mrt<19>, src/collision/include/hippoLBM/collision/mrt.hpp:42-193,
// operator src/collision/mrt.cu:71-89 (Real list, fluid only).
// rates == NULL selects the reference's hard-coded d'Humieres rates
// (mrt.hpp:43-47); otherwise rates = [s1, s2, s4, s9, s10, s13, s16, weps, wepsj, wxx].
// ---------------------------------------------------------------------------
void orc_mrt(void* h, double tau, double Fx, double Fy, double Fz, const double* rates) {
  Grid* g = (Grid*)h;
  const long N = (long)g->N;
  const long nreal = (long)g->real_.size();
#pragma omp parallel for schedule(static)
  for (long r = 0; r < nreal; r++) {
    const long idx = g->real_[r];
    if (g->obst[idx] != FLUID) continue;
    const double rho_m0 = g->m0[idx];
    double* fp = g->f.data();
#define Fq(q) fp[(size_t)N * (q) + idx]
    double s9 = 1 / tau, s13 = s9;
    double s1 = 1.19, s2 = 1.4, s4 = 1.2, s10 = s2, s16 = 1.98;
    double weps = 0, wepsj = -475. / 63., wxx = 0;
    if (rates) {
      s1 = rates[0];
      s2 = rates[1];
      s4 = rates[2];
      s9 = rates[3];
      s10 = rates[4];
      s13 = rates[5];
      s16 = rates[6];
      weps = rates[7];
      wepsj = rates[8];
      wxx = rates[9];
    }
    double rho0, jj, jx2, jy2, jz2;
    double rho, e, eps, jx, qx, jy, qy, jz, qz, pxx3, pixx3, pww, piww, pxy, pyz, pxz, mx, my, mz;
    double eO, epsO, qxO, qyO, qzO, pxx3O, pixx3O, pwwO, piwwO, pxyO, pyzO, pxzO, mxO, myO, mzO;
    double e_eq, eps_eq, pxx3_eq, pww_eq, pxy_eq, pyz_eq, pxz_eq;
    rho = +Fq(0) + Fq(1) + Fq(2) + Fq(3) + Fq(4) + Fq(5) + Fq(6) + Fq(7) + Fq(8) + Fq(9) + Fq(10) + Fq(11) +
          Fq(12) + Fq(13) + Fq(14) + Fq(15) + Fq(16) + Fq(17) + Fq(18);
    e = -30 * Fq(0) - 11 * Fq(1) - 11 * Fq(2) - 11 * Fq(3) - 11 * Fq(4) - 11 * Fq(5) - 11 * Fq(6) + 8 * Fq(7) +
        8 * Fq(8) + 8 * Fq(9) + 8 * Fq(10) + 8 * Fq(11) + 8 * Fq(12) + 8 * Fq(13) + 8 * Fq(14) + 8 * Fq(15) +
        8 * Fq(16) + 8 * Fq(17) + 8 * Fq(18);
    eps = +12 * Fq(0) - 4 * Fq(1) - 4 * Fq(2) - 4 * Fq(3) - 4 * Fq(4) - 4 * Fq(5) - 4 * Fq(6) + Fq(7) + Fq(8) +
          Fq(9) + Fq(10) + Fq(11) + Fq(12) + Fq(13) + Fq(14) + Fq(15) + Fq(16) + Fq(17) + Fq(18);
    jx = +Fq(1) - Fq(2) + Fq(7) - Fq(8) + Fq(9) - Fq(10) + Fq(11) - Fq(12) + Fq(13) - Fq(14);
    qx = -4 * Fq(1) + 4 * Fq(2) + Fq(7) - Fq(8) + Fq(9) - Fq(10) + Fq(11) - Fq(12) + Fq(13) - Fq(14);
    jy = +Fq(3) - Fq(4) + Fq(7) - Fq(8) - Fq(9) + Fq(10) + Fq(15) - Fq(16) + Fq(17) - Fq(18);
    qy = -4 * Fq(3) + 4 * Fq(4) + Fq(7) - Fq(8) - Fq(9) + Fq(10) + Fq(15) - Fq(16) + Fq(17) - Fq(18);
    jz = +Fq(5) - Fq(6) + Fq(11) - Fq(12) - Fq(13) + Fq(14) + Fq(15) - Fq(16) - Fq(17) + Fq(18);
    qz = -4 * Fq(5) + 4 * Fq(6) + Fq(11) - Fq(12) - Fq(13) + Fq(14) + Fq(15) - Fq(16) - Fq(17) + Fq(18);
    pxx3 = +2 * Fq(1) + 2 * Fq(2) - Fq(3) - Fq(4) - Fq(5) - Fq(6) + Fq(7) + Fq(8) + Fq(9) + Fq(10) + Fq(11) +
           Fq(12) + Fq(13) + Fq(14) - 2 * Fq(15) - 2 * Fq(16) - 2 * Fq(17) - 2 * Fq(18);
    pixx3 = -4 * Fq(1) - 4 * Fq(2) + 2 * Fq(3) + 2 * Fq(4) + 2 * Fq(5) + 2 * Fq(6) + Fq(7) + Fq(8) + Fq(9) +
            Fq(10) + Fq(11) + Fq(12) + Fq(13) + Fq(14) - 2 * Fq(15) - 2 * Fq(16) - 2 * Fq(17) - 2 * Fq(18);
    pww = +Fq(3) + Fq(4) - Fq(5) - Fq(6) + Fq(7) + Fq(8) + Fq(9) + Fq(10) - Fq(11) - Fq(12) - Fq(13) - Fq(14);
    piww = -2 * Fq(3) - 2 * Fq(4) + 2 * Fq(5) + 2 * Fq(6) + Fq(7) + Fq(8) + Fq(9) + Fq(10) - Fq(11) - Fq(12) -
           Fq(13) - Fq(14);
    pxy = +Fq(7) + Fq(8) - Fq(9) - Fq(10);
    pyz = +Fq(15) + Fq(16) - Fq(17) - Fq(18);
    pxz = +Fq(11) + Fq(12) - Fq(13) - Fq(14);
    mx = +Fq(7) - Fq(8) + Fq(9) - Fq(10) - Fq(11) + Fq(12) - Fq(13) + Fq(14);
    my = -Fq(7) + Fq(8) + Fq(9) - Fq(10) + Fq(15) - Fq(16) + Fq(17) - Fq(18);
    mz = +Fq(11) - Fq(12) - Fq(13) + Fq(14) - Fq(15) + Fq(16) + Fq(17) - Fq(18);

    rho0 = 1.;
    jx2 = jx * jx;
    jy2 = jy * jy;
    jz2 = jz * jz;
    jj = (jx2 + jy2 + jz2);

    e_eq = -11 * rho + 19 / rho0 * jj;
    eO = e - s1 * (e - e_eq);
    eps_eq = weps * rho + wepsj / rho0 * jj;
    epsO = eps - s2 * (eps - eps_eq);

    qxO = qx - s4 * (qx + 2. / 3. * jx);
    qyO = qy - s4 * (qy + 2. / 3. * jy);
    qzO = qz - s4 * (qz + 2. / 3. * jz);
    pxx3_eq = 1 / (rho0) * (2 * jx2 - (jy2 + jz2));
    pxx3O = pxx3 - s9 * (pxx3 - pxx3_eq);
    pixx3O = pixx3 - s10 * (pixx3 - wxx * pxx3_eq);

    pww_eq = 1 / rho0 * (jy2 - jz2);
    pwwO = pww - s9 * (pww - pww_eq);
    piwwO = piww - s10 * (piww - wxx * pww_eq);
    pxy_eq = 1 / rho0 * jx * jy;
    pxyO = pxy - s13 * (pxy - pxy_eq);
    pyz_eq = 1 / rho0 * jy * jz;
    pyzO = pyz - s13 * (pyz - pyz_eq);
    pxz_eq = 1 / rho0 * jx * jz;
    pxzO = pxz - s13 * (pxz - pxz_eq);
    mxO = mx - s16 * mx;
    myO = my - s16 * my;
    mzO = mz - s16 * mz;

    Fq(0) = +1. / 19 * rho - 5. / 399 * eO + 1. / 21 * epsO;
    Fq(1) = +1. / 19 * rho - 11. / 2394 * eO - 1. / 63 * epsO + 1. / 10 * jx - 1. / 10 * qxO + 1. / 18 * pxx3O -
            1. / 18 * pixx3O;
    Fq(2) = +1. / 19 * rho - 11. / 2394 * eO - 1. / 63 * epsO - 1. / 10 * jx + 1. / 10 * qxO + 1. / 18 * pxx3O -
            1. / 18 * pixx3O;
    Fq(3) = +1. / 19 * rho - 11. / 2394 * eO - 1. / 63 * epsO + 1. / 10 * jy - 1. / 10 * qyO - 1. / 36 * pxx3O +
            1. / 36 * pixx3O + 1. / 12 * pwwO - 1. / 12 * piwwO;
    Fq(4) = +1. / 19 * rho - 11. / 2394 * eO - 1. / 63 * epsO - 1. / 10 * jy + 1. / 10 * qyO - 1. / 36 * pxx3O +
            1. / 36 * pixx3O + 1. / 12 * pwwO - 1. / 12 * piwwO;
    Fq(5) = +1. / 19 * rho - 11. / 2394 * eO - 1. / 63 * epsO + 1. / 10 * jz - 1. / 10 * qzO - 1. / 36 * pxx3O +
            1. / 36 * pixx3O - 1. / 12 * pwwO + 1. / 12 * piwwO;
    Fq(6) = +1. / 19 * rho - 11. / 2394 * eO - 1. / 63 * epsO - 1. / 10 * jz + 1. / 10 * qzO - 1. / 36 * pxx3O +
            1. / 36 * pixx3O - 1. / 12 * pwwO + 1. / 12 * piwwO;
    Fq(7) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO + 1. / 10 * jx + 1. / 40 * qxO + 1. / 10 * jy +
            1. / 40 * qyO + 1. / 36 * pxx3O + 1. / 72 * pixx3O + 1. / 12 * pwwO + 1. / 24 * piwwO + 1. / 4 * pxyO +
            1. / 8 * mxO - 1. / 8 * myO;
    Fq(8) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO - 1. / 10 * jx - 1. / 40 * qxO - 1. / 10 * jy -
            1. / 40 * qyO + 1. / 36 * pxx3O + 1. / 72 * pixx3O + 1. / 12 * pwwO + 1. / 24 * piwwO + 1. / 4 * pxyO -
            1. / 8 * mxO + 1. / 8 * myO;
    Fq(9) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO + 1. / 10 * jx + 1. / 40 * qxO - 1. / 10 * jy -
            1. / 40 * qyO + 1. / 36 * pxx3O + 1. / 72 * pixx3O + 1. / 12 * pwwO + 1. / 24 * piwwO - 1. / 4 * pxyO +
            1. / 8 * mxO + 1. / 8 * myO;
    Fq(10) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO - 1. / 10 * jx - 1. / 40 * qxO + 1. / 10 * jy +
             1. / 40 * qyO + 1. / 36 * pxx3O + 1. / 72 * pixx3O + 1. / 12 * pwwO + 1. / 24 * piwwO - 1. / 4 * pxyO -
             1. / 8 * mxO - 1. / 8 * myO;
    Fq(11) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO + 1. / 10 * jx + 1. / 40 * qxO + 1. / 10 * jz +
             1. / 40 * qzO + 1. / 36 * pxx3O + 1. / 72 * pixx3O - 1. / 12 * pwwO - 1. / 24 * piwwO + 1. / 4 * pxzO -
             1. / 8 * mxO + 1. / 8 * mzO;
    Fq(12) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO - 1. / 10 * jx - 1. / 40 * qxO - 1. / 10 * jz -
             1. / 40 * qzO + 1. / 36 * pxx3O + 1. / 72 * pixx3O - 1. / 12 * pwwO - 1. / 24 * piwwO + 1. / 4 * pxzO +
             1. / 8 * mxO - 1. / 8 * mzO;
    Fq(13) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO + 1. / 10 * jx + 1. / 40 * qxO - 1. / 10 * jz -
             1. / 40 * qzO + 1. / 36 * pxx3O + 1. / 72 * pixx3O - 1. / 12 * pwwO - 1. / 24 * piwwO - 1. / 4 * pxzO -
             1. / 8 * mxO - 1. / 8 * mzO;
    Fq(14) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO - 1. / 10 * jx - 1. / 40 * qxO + 1. / 10 * jz +
             1. / 40 * qzO + 1. / 36 * pxx3O + 1. / 72 * pixx3O - 1. / 12 * pwwO - 1. / 24 * piwwO - 1. / 4 * pxzO +
             1. / 8 * mxO + 1. / 8 * mzO;
    Fq(15) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO + 1. / 10 * jy + 1. / 40 * qyO + 1. / 10 * jz +
             1. / 40 * qzO - 1. / 18 * pxx3O - 1. / 36 * pixx3O + 1. / 4 * pyzO + 1. / 8 * myO - 1. / 8 * mzO;
    Fq(16) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO - 1. / 10 * jy - 1. / 40 * qyO - 1. / 10 * jz -
             1. / 40 * qzO - 1. / 18 * pxx3O - 1. / 36 * pixx3O + 1. / 4 * pyzO - 1. / 8 * myO + 1. / 8 * mzO;
    Fq(17) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO + 1. / 10 * jy + 1. / 40 * qyO - 1. / 10 * jz -
             1. / 40 * qzO - 1. / 18 * pxx3O - 1. / 36 * pixx3O - 1. / 4 * pyzO + 1. / 8 * myO + 1. / 8 * mzO;
    Fq(18) = +1. / 19 * rho + 4. / 1197 * eO + 1. / 252 * epsO - 1. / 10 * jy - 1. / 40 * qyO + 1. / 10 * jz +
             1. / 40 * qzO - 1. / 18 * pxx3O - 1. / 36 * pixx3O - 1. / 4 * pyzO - 1. / 8 * myO - 1. / 8 * mzO;
    // mrt.hpp:187-191, force term with rho from m0
    for (int i = 0; i < Q; i++) {
      const double ef = EX[i] * Fx + EY[i] * Fy + EZ[i] * Fz;
      Fq(i) += 3. * rho_m0 * W(i) * ef;
    }
#undef Fq
  }
}

// ---------------------------------------------------------------------------
// streaming, src/collision/streaming.cu:107-111 and
// src/collision/include/hippoLBM/collision/streaming.hpp:35-67
// ---------------------------------------------------------------------------
void orc_update_ghost(void* h) {
  // single rank: for every periodic shift (make_domain.hpp:213-255), copy the
  // opposite two-layer Real slab (operator_ghost_manager.hpp:96-114) into the
  // ghost slab (operator_ghost_manager.hpp:65-89). All packs read Real nodes
  // only and all unpacks write ghost nodes only, so copying directly is
  // equivalent to pack/memcpy/unpack (ghost_manager.hpp:100-152).
  Grid* g = (Grid*)h;
  for (int s = 0; s < 27; s++) {
    int rel[3], t = s;
    for (int d = 0; d < 3; d++) {
      int v = t % 3;
      rel[d] = v - 1;
      t = (t - v) / 3;
    }
    if (rel[0] == 0 && rel[1] == 0 && rel[2] == 0) continue;
    bool exists = true;
    for (int d = 0; d < 3; d++)
      if (rel[d] != 0 && !g->per[d]) exists = false;  // build_comm, operator_ghost_manager.hpp:36-57
    if (!exists) continue;
    int rlo[3], rhi[3], shift[3];
    for (int d = 0; d < 3; d++) {
      const int lower = GHOST, upper = g->L[d] - 1 - GHOST;
      if (rel[d] == -1) {
        rlo[d] = lower - GHOST;
        rhi[d] = lower - GHOST + 1;
        shift[d] = g->n[d];
      } else if (rel[d] == 1) {
        rlo[d] = upper + GHOST - 1;
        rhi[d] = upper + GHOST;
        shift[d] = -g->n[d];
      } else {
        rlo[d] = lower;
        rhi[d] = upper;
        shift[d] = 0;
      }
    }
    for (int z = rlo[2]; z <= rhi[2]; z++)
      for (int y = rlo[1]; y <= rhi[1]; y++)
        for (int x = rlo[0]; x <= rhi[0]; x++) {
          size_t dst = g->idx(x, y, z);
          size_t src = g->idx(x + shift[0], y + shift[1], z + shift[2]);
          for (int i = 0; i < Q; i++) g->F(dst, i) = g->F(src, i);
        }
  }
}

void orc_streaming(void* h) {
  Grid* g = (Grid*)h;
  const long N = (long)g->N;
  double* f = g->f.data();
  // step1 over Real points
#pragma omp parallel for schedule(static)
  for (long id = 0; id < N; id++) {
    if (g->level[id] <= 1)
      for (int i = 1; i < Q; i += 2) std::swap(f[(size_t)N * i + id], f[(size_t)N * (i + 1) + id]);
  }
  orc_update_ghost(h);
  // step2 over Extend points
  const int L0 = g->L[0], L1 = g->L[1];
#pragma omp parallel for schedule(static)
  for (long id = 0; id < N; id++) {
    if (g->level[id] > 2) continue;
    const int z = (int)(id / ((long)L0 * L1));
    const int rem = (int)(id - (long)z * L0 * L1);
    const int y = rem / L0;
    const int x = rem % L0;
    for (int i = 1; i < Q; i += 2) {
      const int nx = x + EX[i], ny = y + EY[i], nz = z + EZ[i];
      if (g->defined(nx, ny, nz)) {
        const size_t nid = g->idx(nx, ny, nz);
        std::swap(f[(size_t)N * (i + 1) + id], f[(size_t)N * i + nid]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// pre_bounce_back / post_bounce_back,
// src/bcs/pre_bounce_back.cu:89-118, src/bcs/post_bounce_back.cu:101-123,
// src/bcs/include/hippoLBM/bcs/bounce_back.hpp:92-156,
// src/bcs/include/hippoLBM/bcs/bounce_back_manager.hpp:83-117
// ---------------------------------------------------------------------------
void orc_pre_bounce_back(void* h) {
  Grid* g = (Grid*)h;
  for (int dim = 0; dim < 3; dim++) {
    if (g->per[dim]) continue;
    for (int side = 0; side < 2; side++) {
      const int b = dim * 2 + side;
      const auto& pl = g->plane[plane_of(dim, side)];
      const size_t P = pl.size();
      if (g->bb[b].size() != P * 5) g->bb[b].assign(P * 5, 0.0);
    }
  }
  for (int b = 0; b < 6; b++) {
    if (g->bb[b].empty()) continue;
    const auto& pl = g->plane[plane_of(b / 2, b % 2)];
    const size_t P = pl.size();
    for (size_t k = 0; k < P; k++)
      for (int i = 0; i < 5; i++) g->bb[b][i * P + k] = g->F(pl[k], PRE_FID[b][i]);
  }
}

void orc_post_bounce_back(void* h) {
  Grid* g = (Grid*)h;
  for (int b = 0; b < 6; b++) {
    if (g->bb[b].empty()) continue;
    const auto& pl = g->plane[plane_of(b / 2, b % 2)];
    const size_t P = pl.size();
    for (size_t k = 0; k < P; k++)
      for (int i = 0; i < 5; i++) g->F(pl[k], POST_FID[b][i]) = g->bb[b][i * P + k];
  }
}

// ---------------------------------------------------------------------------
// cavity_z_0 / cavity_z_l, src/bcs/cavity.cu:66-88 and
// src/bcs/include/hippoLBM/bcs/cavity.hpp:57-87.
// side 0 = cavity_z_0 (dirs 5,14,11,18,15), 1 = cavity_z_l (dirs 6,13,12,17,16).
// (ux,uy,uz) are used as given (no unit conversion), L = global node count in z.
// The obstacle test uses the plane-local index k as a cell index (cavity.hpp:81-82).
// ---------------------------------------------------------------------------
void orc_cavity(void* h, int side, double ux, double uy, double uz) {
  Grid* g = (Grid*)h;
  static const int fid[2][5] = {{5, 14, 11, 18, 15}, {6, 13, 12, 17, 16}};
  const int b = 4 + side;
  const double L = g->n[2];
  const double uxx = ux * (1 + 0.5 / (L - 1));
  const double uyy = uy * (1 + 0.5 / (L - 1));
  const double uzz = uz * (1 + 0.5 / (L - 1));
  double coeff[5];
  for (int i = 0; i < 5; i++) {
    const int q = fid[side][i];
    coeff[i] = 6. * W(q) * (EX[q] * uxx + EY[q] * uyy + EZ[q] * uzz);
  }
  if (g->bb[b].empty()) return;
  const size_t P = g->bb[b].size() / 5;
  for (size_t k = 0; k < P; k++)
    if (g->obst[k] == FLUID)
      for (int i = 0; i < 5; i++) g->bb[b][i * P + k] += coeff[i];
}

void orc_cavity_coeff(int side, double L, double ux, double uy, double uz, double* out) {
  static const int fid[2][5] = {{5, 14, 11, 18, 15}, {6, 13, 12, 17, 16}};
  const double uxx = ux * (1 + 0.5 / (L - 1));
  const double uyy = uy * (1 + 0.5 / (L - 1));
  const double uzz = uz * (1 + 0.5 / (L - 1));
  for (int i = 0; i < 5; i++) {
    const int q = fid[side][i];
    out[i] = 6. * W(q) * (EX[q] * uxx + EY[q] * uyy + EZ[q] * uzz);
  }
}

// ---------------------------------------------------------------------------
// wall_bounce_back, src/bcs/wall_bounce_back.cu:58-73 and
// src/bcs/include/hippoLBM/bcs/bounce_back.hpp:163-185 (3-D over the Extend box)
// ---------------------------------------------------------------------------
void orc_wall_bounce_back(void* h) {
  Grid* g = (Grid*)h;
  for (int z = g->ext_lo[2]; z <= g->ext_hi[2]; z++)
    for (int y = g->ext_lo[1]; y <= g->ext_hi[1]; y++)
      for (int x = g->ext_lo[0]; x <= g->ext_hi[0]; x++) {
        const size_t id = g->idx(x, y, z);
        if (g->obst[id] != WALL) continue;
        for (int i = 1; i < Q; i++) {
          const int nx = x + EX[i], ny = y + EY[i], nz = z + EZ[i];
          if (g->defined(nx, ny, nz)) {
            const size_t nid = g->idx(nx, ny, nz);
            if (g->obst[nid] != WALL) g->F(id, i) = g->F(nid, OPP[i]);
          }
        }
      }
}

// ---------------------------------------------------------------------------
// lid_driven_cavity, src/bcs/lid_driven_cavity.cu:69-93 and
// src/bcs/include/hippoLBM/bcs/lid_driven_cavity.hpp:38-62.
// U in LBM units (already divided by celerity); plane in LBMGridRegion order.
// ---------------------------------------------------------------------------
void orc_lid_driven_cavity(void* h, int plane, double Ux, double Uy, double Uz) {
  Grid* g = (Grid*)h;
  const auto& pl = g->plane[plane];
  const double u_squ = Ux * Ux + Uy * Uy + Uz * Uz;
  for (size_t k = 0; k < pl.size(); k++) {
    const size_t id = pl[k];
    if (g->obst[id] != FLUID) continue;
    double rho = 0.0;
    for (int i = 0; i < Q; i++) rho += g->F(id, i);
    for (int i = 0; i < Q; i++) {
      double eu = EX[i] * Ux + EY[i] * Uy + EZ[i] * Uz;
      g->F(id, i) = W(i) * rho * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * u_squ);
    }
  }
}

// ---------------------------------------------------------------------------
// neumann (Zou/He velocity), src/bcs/neuman.cu:69-122 and
// src/bcs/include/hippoLBM/bcs/neumann.hpp:38-258. U in LBM units.
// ---------------------------------------------------------------------------
// The density (pressure) form has no counterpart in the reference: rho_in is
// prescribed, the tangential velocity is zero and the normal velocity follows
// from the known-population sum K of the velocity form, rho (1 -+ u_n) = K.
static double zou_he_density(double K, double& un, bool high, bool pressure, double rho_in) {
  if (pressure) {
    un = high ? K / rho_in - 1. : 1. - K / rho_in;
    return rho_in;
  }
  return high ? K / (1. + un) : K / (1. - un);
}

static void zou_he(void* h, int plane, double ux, double uy, double uz, bool pressure, double rho_in) {
  Grid* g = (Grid*)h;
  const auto& pl = g->plane[plane];
  if (pressure) ux = uy = uz = 0.0;
  for (size_t k = 0; k < pl.size(); k++) {
    const size_t idx = pl[k];
    if (g->obst[idx] != FLUID) continue;
#define f(q) g->F(idx, q)
    if (plane == 5) {  // plan_yz_l -> neumann_x_l
      const double rho = zou_he_density(f(3) + f(4) + f(5) + f(6) + f(15) + f(17) + f(18) + f(16) + f(0) +
                                            2. * (f(1) + f(7) + f(9) + f(11) + f(13)),
                                        ux, true, pressure, rho_in);
      const double nyx = (1. / 2.) * (f(3) + f(15) + f(17) - (f(4) + f(18) + f(16))) - (1. / 3.) * rho * uy;
      const double nzx = (1. / 2.) * (f(5) + f(18) + f(15) - (f(6) + f(17) + f(16))) - (1. / 3.) * rho * uz;
      f(2) = f(1) - (1. / 3.) * rho * ux;
      f(10) = f(9) + (1. / 6.) * rho * (-ux + uy) - nyx;
      f(8) = f(7) + (1. / 6.) * rho * (-ux - uy) + nyx;
      f(12) = f(11) + (1. / 6.) * rho * (-ux - uz) + nzx;
      f(14) = f(13) + (1. / 6.) * rho * (-ux + uz) - nzx;
    } else if (plane == 4) {  // plan_yz_0 -> neumann_x_0
      const double rho = zou_he_density(f(3) + f(4) + f(5) + f(6) + f(15) + f(17) + f(18) + f(16) + f(0) +
                                            2. * (f(2) + f(10) + f(8) + f(14) + f(12)),
                                        ux, false, pressure, rho_in);
      const double nyx = (1. / 2.) * (f(3) + f(15) + f(17) - (f(4) + f(18) + f(16))) - (1. / 3.) * rho * uy;
      const double nzx = (1. / 2.) * (f(5) + f(18) + f(15) - (f(6) + f(17) + f(16))) - (1. / 3.) * rho * uz;
      f(1) = f(2) + (1. / 3.) * rho * ux;
      f(9) = f(10) + (1. / 6.) * rho * (ux - uy) + nyx;
      f(7) = f(8) + (1. / 6.) * rho * (ux + uy) - nyx;
      f(11) = f(12) + (1. / 6.) * rho * (ux + uz) - nzx;
      f(13) = f(14) + (1. / 6.) * rho * (ux - uz) + nzx;
    } else if (plane == 3) {  // plan_xz_l -> neumann_y_l
      const double rho = zou_he_density(f(1) + f(2) + f(5) + f(6) + f(11) + f(13) + f(14) + f(12) + f(0) +
                                            2. * (f(3) + f(7) + f(10) + f(15) + f(17)),
                                        uy, true, pressure, rho_in);
      const double nxy = (1. / 2.) * (f(1) + f(11) + f(13) - (f(2) + f(14) + f(12))) - (1. / 3.) * rho * ux;
      const double nzy = (1. / 2.) * (f(5) + f(11) + f(14) - (f(6) + f(13) + f(12))) - (1. / 3.) * rho * uz;
      f(4) = f(3) - (1. / 3.) * rho * uy;
      f(8) = f(7) + (1. / 6.) * rho * (-uy - ux) + nxy;
      f(9) = f(10) + (1. / 6.) * rho * (-uy + ux) - nxy;
      f(16) = f(15) + (1. / 6.) * rho * (-uy - uz) + nzy;
      f(18) = f(17) + (1. / 6.) * rho * (-uy + uz) - nzy;
    } else if (plane == 2) {  // plan_xz_0 -> neumann_y_0
      const double rho = zou_he_density(f(1) + f(2) + f(5) + f(6) + f(11) + f(13) + f(14) + f(12) + f(0) +
                                            2. * (f(4) + f(9) + f(8) + f(18) + f(16)),
                                        uy, false, pressure, rho_in);
      const double nxy = (1. / 2.) * (f(1) + f(11) + f(13) - (f(2) + f(14) + f(12))) - (1. / 3.) * rho * ux;
      const double nzy = (1. / 2.) * (f(5) + f(11) + f(14) - (f(6) + f(13) + f(12))) - (1. / 3.) * rho * uz;
      f(3) = f(4) + (1. / 3.) * rho * uy;
      f(7) = f(8) + (1. / 6.) * rho * (uy + ux) - nxy;
      f(10) = f(9) + (1. / 6.) * rho * (uy - ux) + nxy;
      f(15) = f(16) + (1. / 6.) * rho * (uy + uz) - nzy;
      f(17) = f(18) + (1. / 6.) * rho * (uy - uz) + nzy;
    } else if (plane == 1) {  // plan_xy_l -> neumann_z_l
      const double rho = zou_he_density(f(0) + f(1) + f(2) + f(3) + f(4) + f(7) + f(9) + f(10) + f(8) +
                                            2. * (f(5) + f(11) + f(14) + f(15) + f(18)),
                                        uz, true, pressure, rho_in);
      const double nxz = (1. / 2.) * (f(1) + f(7) + f(9) - (f(2) + f(10) + f(8))) - (1. / 3.) * rho * ux;
      const double nyz = (1. / 2.) * (f(3) + f(7) + f(10) - (f(4) + f(9) + f(8))) - (1. / 3.) * rho * uy;
      f(6) = f(5) - (1. / 3.) * rho * uz;
      f(13) = f(14) + (1. / 6.) * rho * (-uz + ux) - nxz;
      f(12) = f(11) + (1. / 6.) * rho * (-uz - ux) + nxz;
      f(17) = f(18) + (1. / 6.) * rho * (-uz + uy) - nyz;
      f(16) = f(15) + (1. / 6.) * rho * (-uz - uy) + nyz;
    } else {  // plane 0, plan_xy_0 -> neumann_z_0
      const double rho = zou_he_density(f(0) + f(1) + f(2) + f(3) + f(4) + f(7) + f(9) + f(10) + f(8) +
                                            2. * (f(6) + f(13) + f(12) + f(17) + f(16)),
                                        uz, false, pressure, rho_in);
      const double nxz = (1. / 2.) * (f(1) + f(7) + f(9) - (f(2) + f(10) + f(8))) - (1. / 3.) * rho * ux;
      const double nyz = (1. / 2.) * (f(3) + f(7) + f(10) - (f(4) + f(9) + f(8))) - (1. / 3.) * rho * uy;
      f(5) = f(6) + (1. / 3.) * rho * uz;
      f(11) = f(12) + (1. / 6.) * rho * (uz + ux) - nxz;
      f(14) = f(13) + (1. / 6.) * rho * (uz - ux) + nxz;
      f(15) = f(16) + (1. / 6.) * rho * (uz + uy) - nyz;
      f(18) = f(17) + (1. / 6.) * rho * (uz - uy) + nyz;
    }
#undef f
  }
}

void orc_neumann(void* h, int plane, double ux, double uy, double uz) { zou_he(h, plane, ux, uy, uz, false, 1.0); }

void orc_pressure_inlet(void* h, int plane, double rho_in) { zou_he(h, plane, 0.0, 0.0, 0.0, true, rho_in); }

// outflow (no counterpart in the reference): at the fluid nodes of a face, the
// populations entering from outside copy the current (post-stream, post-BC)
// values of the node one layer inside.
void orc_outflow(void* h, int plane) {
  Grid* g = (Grid*)h;
  const auto& pl = g->plane[plane];
  const int dim = plane < 2 ? 2 : (plane < 4 ? 1 : 0);
  const int inward = (plane & 1) ? -1 : 1;
  const long step = dim == 0 ? 1 : (dim == 1 ? (long)g->L[0] : (long)g->L[0] * g->L[1]);
  for (size_t k = 0; k < pl.size(); k++) {
    const size_t idx = pl[k];
    if (g->obst[idx] != FLUID) continue;
    const size_t nid = (size_t)((long)idx + inward * step);
    for (int i = 1; i < Q; i++) {
      const int cd = dim == 0 ? EX[i] : (dim == 1 ? EY[i] : EZ[i]);
      if (cd == inward) g->F(idx, i) = g->F(nid, i);
    }
  }
}

// ---------------------------------------------------------------------------
// simulation_state, src/io/simulation_state.cu:73-98 and
// src/io/include/hippoLBM/io/simulation_state.hpp:47-55 (Real list).
// out = [sum(m0), min |m1|, max |m1|]
// ---------------------------------------------------------------------------
void orc_simulation_state(void* h, double* out) {
  Grid* g = (Grid*)h;
  const long N = (long)g->N;
  double s = 0.0, vmin = std::numeric_limits<double>::max(), vmax = std::numeric_limits<double>::lowest();
  for (int id : g->real_) {
    s += g->m0[id];
    const double x = g->m1[id], y = g->m1[N + id], z = g->m1[2 * N + id];
    const double v = std::sqrt(x * x + y * y + z * z);
    vmin = std::min(vmin, v);
    vmax = std::max(vmax, v);
  }
  out[0] = s;
  out[1] = vmin;
  out[2] = vmax;
}

}  // extern "C"
