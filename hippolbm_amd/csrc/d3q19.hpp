// d3q19.hpp -- D3Q19 lattice constants and the per-node physics shared by the
// HIP kernels (device) and the host library.
//
// Velocity ordering, weights and opposites follow hippoLBM's LBMScheme<19>
// (src/grid/include/hippoLBM/grid/scheme.hpp:52-70): 0 rest; 1/2 +-x; 3/4 +-y;
// 5/6 +-z; 7..18 the twelve diagonals as opposite pairs (odd i, i+1).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define HLBM_HD __host__ __device__
#else
#define HLBM_HD
#endif

namespace hlbm {

constexpr int Q = 19;

struct Lattice {
  static constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  static constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  // weight denominators: w_i = 1.0 / wden[i] (scheme.hpp:32)
  static constexpr int wden[Q] = {3, 18, 18, 18, 18, 18, 18, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36};
  static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
};

HLBM_HD constexpr int CX(int i) { return Lattice::cx[i]; }
HLBM_HD constexpr int CY(int i) { return Lattice::cy[i]; }
HLBM_HD constexpr int CZ(int i) { return Lattice::cz[i]; }
HLBM_HD constexpr int OPP(int i) { return Lattice::opp[i]; }
HLBM_HD constexpr double WD(int i) { return 1.0 / Lattice::wden[i]; }
HLBM_HD constexpr float WF(int i) { return (float)(1.0 / Lattice::wden[i]); }

// Incoming directions on each domain face, in the order the reference's
// post_bounce_back writes them (src/bcs/include/hippoLBM/bcs/bounce_back.hpp:115-140).
// face = dim*2 + side (side 0 = low face, plane "_0"; 1 = high face, "_l").
struct Faces {
  static constexpr int in[6][5] = {{1, 9, 7, 11, 13}, {2, 10, 8, 12, 14}, {3, 7, 10, 15, 17},
                                   {4, 8, 9, 16, 18}, {5, 14, 11, 18, 15}, {6, 13, 12, 17, 16}};
};

// Slot of direction i in the cavity coefficient table of a z face
// (cavity_coeff, src/bcs/include/hippoLBM/bcs/cavity.hpp:39-40): the cavity on
// the low z face adds to directions {5,14,11,18,15}, the high face to {6,13,12,17,16}.
HLBM_HD constexpr int cavity_slot(int i) {
  return i == 5 || i == 6 ? 0 : i == 14 || i == 13 ? 1 : i == 11 || i == 12 ? 2 : i == 18 || i == 17 ? 3 : 4;
}

// Speed class of direction i: 0 rest, 1 axis, 2 diagonal.
HLBM_HD constexpr int dir_class(int i) { return i == 0 ? 0 : (i < 7 ? 1 : 2); }

// The rest state that shifted storage is relative to, g_i = f_i - w_i - sig_i,
// per speed class.  BGK: sig = 0, the rest state is w.  MRT: hippoLBM's
// relaxation drives the epsilon moment of a fluid at rest to weps (mrt.hpp:96-
// 100; weps = 0 with the hard-coded rates) while w has epsilon = 3, so its
// rest state is w + (weps - 3) * (column epsilon of the inverse moment matrix):
// sig = (weps - 3) * {1/21, -1/63, 1/252}.  That column has no mass, momentum
// or e-moment, so rho = 1 + sum g and j = sum c g hold for either shift.
struct RestShift {
  double s[3];
};
inline RestShift rest_shift(int mrt, int mrt_custom, const double* rates) {
  RestShift r{{0.0, 0.0, 0.0}};
  if (mrt) {
    const double d = (mrt_custom ? rates[7] : 0.0) - 3.0;
    r.s[0] = d / 21.0;
    r.s[1] = -d / 63.0;
    r.s[2] = d / 252.0;
  }
  return r;
}

// Node layout in HBM: node (x, y, z) at x*st[0] + y*st[1] + z*st[2].  The axes
// are ordered fastest first by `ord`; the default is x fastest, then y, then z
// (box3d.hpp:127-132).  Other orders put the planes of a chosen axis in
// contiguous memory (e.g. x slowest for x-normal boundary planes).
struct NodeLayout {
  uint32_t st[3];
  int32_t ord[3];
};
inline NodeLayout make_layout(const int n[3], const int ord[3]) {
  NodeLayout L;
  uint32_t s = 1;
  for (int k = 0; k < 3; k++) {
    L.ord[k] = ord[k];
    L.st[ord[k]] = s;
    s *= (uint32_t)n[ord[k]];
  }
  return L;
}
inline NodeLayout default_layout(int nx, int ny, int nz) {
  const int n[3] = {nx, ny, nz}, ord[3] = {0, 1, 2};
  return make_layout(n, ord);
}
HLBM_HD inline uint32_t node_at(const NodeLayout& L, int x, int y, int z) {
  return (uint32_t)x * L.st[0] + (uint32_t)y * L.st[1] + (uint32_t)z * L.st[2];
}

// plane index (hlbm_plane) -> (dim, side)
HLBM_HD constexpr int plane_dim(int p) { return p < 2 ? 2 : (p < 4 ? 1 : 0); }
HLBM_HD constexpr int plane_side(int p) { return p & 1; }

}  // namespace hlbm
