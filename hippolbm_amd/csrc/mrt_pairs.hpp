// mrt_pairs.hpp -- the d'Humieres D3Q19 moment transform and its inverse
// written over opposite-direction pairs, for the FP32-compute forms.
//
// Directions 2k-1 and 2k are opposite.  Every row of the moment matrix M is
// either even (same coefficient on both members of a pair: rho, e, eps, the
// stress moments) or odd (opposite coefficients: j, q, m), so the moments need
// only the pair sums s and differences d, and the inverse f = M^T (D m) (the
// rows of M are orthogonal, D_k = 1 / |M_k|^2) gives each pair as E + O and
// E - O.  About 70 operations for the moments and 100 for the inverse, against
// 220 and 190 for the row-by-row sums the FP64 path keeps (hippoLBM accumulates
// each row in direction order, mrt.hpp:57-94 and 129-173, which the FP64 path
// reproduces bit for bit).  Host and device; tests/test_mrt_pairs.py checks
// both functions against the matrix on random vectors.
#pragma once
#include "d3q19.hpp"

namespace hlbm {

// d'Humieres D3Q19 basis (mrt.hpp:57-94, 129-173): rows of M in the moment order
// rho e eps jx qx jy qy jz qz pxx3 pixx3 pww piww pxy pyz pxz mx my mz
struct MrtBasis {
  static constexpr int M[19][19] = {
      {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
      {-30, -11, -11, -11, -11, -11, -11, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8},
      {12, -4, -4, -4, -4, -4, -4, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
      {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
      {0, -4, 4, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
      {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1},
      {0, 0, 0, -4, 4, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1},
      {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1},
      {0, 0, 0, 0, 0, -4, 4, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1},
      {0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
      {0, -4, -4, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
      {0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
      {0, 0, 0, -2, -2, 2, 2, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1},
      {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, -1, 1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
      {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1},
  };
  // inverse transform: f_i = sum_k sign*num/den * m_k; stored as signed numerator and denominator
  static constexpr int NUM[19][19] = {
      {1, -5, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
      {1, -11, -1, 1, -1, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
      {1, -11, -1, -1, 1, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
      {1, -11, -1, 0, 0, 1, -1, 0, 0, -1, 1, 1, -1, 0, 0, 0, 0, 0, 0},
      {1, -11, -1, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, 0, 0, 0, 0},
      {1, -11, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0, 0, 0},
      {1, -11, -1, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, 0, 0, 0, 0, 0, 0},
      {1, 4, 1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 0, 0, 1, -1, 0},
      {1, 4, 1, -1, -1, -1, -1, 0, 0, 1, 1, 1, 1, 1, 0, 0, -1, 1, 0},
      {1, 4, 1, 1, 1, -1, -1, 0, 0, 1, 1, 1, 1, -1, 0, 0, 1, 1, 0},
      {1, 4, 1, -1, -1, 1, 1, 0, 0, 1, 1, 1, 1, -1, 0, 0, -1, -1, 0},
      {1, 4, 1, 1, 1, 0, 0, 1, 1, 1, 1, -1, -1, 0, 0, 1, -1, 0, 1},
      {1, 4, 1, -1, -1, 0, 0, -1, -1, 1, 1, -1, -1, 0, 0, 1, 1, 0, -1},
      {1, 4, 1, 1, 1, 0, 0, -1, -1, 1, 1, -1, -1, 0, 0, -1, -1, 0, -1},
      {1, 4, 1, -1, -1, 0, 0, 1, 1, 1, 1, -1, -1, 0, 0, -1, 1, 0, 1},
      {1, 4, 1, 0, 0, 1, 1, 1, 1, -1, -1, 0, 0, 0, 1, 0, 0, 1, -1},
      {1, 4, 1, 0, 0, -1, -1, -1, -1, -1, -1, 0, 0, 0, 1, 0, 0, -1, 1},
      {1, 4, 1, 0, 0, 1, 1, -1, -1, -1, -1, 0, 0, 0, -1, 0, 0, 1, 1},
      {1, 4, 1, 0, 0, -1, -1, 1, 1, -1, -1, 0, 0, 0, -1, 0, 0, -1, -1},
  };
  static constexpr int DEN[19][19] = {
      {19, 399, 21, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
      {19, 2394, 63, 10, 10, 1, 1, 1, 1, 18, 18, 1, 1, 1, 1, 1, 1, 1, 1},
      {19, 2394, 63, 10, 10, 1, 1, 1, 1, 18, 18, 1, 1, 1, 1, 1, 1, 1, 1},
      {19, 2394, 63, 1, 1, 10, 10, 1, 1, 36, 36, 12, 12, 1, 1, 1, 1, 1, 1},
      {19, 2394, 63, 1, 1, 10, 10, 1, 1, 36, 36, 12, 12, 1, 1, 1, 1, 1, 1},
      {19, 2394, 63, 1, 1, 1, 1, 10, 10, 36, 36, 12, 12, 1, 1, 1, 1, 1, 1},
      {19, 2394, 63, 1, 1, 1, 1, 10, 10, 36, 36, 12, 12, 1, 1, 1, 1, 1, 1},
      {19, 1197, 252, 10, 40, 10, 40, 1, 1, 36, 72, 12, 24, 4, 1, 1, 8, 8, 1},
      {19, 1197, 252, 10, 40, 10, 40, 1, 1, 36, 72, 12, 24, 4, 1, 1, 8, 8, 1},
      {19, 1197, 252, 10, 40, 10, 40, 1, 1, 36, 72, 12, 24, 4, 1, 1, 8, 8, 1},
      {19, 1197, 252, 10, 40, 10, 40, 1, 1, 36, 72, 12, 24, 4, 1, 1, 8, 8, 1},
      {19, 1197, 252, 10, 40, 1, 1, 10, 40, 36, 72, 12, 24, 1, 1, 4, 8, 1, 8},
      {19, 1197, 252, 10, 40, 1, 1, 10, 40, 36, 72, 12, 24, 1, 1, 4, 8, 1, 8},
      {19, 1197, 252, 10, 40, 1, 1, 10, 40, 36, 72, 12, 24, 1, 1, 4, 8, 1, 8},
      {19, 1197, 252, 10, 40, 1, 1, 10, 40, 36, 72, 12, 24, 1, 1, 4, 8, 1, 8},
      {19, 1197, 252, 1, 1, 10, 40, 10, 40, 18, 36, 1, 1, 1, 4, 1, 1, 8, 8},
      {19, 1197, 252, 1, 1, 10, 40, 10, 40, 18, 36, 1, 1, 1, 4, 1, 1, 8, 8},
      {19, 1197, 252, 1, 1, 10, 40, 10, 40, 18, 36, 1, 1, 1, 4, 1, 1, 8, 8},
      {19, 1197, 252, 1, 1, 10, 40, 10, 40, 18, 36, 1, 1, 1, 4, 1, 1, 8, 8},
  };
};

// Moments m[0..18] (order rho e eps jx qx jy qy jz qz pxx3 pixx3 pww piww pxy
// pyz pxz mx my mz) of the populations g[0..18].
template <class C>
HLBM_HD inline void mrt_moments_pairs(const C* g, C* m) {
  const C sx = g[1] + g[2], dx = g[1] - g[2];      // (+1, 0, 0)
  const C sy = g[3] + g[4], dy = g[3] - g[4];      // (0, +1, 0)
  const C sz = g[5] + g[6], dz = g[5] - g[6];      // (0, 0, +1)
  const C sa = g[7] + g[8], da = g[7] - g[8];      // (+1, +1, 0)
  const C sb = g[9] + g[10], db = g[9] - g[10];    // (+1, -1, 0)
  const C sc = g[11] + g[12], dc = g[11] - g[12];  // (+1, 0, +1)
  const C sd = g[13] + g[14], dd = g[13] - g[14];  // (+1, 0, -1)
  const C se = g[15] + g[16], de = g[15] - g[16];  // (0, +1, +1)
  const C sf = g[17] + g[18], df = g[17] - g[18];  // (0, +1, -1)
  const C A = sx + sy + sz;
  const C Sxy = sa + sb, Sxz = sc + sd, Syz = se + sf;
  const C D = Sxy + Sxz + Syz;
  m[0] = g[0] + A + D;
  m[1] = C(-30) * g[0] - C(11) * A + C(8) * D;
  m[2] = C(12) * g[0] - C(4) * A + D;
  const C Pxy = da + db, Qxy = da - db;
  const C Pxz = dc + dd, Qxz = dc - dd;
  const C Pyz = de + df, Qyz = de - df;
  m[3] = dx + Pxy + Pxz;
  m[4] = m[3] - C(5) * dx;
  m[5] = dy + Qxy + Pyz;
  m[6] = m[5] - C(5) * dy;
  m[7] = dz + Qxz + Qyz;
  m[8] = m[7] - C(5) * dz;
  const C T = Sxy + Sxz, U = sy + sz;
  m[9] = C(2) * (sx - Syz) + (T - U);
  m[10] = C(-4) * sx + C(2) * (U - Syz) + T;
  const C V = sy - sz, W = Sxy - Sxz;
  m[11] = V + W;
  m[12] = C(-2) * V + W;
  m[13] = sa - sb;
  m[14] = se - sf;
  m[15] = sc - sd;
  m[16] = Pxy - Pxz;
  m[17] = Pyz - Qxy;
  m[18] = Qxz - Qyz;
}

// Populations f[0..18] of the moments m[0..18]: f = M^T (D m).
template <class C>
HLBM_HD inline void mrt_inverse_pairs(const C* m, C* f) {
  const C r = m[0] * C(1.0 / 19.0), e = m[1] * C(1.0 / 2394.0), ep = m[2] * C(1.0 / 252.0);
  const C jx = m[3] * C(0.1), qx = m[4] * C(0.025);
  const C jy = m[5] * C(0.1), qy = m[6] * C(0.025);
  const C jz = m[7] * C(0.1), qz = m[8] * C(0.025);
  const C p9 = m[9] * C(1.0 / 36.0), p10 = m[10] * C(1.0 / 72.0);
  const C p11 = m[11] * C(1.0 / 12.0), p12 = m[12] * C(1.0 / 24.0);
  const C p13 = m[13] * C(0.25), p14 = m[14] * C(0.25), p15 = m[15] * C(0.25);
  const C p16 = m[16] * C(0.125), p17 = m[17] * C(0.125), p18 = m[18] * C(0.125);
  f[0] = r - C(30) * e + C(12) * ep;
  // axis directions: M[.][1] = (1, -11, -4, 1, -4, 0, 0, 0, 0, 2, -4, 0, 0, ...)
  const C ba = r - C(11) * e - C(4) * ep;
  const C Ex = ba + C(2) * p9 - C(4) * p10;
  const C Y = C(2) * p10 - p9, Z = p11 - C(2) * p12;
  const C Ey = ba + Y + Z, Ez = ba + Y - Z;
  const C Ox = jx - C(4) * qx, Oy = jy - C(4) * qy, Oz = jz - C(4) * qz;
  f[1] = Ex + Ox;
  f[2] = Ex - Ox;
  f[3] = Ey + Oy;
  f[4] = Ey - Oy;
  f[5] = Ez + Oz;
  f[6] = Ez - Oz;
  // diagonal directions: M[.][7] = (1, 8, 1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 0, 0, 1, -1, 0)
  const C bd = r + C(8) * e + ep;
  const C P = p9 + p10, Wd = p11 + p12;
  const C Exy = bd + P + Wd, Exz = bd + P - Wd, Eyz = bd - C(2) * P;
  const C Jx = jx + qx, Jy = jy + qy, Jz = jz + qz;
  const C O7 = (Jx + Jy) + (p16 - p17), O9 = (Jx - Jy) + (p16 + p17);
  const C O11 = (Jx + Jz) + (p18 - p16), O13 = (Jx - Jz) - (p16 + p18);
  const C O15 = (Jy + Jz) + (p17 - p18), O17 = (Jy - Jz) + (p17 + p18);
  const C E7 = Exy + p13, E9 = Exy - p13;
  const C E11 = Exz + p15, E13 = Exz - p15;
  const C E15 = Eyz + p14, E17 = Eyz - p14;
  f[7] = E7 + O7;
  f[8] = E7 - O7;
  f[9] = E9 + O9;
  f[10] = E9 - O9;
  f[11] = E11 + O11;
  f[12] = E11 - O11;
  f[13] = E13 + O13;
  f[14] = E13 - O13;
  f[15] = E15 + O15;
  f[16] = E15 - O15;
  f[17] = E17 + O17;
  f[18] = E17 - O17;
}

}  // namespace hlbm
