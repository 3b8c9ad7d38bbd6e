// step_kernel.hpp -- the fused pull-scheme D3Q19 step for gfx950.
//
// One thread owns one lattice node.  Per step it
//   1. gathers the 19 post-collision populations streaming into the node
//      (pull: f_i(x) <- f*_i(x - c_i)), resolving in registers
//        - periodic axes by index wrap (replaces streaming + update_ghost,
//          src/collision/streaming.cu:107-111),
//        - non-periodic domain faces by on-node bounce-back f_i(x) <- f*_opp(i)(x)
//          (the implicit bounce-back of the reference's swap streaming,
//          src/collision/include/hippoLBM/collision/streaming.hpp:57-66, which
//          pre_bounce_back/post_bounce_back re-impose, bounce_back.hpp:92-156),
//          plus the moving-wall term of cavity_z_0 / cavity_z_l (cavity.hpp:57-87),
//        - obstacle cells by half-way bounce-back f_i(x) <- f*_opp(i)(x) when the
//          upstream node is a wall (wall_bounce_back, bounce_back.hpp:163-185),
//          except across a periodic face, where special_kernel reproduces the
//          reference's bounce-back off a stale ghost node;
//   2. runs the post-stream boundary program (post_bounce_back re-application,
//      lid_driven_cavity, neumann) in the operator-list order of the deck;
//   3. computes rho and u (macro_variables, macro_variables.hpp:48-79);
//   4. collides (bgk.hpp:47-62 or mrt.hpp:42-193) and stores f*.
//
// The FP64 instantiation keeps the reference's evaluation order and is built
// without floating-point contraction, so its populations are bit-identical to
// the reference's CPU path.  The reduced-precision instantiations compute in
// FP32 on shifted populations g_i = f_i - w_i stored as FP32/FP16/BF16.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "d3q19.hpp"
#include "mrt_pairs.hpp"

namespace hlbm {

// A/B modes read buffer A and write buffer B.  AA modes work in place on one
// buffer: AA_EVEN reads a node's own populations and writes them back into the
// opposite slots; AA_ODD gathers from those swapped slots and pushes its
// post-collision populations to the neighbours' natural slots (bounce-back
// targets stay at the node); AA_STREAM is AA_ODD without collision into a
// second buffer; AA_FIX applies the boundary program to a natural state.
enum StepMode { MODE_FUSED = 0, MODE_COLLIDE = 1, MODE_STREAM = 2, MODE_AA_EVEN = 3, MODE_AA_ODD = 4,
                MODE_AA_STREAM = 5, MODE_AA_FIX = 6 };

__host__ __device__ constexpr bool mode_reads_own(int m) {
  return m == MODE_COLLIDE || m == MODE_AA_EVEN || m == MODE_AA_FIX;
}
__host__ __device__ constexpr bool mode_collides(int m) {
  return m == MODE_FUSED || m == MODE_COLLIDE || m == MODE_AA_EVEN || m == MODE_AA_ODD;
}
__host__ __device__ constexpr bool mode_aa(int m) { return m >= MODE_AA_EVEN; }
// in-place AA steps leave wall nodes alone when wall_bounce_back is active:
// nothing ever reads a wall's populations then
__host__ __device__ constexpr bool mode_skips_walls(int m) { return m == MODE_AA_EVEN || m == MODE_AA_ODD; }
// The kernels of one step: the AA outflow capture, the program-plane kernel,
// and the bulk kernel with special_kernel.  The host may launch the program
// planes on a second stream beside the bulk (they touch disjoint nodes).
enum StepPart { PART_CAPTURE = 1, PART_BOUNDARY = 2, PART_BULK = 4, PART_ALL = 7 };
enum ProgKind { PK_LID = 1, PK_NEUMANN = 2, PK_REBB = 3, PK_PRESSURE = 4, PK_OUTFLOW = 5 };
enum CollKind { COLL_BGK = 0, COLL_MRT = 1 };

constexpr int MAX_PROG = 12;

struct ProgEntry {
  int32_t kind;
  int32_t plane;  // hlbm_plane for LID/NEUMANN; unused for REBB
  double U[3];
};

constexpr int MAX_SET = 4;

struct StepParams {
  int32_t nx, ny, nz;
  int32_t per[3];
  uint32_t N;           // nodes
  uint32_t pad0;
  uint64_t stride;      // elements between two population planes
  NodeLayout lay;       // node (x, y, z) at lay.st . (x, y, z); lay.ord: the axes, fastest first
  int32_t ln[3];        // nodes along lay.ord[0..2] (the bulk kernel's lane, row and grid axes)
  const void* fin;
  void* fout;
  const uint8_t* flags; // 0 fluid, 1 wall
  const uint32_t* bbm;  // per node: bit 0 wall, bit i (1..18) f_i bounces off an upstream wall (wall_bb)
  double* m0;           // rho (written when write_macro)
  double* m1;           // u, SoA 3*N
  int32_t write_macro;
  int32_t wall_bb;      // wall_bounce_back active
  int32_t plane_bb;     // pre_bounce_back + post_bounce_back active
  int32_t cav_on[2];    // cavity on the low / high z face
  int32_t cav_gather;   // post_bounce_back is the first post-stream operator: the gather adds the cavity term
  int32_t nprog;
  int32_t prog_planes;  // OR of plane bits used by the program
  int32_t mrt_custom;
  int32_t gx, gy, gz;   // tile grid of the bulk kernel (tiles of blockDim.x * blockDim.y nodes)
  int32_t xcd_remap;    // map consecutive dispatch slots to the same XCD's run of tiles
  int32_t skip_program; // boundary nodes processed without the post-stream program (AA_EVEN on a clean state)
  int32_t any_wall;     // some node is a wall
  int32_t oflow_n;      // nodes per plane slot of oflow
  void* oflow;          // outflow values captured before an in-place AA step (C type), or null
  uint32_t seed;        // stochastic rounding: hash of the time step (0 with sr = 0)
  int32_t sr;           // 16-bit storage rounds stochastically
  int32_t has_force;    // some c_i . F is non-zero (otherwise the force term adds exactly +0)
  int32_t force_wide;   // take the 64-bit-address bulk variant even when 32-bit offsets would do
  int32_t pair32;       // (stride + N) * element size < 2^32: bounce-back waves gather without branches
  float k1;             // 1 - 1/tau (FP32 BGK: g' = k1 g + kw[class] (dr + rho t))
  float kw[3];          // w_class / tau for the rest, axis and diagonal classes
  float sig[3];         // shifted storage: rest state w_i + sig[class(i)] (RestShift)
  float f3[3];          // 3 F for the FP32-compute forms (the force term is rho w_i c_i . 3F)
  float cavf[2][5];     // cav in FP32 for the FP32-compute forms (uniform registers, no conversions)
  float pad_[3];        // keeps the FP64 fields below at the offsets the FP64 kernels were tuned with
  double cav[2][5];
  double tau;
  double inv_tau;       // 1 / tau (the MRT default s9 = s13, mrt.hpp:43)
  double F[3];
  double ef[Q];         // c_i . F, evaluated in the reference's order on the host
  double mrt[10];       // s1, s2, s4, s9, s10, s13, s16, weps, wepsj, wxx (mrt_custom)
  ProgEntry prog[MAX_PROG];
  // nodes next to a wall across a periodic face (wall_bounce_back): special_kernel
  uint32_t nspec;              // listed nodes
  const uint32_t* spec_idx;    // node index
  const uint32_t* spec_mask;   // bit i: f_i takes the node's own f_opp(i) of the previous step
  void* spec_val;              // those values (C type), [i * nspec + k]
  int32_t boundary_first;      // launch the program-plane kernel before the bulk kernel
  int32_t parts;               // kernels of the step to launch (StepPart bits)
  int32_t nset;                // start_iteration set_distribution boxes
  int32_t set_box[MAX_SET][6]; // node box, inclusive: x0 y0 z0 x1 y1 z1
  double set_val[MAX_SET];     // f_i = value * w_i
  float mrtf[10];              // the MRT rates in FP32 (defaults or custom), for the FP32-compute forms
};

// ---------------------------------------------------------------------------
// storage types
// ---------------------------------------------------------------------------
struct bf16s {
  uint16_t v;
};

template <class S>
struct StoreOps;

// ld/st convert between the stored and the compute type with round-to-nearest-
// even; st_sr rounds stochastically with the uniform 32-bit word r: the bits
// the narrower format drops are replaced by random ones before truncation, so
// a value is rounded up with probability equal to its distance from the lower
// neighbour (unbiased, and exact for representable values).  FP16 stores
// g * 2^13: the shifted populations are small (|g| < 0.1), and the scale keeps
// them above FP16's subnormal range, where its 11-bit precision holds.
template <>
struct StoreOps<double> {
  using C = double;
  static constexpr bool shifted = false;
  static constexpr bool narrow16 = false;
  static __device__ __forceinline__ C ld(const double* p, size_t i) { return p[i]; }
  static __device__ __forceinline__ void st(double* p, size_t i, C v) { p[i] = v; }
  static __device__ __forceinline__ void st_sr(double* p, size_t i, C v, uint32_t) { p[i] = v; }
};
// The reduced-precision forms store with the non-temporal hint (nothing reads
// the new populations before the next step; measured faster for FP32 and
// FP16, slower for FP64, which therefore stores normally).
template <>
struct StoreOps<float> {
  using C = float;
  static constexpr bool shifted = true;
  static constexpr bool narrow16 = false;
  static __device__ __forceinline__ C ld(const float* p, size_t i) { return p[i]; }
  static __device__ __forceinline__ void st(float* p, size_t i, C v) { __builtin_nontemporal_store(v, p + i); }
  static __device__ __forceinline__ void st_sr(float* p, size_t i, C v, uint32_t) {
    __builtin_nontemporal_store(v, p + i);
  }
};
template <>
struct StoreOps<_Float16> {
  using C = float;
  static constexpr bool shifted = true;
  static constexpr bool narrow16 = true;
  static constexpr float SCALE = 8192.0f, INV_SCALE = 1.0f / 8192.0f;  // 2^13, 2^-13
  // fma(h, 2^-13, +0) is h * 2^-13 (exact) except that -0 loads as +0, and it
  // is one v_fma_mix_f32 (conversion and scaling together) instead of two
  static __device__ __forceinline__ C ld(const _Float16* p, size_t i) {
    return __builtin_fmaf((float)p[i], INV_SCALE, 0.0f);
  }
  static __device__ __forceinline__ void st(_Float16* p, size_t i, C v) {
    __builtin_nontemporal_store((_Float16)(v * SCALE), p + i);
  }
  static __device__ __forceinline__ void st_sr(_Float16* p, size_t i, C v, uint32_t r) {
    // FP16 keeps 10 of FP32's 23 mantissa bits: the 13 dropped ones are dithered
    // with the top 13 bits of r, i.e. (bits(v) + (r >> 19)) & ~0x1FFF, then
    // converted; gfx950's v_cvt_sr_f16_f32 does exactly this in one instruction
    st_sr_scaled(p, i, v * SCALE, r);
  }
  // the same for a value already multiplied by SCALE (the FP32 BGK collision
  // can fold the scale into its constants).  The instruction keeps the upper
  // half of its destination, which is not stored: the scaled value's own
  // register serves, no zeroed one is needed.
  static __device__ __forceinline__ void st_sr_scaled(_Float16* p, size_t i, C s, uint32_t r) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 h = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(h2, s), s, r, false);
    __builtin_nontemporal_store(h[0], p + i);
  }
};
template <>
struct StoreOps<bf16s> {
  using C = float;
  static constexpr bool shifted = true;
  static constexpr bool narrow16 = true;
  static __device__ __forceinline__ C ld(const bf16s* p, size_t i) {
    return __uint_as_float(((uint32_t)p[i].v) << 16);
  }
  static __device__ __forceinline__ void st(bf16s* p, size_t i, C v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    __builtin_nontemporal_store((uint16_t)(u >> 16), (uint16_t*)(p + i));
  }
  static __device__ __forceinline__ void st_sr(bf16s* p, size_t i, C v, uint32_t r) {
    // the 16 dropped bits dithered with the top 16 bits of r, (bits(v) + (r >> 16)) >> 16:
    // gfx950's v_cvt_sr_bf16_f32
    typedef short s2 __attribute__((ext_vector_type(2)));
    const s2 b = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(s2, v), v, r, false);
    __builtin_nontemporal_store((uint16_t)b[0], (uint16_t*)(p + i));
  }
};

// Random words for stochastic rounding: one integer hash per (node, step), then
// a Weyl sequence over the directions (the top bits, which the stores use, are
// uniform for each direction).
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
// The rounding word of direction i at a node: random with stochastic rounding,
// else half an ulp of the dropped bits (round to nearest, ties away from zero).
template <class S>
struct Rounder {
  uint32_t base = 0, inc = 0;
  template <class P>
  __device__ __forceinline__ Rounder(const P& p, uint32_t idx) {
    if constexpr (StoreOps<S>::narrow16) {
      if (p.sr) {
        base = mix32(idx ^ p.seed);
        inc = 0x9E3779B9u;
      } else {
        base = 0x80000000u;
      }
    }
  }
  __device__ __forceinline__ uint32_t word(int i) const { return base + (uint32_t)i * inc; }
};

// ---------------------------------------------------------------------------
// macroscopic variables
// ---------------------------------------------------------------------------
// FP64: rho and j accumulated in direction order (macro_variables.hpp:56-62).
// Terms with a zero lattice coefficient are skipped and +-1 coefficients become
// add/subtract; both are exact, so the sums are the reference's bit for bit.
__device__ __forceinline__ void macro_f64(const double* f, double& rho, double& ux, double& uy, double& uz) {
  rho = 0.0;
  ux = 0.0;
  uy = 0.0;
  uz = 0.0;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const double s = f[i];
    if (CX(i) == 1) ux += s;
    if (CX(i) == -1) ux -= s;
    if (CY(i) == 1) uy += s;
    if (CY(i) == -1) uy -= s;
    if (CZ(i) == 1) uz += s;
    if (CZ(i) == -1) uz -= s;
    rho += s;
  }
  if (rho > 1.0e-14) {
    ux /= rho;
    uy /= rho;
    uz /= rho;
  }
}

// Moving-wall coefficient j of the cavity face `side` in the compute type.
template <class C>
__device__ __forceinline__ C cav_coef(const StepParams& p, int side, int j) {
  if constexpr (std::is_same<C, double>::value)
    return p.cav[side][j];
  else
    return p.cavf[side][j];
}

// c_i . v in the reference's order (ex*vx + ey*vy) + ez*vz with zero terms dropped.
template <int i, class C>
__device__ __forceinline__ C cdot(C vx, C vy, C vz) {
  C r = C(0);
  bool first = true;
  if (CX(i) != 0) {
    r = CX(i) == 1 ? vx : -vx;
    first = false;
  }
  if (CY(i) != 0) {
    r = first ? (CY(i) == 1 ? vy : -vy) : (CY(i) == 1 ? r + vy : r - vy);
    first = false;
  }
  if (CZ(i) != 0) {
    r = first ? (CZ(i) == 1 ? vz : -vz) : (CZ(i) == 1 ? r + vz : r - vz);
  }
  return r;
}

template <int I = 0, int END = Q, class F>
__device__ __forceinline__ void static_for(F&& fn) {
  if constexpr (I < END) {
    fn(std::integral_constant<int, I>{});
    static_for<I + 1, END>(fn);
  }
}

// ---------------------------------------------------------------------------
// BGK (bgk.hpp:47-62), FP64 order-faithful
// ---------------------------------------------------------------------------
// a / tau, correctly rounded, from y = RN(1/tau) (an IEEE division on the host):
// q = RN(a y) is within one ulp of a/tau, r = a - q tau is exact in one fused
// multiply-add, and RN(q + r y) is the correctly rounded quotient (Markstein,
// IBM J. Res. Dev. 34 (1990) 111-119; a CPU check of 4.4e8 quotients against
// IEEE division found no difference).  Three instructions instead of the
// general division's scale / reciprocal / refinement sequence.  (For a = -0 it
// returns +0 where IEEE gives -0; f + (+-0) is the same for every f != -0.)
__device__ __forceinline__ double div_tau(double a, const StepParams& p) {
  const double q = a * p.inv_tau;
  const double r = __builtin_fma(-q, p.tau, a);
  return __builtin_fma(r, p.inv_tau, q);
}

__device__ __forceinline__ void bgk_f64(double* f, double rho, double ux, double uy, double uz, const StepParams& p) {
  const double u_squ = (ux * ux + uy * uy + uz * uz);
  static_for([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    const double w = WD(i);
    const double eu = cdot<i>(ux, uy, uz);
    const double feq = w * rho * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * u_squ);
    // F = 0 makes the force term exactly +0 (c_i . F is evaluated to +0.0), so
    // dropping it leaves every sum bit-identical
    if (p.has_force)
      f[i] += div_tau(feq - f[i], p) + 3. * rho * w * p.ef[i];
    else
      f[i] += div_tau(feq - f[i], p);
  });
}

// BGK on shifted populations g = f - w, FP32:
//   g' = g + omega (g_eq - g) = k1 g + kw_c (dr + rho t_i),
//   t_i = eu (3 + 4.5 eu) - 1.5 u^2,  g_eq = w_i (dr + rho t_i),
// with k1 = 1 - omega and kw_c = omega w_c per weight class; kw_c rho and
// kw_c dr are formed once per node, so a population costs four fused
// operations.  SO (a power of two) scales the result exactly: with FP16's
// storage scale the store needs no multiply, and every form rounds alike.
// The body-force term rho w_i c_i . 3F is a separate pass under the solve's
// uniform force flag (inside the update it would cost a multiply-add per
// population on every node and hold 19 scalar registers).
template <bool SCALED>
__device__ __forceinline__ void bgk_shift(float* g, float dr, float ux, float uy, float uz, const StepParams& p) {
  constexpr float SO = SCALED ? StoreOps<_Float16>::SCALE : 1.0f;
  const float rho = 1.0f + dr;
  const float usq15 = 1.5f * (ux * ux + uy * uy + uz * uz);
  const float k1 = p.k1 * SO;
  // the classes are contiguous (0 | 1..6 | 7..18): their constants are formed
  // as each begins, so only one pair is live
  float kr = 0.0f, kd = 0.0f;
  static_for([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    if constexpr (i == 0 || i == 1 || i == 7) {
      const float kw = p.kw[dir_class(i)] * SO;
      kr = kw * rho;
      kd = kw * dr;
    }
    const float eu = cdot<i>(ux, uy, uz);
    const float t = eu * (3.0f + 4.5f * eu) - usq15;
    g[i] = kr * t + (k1 * g[i] + kd);
  });
  if (p.has_force) {
    const float fx = p.f3[0], fy = p.f3[1], fz = p.f3[2];
    static_for([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      g[i] += (rho * (WF(i) * SO)) * cdot<i>(fx, fy, fz);
    });
  }
}

// rho and j = sum c_i g_i of FP32 populations: zero coefficients skipped and
// +-1 coefficients as add / subtract.  The sums start from -0, which x + (-0)
// leaves unchanged for every x, so the first term costs nothing (a start from
// +0, or x * 0 terms, are not folded under IEEE semantics: 30 instructions per
// node).
__device__ __forceinline__ void moments_f32(const float* g, float& dr, float& jx, float& jy, float& jz) {
  dr = jx = jy = jz = -0.0f;
  static_for([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    dr += g[i];
    if constexpr (CX(i) > 0) jx += g[i];
    if constexpr (CX(i) < 0) jx -= g[i];
    if constexpr (CY(i) > 0) jy += g[i];
    if constexpr (CY(i) < 0) jy -= g[i];
    if constexpr (CZ(i) > 0) jz += g[i];
    if constexpr (CZ(i) < 0) jz -= g[i];
  });
}

// 1 / x for x near 1 (a lattice density): the hardware reciprocal (1 ulp)
// and one Newton step, which leaves it within half an ulp plus a few units of
// 2^-46; three instructions instead of the IEEE division's eleven.
__device__ __forceinline__ float recip_f32(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
}

// ---------------------------------------------------------------------------
// MRT (mrt.hpp:42-193).  Moment rows and the inverse transform are the
// d'Humieres D3Q19 basis; the FP64 path accumulates each row in increasing
// direction order and each inverse row in moment order, as the reference does.
// moment order: rho e eps jx qx jy qy jz qz pxx3 pixx3 pww piww pxy pyz pxz mx my mz
// ---------------------------------------------------------------------------
// MrtBasis (the matrix, its exact inverse as numerators/denominators): mrt_pairs.hpp

constexpr bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// acc +- c * v for a coefficient c = sgn * 2^e: the product is exact, so one
// fused multiply-add rounds exactly as the separate product and sum do.  (For
// populations and moments of a lattice fluid the product never reaches the
// subnormal range, where it would stop being exact.)
template <class C>
__device__ __forceinline__ C acc_exact(C acc, C c, C v) {
  if constexpr (std::is_same<C, double>::value)
    return __builtin_fma(c, v, acc);
  else
    return acc + c * v;
}

// Row k of M applied to f, accumulated left to right over the non-zero entries.
template <int k, class C>
__device__ __forceinline__ C moment(const C* f) {
  C acc = C(0);
  bool first = true;
  static_for([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    constexpr int c = MrtBasis::M[k][i];
    if constexpr (c != 0) {
      constexpr int a = c < 0 ? -c : c;
      if (first) {
        acc = a == 1 ? (c < 0 ? -f[i] : f[i]) : C(c) * f[i];
        first = false;
      } else if constexpr (a == 1) {
        acc = c < 0 ? acc - f[i] : acc + f[i];
      } else if constexpr (pow2(a)) {
        acc = acc_exact(acc, C(c), f[i]);
      } else {
        const C term = C(a) * f[i];
        acc = c < 0 ? acc - term : acc + term;
      }
    }
  });
  return acc;
}

// Inverse row i applied to the post-collision moments mo[19].
template <int i, class C>
__device__ __forceinline__ C inverse_row(const C* mo) {
  C acc = C(0);
  bool first = true;
  static_for([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    constexpr int num = MrtBasis::NUM[i][k];
    if constexpr (num != 0) {
      constexpr int an = num < 0 ? -num : num;
      constexpr int den = MrtBasis::DEN[i][k];
      // |num|/den is a power of two (1, 1/4, 1/8) when den = |num| * 2^e
      constexpr bool exact = den % an == 0 && pow2(den / an);
      const C coef = C((double)an / (double)den);
      if (first) {
        const C term = coef * mo[k];
        acc = num < 0 ? -term : term;
        first = false;
      } else if constexpr (exact) {
        acc = acc_exact(acc, num < 0 ? -coef : coef, mo[k]);
      } else {
        const C term = coef * mo[k];
        acc = num < 0 ? acc - term : acc + term;
      }
    }
  });
  return acc;
}

struct MrtRates {
  double s1, s2, s4, s9, s10, s13, s16, weps, wepsj, wxx;
};
struct MrtRatesF {
  float s1, s2, s4, s9, s10, s13, s16, weps, wepsj, wxx;
};

__device__ __forceinline__ MrtRatesF mrt_rates_f32(const StepParams& p) {
  const float* r = p.mrtf;
  return MrtRatesF{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9]};
}

__device__ __forceinline__ MrtRates mrt_rates(const StepParams& p) {
  MrtRates r;
  if (p.mrt_custom) {
    r.s1 = p.mrt[0];
    r.s2 = p.mrt[1];
    r.s4 = p.mrt[2];
    r.s9 = p.mrt[3];
    r.s10 = p.mrt[4];
    r.s13 = p.mrt[5];
    r.s16 = p.mrt[6];
    r.weps = p.mrt[7];
    r.wepsj = p.mrt[8];
    r.wxx = p.mrt[9];
  } else {  // mrt.hpp:43-47; 1 / tau computed once on the host (IEEE division either way)
    r.s9 = p.inv_tau;
    r.s13 = r.s9;
    r.s1 = 1.19;
    r.s2 = 1.4;
    r.s4 = 1.2;
    r.s10 = r.s2;
    r.s16 = 1.98;
    r.weps = 0;
    r.wepsj = -475. / 63.;
    r.wxx = 0;
  }
  return r;
}

// Relaxation of the moments (mrt.hpp:96-126).  For the shifted form the
// caller passes the moments of g = f - w; the rest state w has the moments
// rho = 1, e = -11, eps = 3 (all others 0), which enter through rest_e/rest_eps
// and rest_rho (all zero for the unshifted FP64 path).
template <class C, class R>
__device__ __forceinline__ void mrt_relax(C* m, const R& r, C rest_rho, C rest_e, C rest_eps) {
  const C rho = m[0] + rest_rho;
  const C e = m[1], eps = m[2], jx = m[3], qx = m[4], jy = m[5], qy = m[6], jz = m[7], qz = m[8];
  const C pxx3 = m[9], pixx3 = m[10], pww = m[11], piww = m[12], pxy = m[13], pyz = m[14], pxz = m[15];
  const C mx = m[16], my = m[17], mz = m[18];
  const C s1 = C(r.s1), s2 = C(r.s2), s4 = C(r.s4), s9 = C(r.s9), s10 = C(r.s10), s13 = C(r.s13), s16 = C(r.s16);
  const C jx2 = jx * jx, jy2 = jy * jy, jz2 = jz * jz;
  const C jj = (jx2 + jy2 + jz2);
  // rho0 = 1: 19/rho0, wepsj/rho0 and 1/rho0 are exact
  const C e_eq = C(-11) * rho + C(19) * jj - rest_e;
  m[1] = e - s1 * (e - e_eq);
  const C eps_eq = C(r.weps) * rho + C(r.wepsj) * jj - rest_eps;
  m[2] = eps - s2 * (eps - eps_eq);
  m[4] = qx - s4 * (qx + C(2. / 3.) * jx);
  m[6] = qy - s4 * (qy + C(2. / 3.) * jy);
  m[8] = qz - s4 * (qz + C(2. / 3.) * jz);
  const C pxx3_eq = (C(2) * jx2 - (jy2 + jz2));
  m[9] = pxx3 - s9 * (pxx3 - pxx3_eq);
  m[10] = pixx3 - s10 * (pixx3 - C(r.wxx) * pxx3_eq);
  const C pww_eq = (jy2 - jz2);
  m[11] = pww - s9 * (pww - pww_eq);
  m[12] = piww - s10 * (piww - C(r.wxx) * pww_eq);
  const C pxy_eq = jx * jy;
  m[13] = pxy - s13 * (pxy - pxy_eq);
  const C pyz_eq = jy * jz;
  m[14] = pyz - s13 * (pyz - pyz_eq);
  const C pxz_eq = jx * jz;
  m[15] = pxz - s13 * (pxz - pxz_eq);
  m[16] = mx - s16 * mx;
  m[17] = my - s16 * my;
  m[18] = mz - s16 * mz;
}

template <class C>
__device__ __forceinline__ void mrt_collide(C* f, C rho_force, const StepParams& p, bool shifted) {
  C m[19];
  if constexpr (std::is_same<C, double>::value) {
    static_for([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      m[k] = moment<k>(f);
    });
    const MrtRates r = mrt_rates(p);
    // shifted storage holds g = f - (w + sig): the rest state's moments are
    // rho = 1, e = -11 and eps = weps (RestShift)
    if (shifted)
      mrt_relax<C>(m, r, C(1), C(-11), C(r.weps));
    else
      mrt_relax<C>(m, r, C(0), C(0), C(0));
    static_for([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      f[i] = inverse_row<i>(m);
    });
    // mrt.hpp:187-191, force term with rho from macro_variables (exactly +0 when F = 0)
    if (p.has_force) {
      static_for([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        f[i] += C(3.) * rho_force * C(WD(i)) * C(p.ef[i]);
      });
    }
  } else {
    // FP32: the pair forms of the transform and FP32 rates (same basis and
    // relaxation, far fewer operations); the force term from 3F as for BGK
    mrt_moments_pairs(f, m);
    const MrtRatesF r = mrt_rates_f32(p);
    if (shifted)
      mrt_relax<C>(m, r, C(1), C(-11), C(r.weps));
    else
      mrt_relax<C>(m, r, C(0), C(0), C(0));
    mrt_inverse_pairs(m, f);
    if (p.has_force) {
      const float fx = p.f3[0], fy = p.f3[1], fz = p.f3[2];
      static_for([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        f[i] += (rho_force * WF(i)) * cdot<i>(fx, fy, fz);
      });
    }
  }
}

// ---------------------------------------------------------------------------
// post-stream boundary operators on one node
// ---------------------------------------------------------------------------
// lid_driven_cavity (lid_driven_cavity.hpp:52-62): f_i = feq(rho_local, U)
template <class C, bool SHIFT>
__device__ __forceinline__ void lid_node(C* f, const double* U, const float* sig) {
  C rho = C(0);
#pragma unroll
  for (int i = 0; i < Q; i++) rho += f[i];
  const C Ux = C(U[0]), Uy = C(U[1]), Uz = C(U[2]);
  const C u_squ = Ux * Ux + Uy * Uy + Uz * Uz;
  if constexpr (!SHIFT) {
    static_for([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const C eu = C(CX(i)) * Ux + C(CY(i)) * Uy + C(CZ(i)) * Uz;
      f[i] = C(WD(i)) * rho * (C(1.) + C(3.) * eu + C(4.5) * eu * eu - C(1.5) * u_squ);
    });
  } else {
    const C rhot = C(1) + rho;  // rho holds sum g = rho - 1
    static_for([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const C eu = C(CX(i)) * Ux + C(CY(i)) * Uy + C(CZ(i)) * Uz;
      f[i] = C(WD(i)) * (rho + rhot * (C(3.) * eu + C(4.5) * eu * eu - C(1.5) * u_squ)) - C(sig[dir_class(i)]);
    });
  }
}

// neumann (Zou/He velocity boundary, neumann.hpp:38-258), one face.
// SHIFT adds the rest-state weight sum (=1) to the density numerator.
// PRESSURE: the Zou/He density boundary (no counterpart in the reference):
// rho = rho_in is prescribed, the tangential velocity is zero and the normal
// velocity follows from the same known-population sum K as the velocity form,
// rho (1 -+ u_n) = K, i.e. u_n = +-(1 - K / rho_in); the unknown populations
// then take the velocity form's expressions.
template <class C, bool SHIFT, bool PRESSURE = false>
__device__ __forceinline__ void neumann_node(C* f, int plane, const double* U, double rho_in = 1.0) {
  C ux = PRESSURE ? C(0) : C(U[0]), uy = PRESSURE ? C(0) : C(U[1]), uz = PRESSURE ? C(0) : C(U[2]);
  const C base = SHIFT ? C(1) : C(0);
  const C h = C(1. / 2.), t = C(1. / 3.), s = C(1. / 6.);
  const C rin = C(rho_in);
  // rho from the known sum K over (1 + un) (high face) or (1 - un) (low face)
  auto density = [&](C K, C& un, bool high) -> C {
    if constexpr (PRESSURE) {
      un = high ? K / rin - C(1.) : C(1.) - K / rin;
      return rin;
    } else {
      return high ? K / (C(1.) + un) : K / (C(1.) - un);
    }
  };
  switch (plane) {
    case 5: {  // plan_yz_l
      const C rho = density((f[3] + f[4] + f[5] + f[6] + f[15] + f[17] + f[18] + f[16] + f[0] +
                             C(2.) * (f[1] + f[7] + f[9] + f[11] + f[13])) + base, ux, true);
      const C nyx = h * (f[3] + f[15] + f[17] - (f[4] + f[18] + f[16])) - t * rho * uy;
      const C nzx = h * (f[5] + f[18] + f[15] - (f[6] + f[17] + f[16])) - t * rho * uz;
      f[2] = f[1] - t * rho * ux;
      f[10] = f[9] + s * rho * (-ux + uy) - nyx;
      f[8] = f[7] + s * rho * (-ux - uy) + nyx;
      f[12] = f[11] + s * rho * (-ux - uz) + nzx;
      f[14] = f[13] + s * rho * (-ux + uz) - nzx;
      break;
    }
    case 4: {  // plan_yz_0
      const C rho = density((f[3] + f[4] + f[5] + f[6] + f[15] + f[17] + f[18] + f[16] + f[0] +
                             C(2.) * (f[2] + f[10] + f[8] + f[14] + f[12])) + base, ux, false);
      const C nyx = h * (f[3] + f[15] + f[17] - (f[4] + f[18] + f[16])) - t * rho * uy;
      const C nzx = h * (f[5] + f[18] + f[15] - (f[6] + f[17] + f[16])) - t * rho * uz;
      f[1] = f[2] + t * rho * ux;
      f[9] = f[10] + s * rho * (ux - uy) + nyx;
      f[7] = f[8] + s * rho * (ux + uy) - nyx;
      f[11] = f[12] + s * rho * (ux + uz) - nzx;
      f[13] = f[14] + s * rho * (ux - uz) + nzx;
      break;
    }
    case 3: {  // plan_xz_l
      const C rho = density((f[1] + f[2] + f[5] + f[6] + f[11] + f[13] + f[14] + f[12] + f[0] +
                             C(2.) * (f[3] + f[7] + f[10] + f[15] + f[17])) + base, uy, true);
      const C nxy = h * (f[1] + f[11] + f[13] - (f[2] + f[14] + f[12])) - t * rho * ux;
      const C nzy = h * (f[5] + f[11] + f[14] - (f[6] + f[13] + f[12])) - t * rho * uz;
      f[4] = f[3] - t * rho * uy;
      f[8] = f[7] + s * rho * (-uy - ux) + nxy;
      f[9] = f[10] + s * rho * (-uy + ux) - nxy;
      f[16] = f[15] + s * rho * (-uy - uz) + nzy;
      f[18] = f[17] + s * rho * (-uy + uz) - nzy;
      break;
    }
    case 2: {  // plan_xz_0
      const C rho = density((f[1] + f[2] + f[5] + f[6] + f[11] + f[13] + f[14] + f[12] + f[0] +
                             C(2.) * (f[4] + f[9] + f[8] + f[18] + f[16])) + base, uy, false);
      const C nxy = h * (f[1] + f[11] + f[13] - (f[2] + f[14] + f[12])) - t * rho * ux;
      const C nzy = h * (f[5] + f[11] + f[14] - (f[6] + f[13] + f[12])) - t * rho * uz;
      f[3] = f[4] + t * rho * uy;
      f[7] = f[8] + s * rho * (uy + ux) - nxy;
      f[10] = f[9] + s * rho * (uy - ux) + nxy;
      f[15] = f[16] + s * rho * (uy + uz) - nzy;
      f[17] = f[18] + s * rho * (uy - uz) + nzy;
      break;
    }
    case 1: {  // plan_xy_l
      const C rho = density((f[0] + f[1] + f[2] + f[3] + f[4] + f[7] + f[9] + f[10] + f[8] +
                             C(2.) * (f[5] + f[11] + f[14] + f[15] + f[18])) + base, uz, true);
      const C nxz = h * (f[1] + f[7] + f[9] - (f[2] + f[10] + f[8])) - t * rho * ux;
      const C nyz = h * (f[3] + f[7] + f[10] - (f[4] + f[9] + f[8])) - t * rho * uy;
      f[6] = f[5] - t * rho * uz;
      f[13] = f[14] + s * rho * (-uz + ux) - nxz;
      f[12] = f[11] + s * rho * (-uz - ux) + nxz;
      f[17] = f[18] + s * rho * (-uz + uy) - nyz;
      f[16] = f[15] + s * rho * (-uz - uy) + nyz;
      break;
    }
    default: {  // 0: plan_xy_0
      const C rho = density((f[0] + f[1] + f[2] + f[3] + f[4] + f[7] + f[9] + f[10] + f[8] +
                             C(2.) * (f[6] + f[13] + f[12] + f[17] + f[16])) + base, uz, false);
      const C nxz = h * (f[1] + f[7] + f[9] - (f[2] + f[10] + f[8])) - t * rho * ux;
      const C nyz = h * (f[3] + f[7] + f[10] - (f[4] + f[9] + f[8])) - t * rho * uy;
      f[5] = f[6] + t * rho * uz;
      f[11] = f[12] + s * rho * (uz + ux) - nxz;
      f[14] = f[13] + s * rho * (uz - ux) + nxz;
      f[15] = f[16] + s * rho * (uz + uy) - nyz;
      f[18] = f[17] + s * rho * (uz - uy) + nyz;
      break;
    }
  }
}

// ---------------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------------
// Plane bits of a node: bit p set when the node lies on plane p (hlbm_plane order).
// Index of node (x, y, z) inside plane pl: a + b*na, a the fastest in-plane
// axis (the order boundary_kernel walks a plane in).
__device__ __forceinline__ uint32_t plane_slot(int pl, int x, int y, int z, int nx, int ny) {
  const int dim = plane_dim(pl);
  const uint32_t na = dim == 0 ? (uint32_t)ny : (uint32_t)nx;
  const uint32_t a = dim == 0 ? (uint32_t)y : (uint32_t)x;
  const uint32_t b = dim == 2 ? (uint32_t)y : (uint32_t)z;
  return b * na + a;
}

__device__ __forceinline__ uint32_t node_index(const StepParams& p, int x, int y, int z) {
  return node_at(p.lay, x, y, z);
}

// The node's index in the default layout (x fastest): stochastic rounding is
// keyed on it, so reduced-precision results do not depend on the layout.
__device__ __forceinline__ uint32_t canonical_index(const StepParams& p, int x, int y, int z) {
  return ((uint32_t)z * (uint32_t)p.ny + (uint32_t)y) * (uint32_t)p.nx + (uint32_t)x;
}

// The coordinates of a node from its index (lay.ord picks which of the three
// layout coordinates is x, y and z; the choices are uniform selects).
__device__ __forceinline__ void node_coords(const StepParams& p, uint32_t idx, int& x, int& y, int& z) {
  const uint32_t n0 = (uint32_t)p.ln[0], n01 = n0 * (uint32_t)p.ln[1];
  const uint32_t c2 = idx / n01, r = idx - c2 * n01, c1 = r / n0, c0 = r - c1 * n0;
  x = (int)(p.lay.ord[0] == 0 ? c0 : (p.lay.ord[1] == 0 ? c1 : c2));
  y = (int)(p.lay.ord[0] == 1 ? c0 : (p.lay.ord[1] == 1 ? c1 : c2));
  z = (int)(p.lay.ord[0] == 2 ? c0 : (p.lay.ord[1] == 2 ? c1 : c2));
}

__device__ __forceinline__ int node_planes(int x, int y, int z, int nx, int ny, int nz) {
  return (z == 0 ? 1 : 0) | (z == nz - 1 ? 2 : 0) | (y == 0 ? 4 : 0) | (y == ny - 1 ? 8 : 0) | (x == 0 ? 16 : 0) |
         (x == nx - 1 ? 32 : 0);
}

// Phase 1: the 19 populations arriving at the node (pull + bounce-back), or
// the node's own populations when the state is pre-collision (MODE_COLLIDE).
//
// pull_one: the post-stream value of direction I at node (x, y, z).
// A/B: f*_i(s) sits in plane i at s and f*_opp(i)(x) in plane opp(i) at x.
// AA (swapped state): f*_i(s) sits in plane opp(i) at s, f*_opp(i)(x) in plane i at x.
// Modes that read the node's own populations return its natural slot.
template <class S, int MODE, int I>
__device__ __forceinline__ typename StoreOps<S>::C pull_one(const StepParams& p, int x, int y, int z, uint32_t idx,
                                                            bool wall) {
  using O = StoreOps<S>;
  using C = typename O::C;
  const S* __restrict__ fin = (const S*)p.fin;
  const size_t st = p.stride;
  if constexpr (mode_reads_own(MODE)) {
    return O::ld(fin, (size_t)I * st + idx);
  } else {
    constexpr bool AA = mode_aa(MODE);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    // upstream coordinates; -1 marks a non-periodic face (bounce-back)
    int sx = x, sy = y, sz = z;
    if constexpr (CX(I) == 1) sx = x > 0 ? x - 1 : (p.per[0] ? nx - 1 : -1);
    if constexpr (CX(I) == -1) sx = x < nx - 1 ? x + 1 : (p.per[0] ? 0 : -1);
    if constexpr (CY(I) == 1) sy = y > 0 ? y - 1 : (p.per[1] ? ny - 1 : -1);
    if constexpr (CY(I) == -1) sy = y < ny - 1 ? y + 1 : (p.per[1] ? 0 : -1);
    if constexpr (CZ(I) == 1) sz = z > 0 ? z - 1 : (p.per[2] ? nz - 1 : -1);
    if constexpr (CZ(I) == -1) sz = z < nz - 1 ? z + 1 : (p.per[2] ? 0 : -1);
    if (sx < 0 || sy < 0 || sz < 0) {
      C v = O::ld(fin, (size_t)(AA ? I : OPP(I)) * st + idx);
      if constexpr (CZ(I) != 0) {
        constexpr int side = CZ(I) == 1 ? 0 : 1;
        if (sz < 0 && p.cav_gather && p.cav_on[side]) v += cav_coef<C>(p, side, cavity_slot(I));
      }
      return v;
    }
    const uint32_t sidx = node_index(p, sx, sy, sz);
    if (I != 0 && p.wall_bb && !wall && p.flags[sidx] != 0) return O::ld(fin, (size_t)(AA ? I : OPP(I)) * st + idx);
    return O::ld(fin, (size_t)(AA ? OPP(I) : I) * st + sidx);
  }
}

template <class S, int MODE>
__device__ __forceinline__ void gather_node(const StepParams& p, int x, int y, int z, uint32_t idx, bool wall,
                                            typename StoreOps<S>::C* f) {
  static_for([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    f[i] = pull_one<S, MODE, i>(p, x, y, z, idx, wall);
  });
}

// Phase 2: the post-stream boundary program (post_bounce_back re-application,
// lid_driven_cavity, neumann) in deck order, for a node on a program plane.
template <class S, int MODE>
__device__ __forceinline__ void apply_program(const StepParams& p, int x, int y, int z, uint32_t idx, bool wall,
                                              typename StoreOps<S>::C* f) {
  using O = StoreOps<S>;
  using C = typename O::C;
  constexpr bool SHIFT = O::shifted;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const S* __restrict__ fin = (const S*)p.fin;
  const size_t st = p.stride;
  {
      const int bmask = node_planes(x, y, z, nx, ny, nz);
      {
        for (int k = 0; k < p.nprog; k++) {
          const ProgEntry& e = p.prog[k];
          if (e.kind == PK_REBB) {
            // post_bounce_back: faces in x-left, x-right, y-left, y-right, z-left, z-right order
            static_for<0, 6>([&](auto fc) {
              constexpr int face = decltype(fc)::value;
              constexpr int dim = face >> 1, side = face & 1;
              constexpr int pl = dim == 2 ? side : (dim == 1 ? 2 + side : 4 + side);
              if (!p.per[dim] && ((bmask >> pl) & 1)) {
                static_for<0, 5>([&](auto jc) {
                  constexpr int j = decltype(jc)::value;
                  constexpr int qi = Faces::in[face][j];
                  // f*_opp(qi)(x): plane opp(qi) (A/B), plane qi (AA: swapped or natural bounce-back slot)
                  C v = O::ld(fin, (size_t)(mode_aa(MODE) ? qi : OPP(qi)) * st + idx);
                  // the natural AA state (AA_EVEN, AA_FIX) holds the value the odd step bounced
                  // into slot qi, which already carries the moving-wall term when the gather
                  // adds it (cav_gather); every other state holds the bare f*_opp(qi)(x)
                  constexpr bool natural = MODE == MODE_AA_EVEN || MODE == MODE_AA_FIX;
                  if (dim == 2 && p.cav_on[side] && !(natural && p.cav_gather)) v += cav_coef<C>(p, side, j);
                  f[qi] = v;
                });
              }
            });
          } else if (((bmask >> e.plane) & 1) && !wall) {
            if (e.kind == PK_LID) {
              lid_node<C, SHIFT>(f, e.U, p.sig);
            } else if (e.kind == PK_NEUMANN) {
              neumann_node<C, SHIFT>(f, e.plane, e.U);
            } else if (e.kind == PK_PRESSURE) {
              neumann_node<C, SHIFT, true>(f, e.plane, e.U, e.U[0]);
            } else {
              // outflow (no counterpart in the reference): the populations entering
              // from outside take their post-stream values at the first interior node
              // (one copy of the code for the six planes keeps the register count down)
              const int pl = e.plane;
              const int dim = pl < 2 ? 2 : (pl < 4 ? 1 : 0);
              const int inward = (pl & 1) ? -1 : 1;
              if (p.oflow) {
                // in-place AA step: the values were captured before any node was updated
                const C* cap =
                    (const C*)p.oflow + (size_t)e.U[0] * 5 * (size_t)p.oflow_n + plane_slot(pl, x, y, z, nx, ny);
                int j = 0;
                static_for<1, Q>([&](auto ic) {
                  constexpr int i = decltype(ic)::value;
                  const int cd = dim == 0 ? CX(i) : (dim == 1 ? CY(i) : CZ(i));
                  if (cd == inward) f[i] = cap[(size_t)(j++) * p.oflow_n];
                });
              } else {
                const int ax = x + (dim == 0 ? inward : 0), ay = y + (dim == 1 ? inward : 0),
                          az = z + (dim == 2 ? inward : 0);
                const uint32_t aidx = node_index(p, ax, ay, az);
                const bool awall = p.flags[aidx] != 0;
                static_for<1, Q>([&](auto ic) {
                  constexpr int i = decltype(ic)::value;
                  const int cd = dim == 0 ? CX(i) : (dim == 1 ? CY(i) : CZ(i));
                  if (cd == inward) f[i] = pull_one<S, MODE, i>(p, ax, ay, az, aidx, awall);
                });
              }
            }
          }
        }
      }
    }
}

// Phase 3: macro_variables (optionally written out) and collision.
// RAW (FP16 BGK in the A/B bulk paths only): the populations leave already
// multiplied by the storage scale, for store_bulk<..., RAW>; wall lanes are
// scaled as they are.
template <class S, int COLL, int MODE, bool RAW = false>
__device__ __forceinline__ void collide_node(const StepParams& p, uint32_t idx, bool wall,
                                             typename StoreOps<S>::C* f) {
  using O = StoreOps<S>;
  constexpr bool SHIFT = O::shifted;
  if constexpr (mode_collides(MODE)) {
    if constexpr (!SHIFT) {
      double rho, ux, uy, uz;
      macro_f64((const double*)f, rho, ux, uy, uz);
      if (p.write_macro) {
        if (!wall) {
          p.m0[idx] = rho;
          p.m1[idx] = ux;
          p.m1[(size_t)p.N + idx] = uy;
          p.m1[2 * (size_t)p.N + idx] = uz;
        } else {
          p.m1[idx] = 0.0;
          p.m1[(size_t)p.N + idx] = 0.0;
          p.m1[2 * (size_t)p.N + idx] = 0.0;
        }
      }
      if (!wall) {
        if constexpr (COLL == COLL_BGK)
          bgk_f64((double*)f, rho, ux, uy, uz, p);
        else
          mrt_collide<double>((double*)f, rho, p, false);
      }
    } else {
      if constexpr (RAW) {
        if (wall) {  // not collided: stored as they are, in the scaled form
#pragma unroll
          for (int i = 0; i < Q; i++) f[i] *= StoreOps<S>::SCALE;
        }
      }
      float dr, jx, jy, jz;
      moments_f32((const float*)f, dr, jx, jy, jz);
      const float rho = 1.0f + dr;
      const float inv = recip_f32(rho);
      const float ux = jx * inv, uy = jy * inv, uz = jz * inv;
      if (p.write_macro) {
        if (!wall) {
          p.m0[idx] = (double)rho;
          p.m1[idx] = (double)ux;
          p.m1[(size_t)p.N + idx] = (double)uy;
          p.m1[2 * (size_t)p.N + idx] = (double)uz;
        } else {
          p.m1[idx] = 0.0;
          p.m1[(size_t)p.N + idx] = 0.0;
          p.m1[2 * (size_t)p.N + idx] = 0.0;
        }
      }
      if (!wall) {
        if constexpr (COLL == COLL_BGK) {
          bgk_shift<RAW>((float*)f, dr, ux, uy, uz, p);
        } else {
          mrt_collide<float>((float*)f, rho, p, true);
        }
      }
    }
  }
}

// Phase 4: store (64-bit per-lane addressing; any node).
template <class S, int MODE>
__device__ __forceinline__ void store_node(const StepParams& p, int x, int y, int z, uint32_t idx, bool wall,
                                           typename StoreOps<S>::C* f) {
  using O = StoreOps<S>;
  using C = typename O::C;
  S* __restrict__ fout = (S*)p.fout;
  const size_t st = p.stride;
  const Rounder<S> R(p, StoreOps<S>::narrow16 ? canonical_index(p, x, y, z) : idx);
  if constexpr (MODE == MODE_AA_EVEN) {
#pragma unroll
    for (int i = 0; i < Q; i++) O::st_sr(fout, (size_t)OPP(i) * st + idx, f[i], R.word(i));
  } else if constexpr (MODE == MODE_AA_ODD) {
    // push f*_i(x) to the natural slot of x + c_i; a non-periodic face or (with
    // wall_bounce_back) a wall downstream bounces it into slot opp(i) of x,
    // adding the moving-wall term the gather would add on a z face
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int xm = x > 0 ? x - 1 : (p.per[0] ? nx - 1 : -1);
    const int xp = x < nx - 1 ? x + 1 : (p.per[0] ? 0 : -1);
    const int ym = y > 0 ? y - 1 : (p.per[1] ? ny - 1 : -1);
    const int yp = y < ny - 1 ? y + 1 : (p.per[1] ? 0 : -1);
    const int zm = z > 0 ? z - 1 : (p.per[2] ? nz - 1 : -1);
    const int zp = z < nz - 1 ? z + 1 : (p.per[2] ? 0 : -1);
    const bool wbb = p.wall_bb && !wall;
    static_for([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int tx = CX(i) == 0 ? x : (CX(i) == 1 ? xp : xm);
      const int ty = CY(i) == 0 ? y : (CY(i) == 1 ? yp : ym);
      const int tz = CZ(i) == 0 ? z : (CZ(i) == 1 ? zp : zm);
      if (tx < 0 || ty < 0 || tz < 0) {
        C v = f[i];
        if constexpr (CZ(i) != 0) {
          constexpr int side = CZ(i) == -1 ? 0 : 1;  // face of the incoming direction opp(i)
          if (tz < 0 && p.cav_gather && p.cav_on[side]) v += cav_coef<C>(p, side, cavity_slot(OPP(i)));
        }
        O::st_sr(fout, (size_t)OPP(i) * st + idx, v, R.word(i));
      } else {
        const uint32_t tidx = node_index(p, tx, ty, tz);
        if (i != 0 && wbb && p.flags[tidx] != 0)
          O::st_sr(fout, (size_t)OPP(i) * st + idx, f[i], R.word(i));
        else
          O::st_sr(fout, (size_t)i * st + tidx, f[i], R.word(i));
      }
    });
  } else {
#pragma unroll
    for (int i = 0; i < Q; i++) O::st_sr(fout, (size_t)i * st + idx, f[i], R.word(i));
  }
}

// start_iteration set_distribution (set_distribution.hpp:38-41): before the
// macro variables of a step, f_i = value * w_i on the nodes of each box, in
// list order; the shifted storage holds f_i - w_i.
template <class S, int MODE>
__device__ __forceinline__ void apply_sets(const StepParams& p, int x, int y, int z, typename StoreOps<S>::C* f) {
  using C = typename StoreOps<S>::C;
  if constexpr (mode_collides(MODE)) {
    for (int b = 0; b < p.nset; b++) {
      const int* B = p.set_box[b];
      const bool in = x >= B[0] && x <= B[3] && y >= B[1] && y <= B[4] && z >= B[2] && z <= B[5];
      const double v = p.set_val[b];
#pragma unroll
      for (int i = 0; i < Q; i++) {
        const C nv = StoreOps<S>::shifted ? C(v * WD(i) - WD(i)) - C(p.sig[dir_class(i)]) : C(v * WD(i));
        f[i] = in ? nv : f[i];
      }
    }
  }
}

template <class S, int COLL, int MODE>
__device__ __forceinline__ void finish_node(const StepParams& p, int x, int y, int z, uint32_t idx, bool wall,
                                            typename StoreOps<S>::C* f) {
  if (p.nset) apply_sets<S, MODE>(p, x, y, z, f);
  collide_node<S, COLL, MODE>(p, idx, wall, f);
  store_node<S, MODE>(p, x, y, z, idx, wall, f);
}

// ---------------------------------------------------------------------------
// bulk-node addressing
// ---------------------------------------------------------------------------
// In the bulk kernel every population access is a wave-uniform 64-bit plane
// base (scalar registers) plus a per-lane 32-bit byte offset: one vector add
// per direction instead of a 64-bit address per lane.  Valid while every byte
// offset of a plane, (N-1)*sizeof(S), fits in 32 bits (the host picks the
// 64-bit variant of the bulk kernel otherwise).
template <class T>
__device__ __forceinline__ T* at_bytes(T* base, uint32_t off) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return (T*)((B*)base + off);
}
// Plane bases are walked with one running pointer held in scalar registers and
// opaque to the optimizer: otherwise it reassociates base + i*stride + offset
// into a per-lane 64-bit address and keeps 19 products i*stride live.
template <class T>
__device__ __forceinline__ T* scalar_ptr(T* p) {
  auto g = (__attribute__((address_space(1))) T*)p;  // global memory: keeps global_* (not flat_*) accesses
  asm("" : "+s"(g));
  return (T*)g;
}
// A pointer the compiler cannot prove wave-uniform (loop-carried through
// divergent control flow) but that is: read from the first active lane.
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (T*)(((uint64_t)hi << 32) | lo);
}

// A per-lane byte offset defined anew where it is used, so the selector sees
// (scalar base + zero-extended 32-bit offset) in the same block and emits the
// scalar-base form of the access.
__device__ __forceinline__ uint32_t lane_offset(uint32_t o) {
  asm("" : "+v"(o));
  return o;
}

// Byte offsets from a node to its two neighbours on each axis, with the
// periodic wrap applied (arithmetic modulo 2^32).  A neighbour outside a
// non-periodic face gets offset 0 (the node itself: a valid address whose value
// the bounce-back replaces) and sets its face bit.
struct NodeNbrs {
  uint32_t dm[3];  // to the node at -1 on the axis
  uint32_t dp[3];  // to the node at +1 on the axis
  bool fm[3];      // -1 neighbour is outside a non-periodic face
  bool fp[3];      // +1 neighbour is outside a non-periodic face
};

template <class S>
__device__ __forceinline__ NodeNbrs node_nbrs(const StepParams& p, int x, int y, int z) {
  NodeNbrs a;
  const int c[3] = {x, y, z};
  const int n[3] = {p.nx, p.ny, p.nz};
  const uint32_t unit[3] = {(uint32_t)sizeof(S) * p.lay.st[0], (uint32_t)sizeof(S) * p.lay.st[1],
                            (uint32_t)sizeof(S) * p.lay.st[2]};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const bool lo = c[d] == 0, hi = c[d] == n[d] - 1;
    const uint32_t span = (uint32_t)(n[d] - 1) * unit[d];
    a.dm[d] = lo ? (p.per[d] ? span : 0u) : 0u - unit[d];
    a.dp[d] = hi ? (p.per[d] ? 0u - span : 0u) : unit[d];
    a.fm[d] = lo && !p.per[d];
    a.fp[d] = hi && !p.per[d];
  }
  return a;
}

// offset of x - c_i (UP) or x + c_i (!UP) relative to x
template <int I, bool UP>
__device__ __forceinline__ uint32_t nbr_offset(const NodeNbrs& a) {
  uint32_t o = 0;
  if constexpr (CX(I) != 0) o += (CX(I) == 1) == UP ? a.dm[0] : a.dp[0];
  if constexpr (CY(I) != 0) o += (CY(I) == 1) == UP ? a.dm[1] : a.dp[1];
  if constexpr (CZ(I) != 0) o += (CZ(I) == 1) == UP ? a.dm[2] : a.dp[2];
  return o;
}
// x - c_i (UP) or x + c_i (!UP) lies outside a non-periodic face; zface: outside a z face
template <int I, bool UP>
__device__ __forceinline__ bool nbr_outside(const NodeNbrs& a, bool& zface) {
  bool o = false;
  zface = false;
  if constexpr (CX(I) != 0) o = o || ((CX(I) == 1) == UP ? a.fm[0] : a.fp[0]);
  if constexpr (CY(I) != 0) o = o || ((CY(I) == 1) == UP ? a.fm[1] : a.fp[1]);
  if constexpr (CZ(I) != 0) {
    zface = (CZ(I) == 1) == UP ? a.fm[2] : a.fp[2];
    o = o || zface;
  }
  return o;
}

// Gather without branches (same result as gather_node), for waves that hold
// bounce-back lanes: the opposite directions (2k-1, 2k) share one scalar base,
// their first plane, so a lane that bounces back (face or wall upstream) loads
// f*_opp(i)(x) in the same instruction as the other lanes' pull, by selecting
// the other offset.  Needs (stride + N) * sizeof(S) < 2^32 (p.pair32).
template <class S, int MODE>
__device__ __forceinline__ void gather_select(const StepParams& p, uint32_t bo, const NodeNbrs& a, uint32_t wbits,
                                              typename StoreOps<S>::C* f) {
  using O = StoreOps<S>;
  using C = typename O::C;
  constexpr bool AA = mode_aa(MODE);
  const S* __restrict__ fin = (const S*)p.fin;
  const uint32_t stb = (uint32_t)(p.stride * sizeof(S));
  bo = lane_offset(bo);
  f[0] = O::ld(at_bytes(fin, bo), 0);
  const S* b = fin;
  static_for<1, Q>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    constexpr int lo = (i & 1) ? i : i - 1;  // first plane of the pair {i, opp(i)}
    if constexpr (i & 1) b = scalar_ptr(b + (size_t)(i == 1 ? 1 : 2) * p.stride);
    bool zf;
    const bool bb = nbr_outside<i, true>(a, zf) || ((wbits >> i) & 1u);
    // A/B: the pull reads plane i, the bounce-back plane opp(i); AA (swapped): the reverse
    const uint32_t pull = ((AA ? OPP(i) : i) - lo) * stb + bo + nbr_offset<i, true>(a);
    const uint32_t back = ((AA ? i : OPP(i)) - lo) * stb + bo;
    C v = O::ld(at_bytes(b, bb ? back : pull), 0);
    if constexpr (CZ(i) != 0) {
      constexpr int side = CZ(i) == 1 ? 0 : 1;
      if (p.cav_gather && p.cav_on[side]) v += (zf ? cav_coef<C>(p, side, cavity_slot(i)) : C(0));
    }
    f[i] = v;
  });
}

// Gather for a bulk node (same result as gather_node): all 19 pulls are issued
// with wrapped offsets, then, only in waves holding a bounce-back lane, the
// lanes whose upstream node is outside a non-periodic face or (wall_bb) a wall
// take f*_opp(i)(x) instead, plus the moving-wall term on a cavity face.
template <class S, int MODE>
__device__ __forceinline__ void gather_bulk(const StepParams& p, uint32_t bo, const NodeNbrs& a, uint32_t wbits,
                                            typename StoreOps<S>::C* f) {
  using O = StoreOps<S>;
  using C = typename O::C;
  const S* __restrict__ fin = (const S*)p.fin;
  const size_t st = p.stride;
  const S* b = fin;
  bo = lane_offset(bo);
  if constexpr (mode_reads_own(MODE)) {
    static_for([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      b = scalar_ptr(j ? b + st : b);
      f[j] = O::ld(at_bytes(b, bo), 0);
    });
  } else {
    constexpr bool AA = mode_aa(MODE);
    const bool edge = a.fm[0] || a.fp[0] || a.fm[1] || a.fp[1] || a.fm[2] || a.fp[2] || wbits > 1u;
    if (p.pair32 && __any(edge)) {
      gather_select<S, MODE>(p, bo, a, wbits, f);
      return;
    }
    // plane j holds direction i = j (A/B) or i = opp(j) (AA, swapped)
    static_for([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      constexpr int i = AA ? OPP(j) : j;
      b = scalar_ptr(j ? b + st : b);
      f[i] = O::ld(at_bytes(b, bo + nbr_offset<i, true>(a)), 0);
    });
    if (__any(edge)) {
      static_for<1, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        bool zf;
        const bool out = nbr_outside<i, true>(a, zf);
        if (out || ((wbits >> i) & 1u)) {
          C v = O::ld(at_bytes(fin + (size_t)(AA ? i : OPP(i)) * st, bo), 0);
          if constexpr (CZ(i) != 0) {
            constexpr int side = CZ(i) == 1 ? 0 : 1;
            if (zf && p.cav_gather && p.cav_on[side]) v += cav_coef<C>(p, side, cavity_slot(i));
          }
          f[i] = v;
        }
      });
    }
  }
}

// Store for a bulk node (same result as store_node).
// The A/B bulk paths of FP16 BGK fold the storage scale into the collision
// (collide_node<..., RAW>) and store without a multiply (store_bulk<..., RAW>).
template <class S, int COLL, int MODE>
constexpr bool raw_store() {
  return std::is_same<S, _Float16>::value && COLL == COLL_BGK && mode_collides(MODE) && !mode_aa(MODE);
}

template <class S, int MODE, bool RAW = false>
__device__ __forceinline__ void store_bulk(const StepParams& p, uint32_t bo, uint32_t cidx, const NodeNbrs& a,
                                           uint32_t wbits, typename StoreOps<S>::C* f) {
  using O = StoreOps<S>;
  using C = typename O::C;
  S* __restrict__ fout = (S*)p.fout;
  const size_t st = p.stride;
  S* b = fout;
  const Rounder<S> R(p, cidx);
  bo = lane_offset(bo);
  if constexpr (MODE == MODE_AA_EVEN) {
    static_for([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      b = scalar_ptr(j ? b + st : b);
      O::st_sr(at_bytes(b, bo), 0, f[OPP(j)], R.word(OPP(j)));
    });
  } else if constexpr (MODE == MODE_AA_ODD) {
    // push to x + c_i; a face or (wall_bb) wall downstream bounces into slot opp(i) of x
    const bool edge = a.fm[0] || a.fp[0] || a.fm[1] || a.fp[1] || a.fm[2] || a.fp[2] || wbits > 1u;
    if (!__any(edge)) {
      static_for([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        b = scalar_ptr(i ? b + st : b);
        O::st_sr(at_bytes(b, bo + nbr_offset<i, false>(a)), 0, f[i], R.word(i));
      });
    } else {
      static_for([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        bool zf;
        const bool out = nbr_outside<i, false>(a, zf);
        if (i != 0 && (out || ((wbits >> OPP(i)) & 1u))) {
          C v = f[i];
          if constexpr (CZ(i) != 0) {
            constexpr int side = CZ(i) == -1 ? 0 : 1;  // face of the incoming direction opp(i)
            if (zf && p.cav_gather && p.cav_on[side]) v += cav_coef<C>(p, side, cavity_slot(OPP(i)));
          }
          O::st_sr(at_bytes(fout + (size_t)OPP(i) * st, bo), 0, v, R.word(i));
        } else {
          O::st_sr(at_bytes(fout + (size_t)i * st, bo + nbr_offset<i, false>(a)), 0, f[i], R.word(i));
        }
      });
    }
  } else {
    static_for([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      b = scalar_ptr(j ? b + st : b);
      if constexpr (RAW)
        O::st_sr_scaled(at_bytes(b, bo), 0, f[j], R.word(j));
      else
        O::st_sr(at_bytes(b, bo), 0, f[j], R.word(j));
    });
  }
}

// Full per-node pipeline.  PROG selects whether the post-stream boundary
// program is compiled in: the bulk kernel never runs it (its nodes are not on
// a program plane), the boundary kernel always may.
template <class S, int COLL, int MODE, bool PROG>
__device__ __forceinline__ void process_node(const StepParams& p, int x, int y, int z) {
  using C = typename StoreOps<S>::C;
  const uint32_t idx = node_index(p, x, y, z);
  const bool wall = p.flags[idx] != 0;
  if (mode_skips_walls(MODE) && wall && p.wall_bb) return;
  if (p.nspec && (p.bbm[idx] >> 31)) return;  // special_kernel's node
  C f[Q];
  gather_node<S, MODE>(p, x, y, z, idx, wall, f);
  if constexpr (PROG && MODE != MODE_COLLIDE) {
    if (!p.skip_program) apply_program<S, MODE>(p, x, y, z, idx, wall, f);
  }
  finish_node<S, COLL, MODE>(p, x, y, z, idx, wall, f);
}

// Bulk kernel: every node that is not on a plane of the boundary program.
// A tile is blockDim.x nodes along x by blockDim.y*NPT rows along y; each
// thread gathers its NPT nodes first (19*NPT loads in flight), then collides
// and stores them.
// PERM: the node layout is not the default one (x fastest); the default layout
// compiles the index arithmetic down to fixed axes.
template <class S, int COLL, int MODE, int NPT, bool WIDE, bool PERM>
__global__ void __launch_bounds__(256) step_kernel(const StepParams p) {
  using C = typename StoreOps<S>::C;
  // 1-D dispatch over tiles.  Workgroups are handed to the 8 XCDs round-robin;
  // with xcd_remap each XCD walks a contiguous run of tiles, so tiles that
  // share cache lines (x neighbours, y/z neighbours' flags) share one L2.
  const uint32_t nt = (uint32_t)p.gx * (uint32_t)p.gy * (uint32_t)p.gz;
  uint32_t t = blockIdx.x;
  if (p.xcd_remap) {
    const uint32_t per = nt / 8u;
    if (t < per * 8u) t = (t % 8u) * per + t / 8u;
  }
  // lanes run along the layout's fastest axis, tile rows along the next one and
  // the grid along the slowest: l, row and c2 below
  const int bx = (int)(t % (uint32_t)p.gx);
  const uint32_t r = t / (uint32_t)p.gx;
  const int by = (int)(r % (uint32_t)p.gy);
  const int c2 = (int)(r / (uint32_t)p.gy);
  const int l = bx * blockDim.x + threadIdx.x;
  if (l >= p.ln[0] || c2 >= p.ln[2]) return;
  const int o0 = PERM ? p.lay.ord[0] : 0, o1 = PERM ? p.lay.ord[1] : 1;
  auto coords = [&](int row, int& x, int& y, int& z) {
    if constexpr (PERM) {
      x = o0 == 0 ? l : (o1 == 0 ? row : c2);
      y = o0 == 1 ? l : (o1 == 1 ? row : c2);
      z = o0 == 2 ? l : (o1 == 2 ? row : c2);
    } else {
      x = l;
      y = row;
      z = c2;
    }
  };
  // the default layout: the node index and the stochastic-rounding key coincide
  auto index_of = [&](int x, int y, int z) -> uint32_t {
    if constexpr (PERM)
      return node_index(p, x, y, z);
    else
      return ((uint32_t)z * (uint32_t)p.ny + (uint32_t)y) * (uint32_t)p.nx + (uint32_t)x;
  };
  if constexpr (!WIDE) {
    C f[NPT][Q];
    uint32_t bo[NPT], wbits[NPT], cidx[NPT];
    NodeNbrs a[NPT];
    bool wall[NPT], act[NPT];
#pragma unroll
    for (int k = 0; k < NPT; k++) {
      const int row = (by * NPT + k) * blockDim.y + threadIdx.y;
      int x, y, z;
      coords(row, x, y, z);
      act[k] = row < p.ln[1] && !(MODE != MODE_COLLIDE && (node_planes(x, y, z, p.nx, p.ny, p.nz) & p.prog_planes));
      const uint32_t idx = index_of(x, y, z);
      bo[k] = idx * (uint32_t)sizeof(S);
      cidx[k] = PERM && StoreOps<S>::narrow16 ? canonical_index(p, x, y, z) : idx;
      wbits[k] = 0u;
      wall[k] = false;
      if (act[k] && p.any_wall) {
        if (p.wall_bb && p.bbm) {  // the host never launches this variant with wall_bb and no mask
          wbits[k] = p.bbm[idx];
          wall[k] = (wbits[k] & 1u) != 0u;
          if (wbits[k] >> 31) act[k] = false;  // special_kernel's node
        } else {
          wall[k] = p.flags[idx] != 0;
        }
        if (mode_skips_walls(MODE) && wall[k] && p.wall_bb) act[k] = false;
      }
      if (act[k]) {
        a[k] = node_nbrs<S>(p, x, y, z);
        gather_bulk<S, MODE>(p, bo[k], a[k], wbits[k], f[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < NPT; k++) {
      if (act[k]) {
        constexpr bool RAW = raw_store<S, COLL, MODE>();
        collide_node<S, COLL, MODE, RAW>(p, bo[k] / (uint32_t)sizeof(S), wall[k], f[k]);
        store_bulk<S, MODE, RAW>(p, bo[k], cidx[k], a[k], wbits[k], f[k]);
      }
    }
  } else {
  C f[NPT][Q];
  uint32_t idx[NPT];
  int xs[NPT], ys[NPT], zs[NPT];
  bool wall[NPT], act[NPT];
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int row = (by * NPT + k) * blockDim.y + threadIdx.y;
    int& x = xs[k];
    int& y = ys[k];
    int& z = zs[k];
    coords(row, x, y, z);
    act[k] = row < p.ln[1] && !(MODE != MODE_COLLIDE && (node_planes(x, y, z, p.nx, p.ny, p.nz) & p.prog_planes));
    idx[k] = index_of(x, y, z);
    if (act[k]) {
      wall[k] = p.flags[idx[k]] != 0;
      if (mode_skips_walls(MODE) && wall[k] && p.wall_bb) act[k] = false;
      if (p.nspec && (p.bbm[idx[k]] >> 31)) act[k] = false;  // special_kernel's node
    }
    if (act[k]) gather_node<S, MODE>(p, x, y, z, idx[k], wall[k], f[k]);
  }
#pragma unroll
  for (int k = 0; k < NPT; k++)
    if (act[k]) finish_node<S, COLL, MODE>(p, xs[k], ys[k], zs[k], idx[k], wall[k], f[k]);
  }
}

// Boundary kernel: the nodes of the program planes, each once (a node on
// several planes is handled by the lowest-numbered plane that holds it).
template <class S, int COLL, int MODE>
__global__ void __launch_bounds__(256) boundary_kernel(const StepParams p) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const int n[3] = {nx, ny, nz};
  for (int pl = 0; pl < 6; pl++) {
    if (!((p.prog_planes >> pl) & 1)) continue;
    const int dim = plane_dim(pl);
    // lanes along the in-plane axis that is faster in memory
    int da = dim == 0 ? 1 : 0, db = dim == 2 ? 1 : 2;
    if (p.lay.st[db] < p.lay.st[da]) {
      const int sw = da;
      da = db;
      db = sw;
    }
    const uint32_t na = (uint32_t)n[da], nb = (uint32_t)n[db];
    const uint32_t cnt = na * nb;
    if (t < cnt) {
      const int a = (int)(t % na), b = (int)(t / na);
      const int c = plane_side(pl) ? n[dim] - 1 : 0;
      const int x = dim == 0 ? c : (da == 0 ? a : b);
      const int y = dim == 1 ? c : (da == 1 ? a : b);
      const int z = dim == 2 ? c : (da == 2 ? a : b);
      const int lower = node_planes(x, y, z, nx, ny, nz) & p.prog_planes & ((1 << pl) - 1);
      if (!lower) process_node<S, COLL, MODE, true>(p, x, y, z);
      return;
    }
    t -= cnt;
  }
}

// Nodes next to a wall across a periodic face (wall_bounce_back; the list and
// its masks come from k_find_special).  hippoLBM bounces such a wall off its
// ghost neighbour, which still holds the populations of the previous step's
// streaming (ghosts are refreshed only inside streaming, streaming.cu:107-111;
// bounce_back.hpp:163-185): the population f_i(x) arriving from the wall is
// f_opp(i)(x) of x's previous pre-collision state, not of the current
// post-collision one.  Each listed node keeps those values in p.spec_val,
// written whenever a kernel forms the node's pre-collision state (collision
// modes, after the boundary program: the ghost gets the same plane operators
// from the same inputs) and read by the next gather.  None of these nodes is a
// wall; the host refuses the operators that treat a node and its ghost
// differently (pre/post_bounce_back faces, programs on planes of periodic axes,
// start_iteration boxes, outflow reading the node).
template <class S, int COLL, int MODE>
__global__ void __launch_bounds__(64) special_kernel(const StepParams p) {
  using C = typename StoreOps<S>::C;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= p.nspec) return;
  const uint32_t idx = p.spec_idx[k];
  const uint32_t mask = p.spec_mask[k];
  int x, y, z;
  node_coords(p, idx, x, y, z);
  C* sv = (C*)p.spec_val;
  const size_t ns = p.nspec;
  C f[Q];
  gather_node<S, MODE>(p, x, y, z, idx, false, f);
  if constexpr (!mode_reads_own(MODE)) {
    static_for<1, Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if ((mask >> i) & 1u) f[i] = sv[(size_t)i * ns + k];
    });
  }
  if constexpr (MODE != MODE_COLLIDE) {
    if (!p.skip_program && (node_planes(x, y, z, p.nx, p.ny, p.nz) & p.prog_planes))
      apply_program<S, MODE>(p, x, y, z, idx, false, f);
  }
  if constexpr (mode_collides(MODE)) {
    static_for<1, Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if ((mask >> i) & 1u) sv[(size_t)i * ns + k] = f[OPP(i)];
    });
  }
  collide_node<S, COLL, MODE>(p, idx, false, f);
  if constexpr (MODE == MODE_AA_ODD) {
    // direction opp(i) runs into the wall and bounces into slot i of the node,
    // which the next (even) step reads as f_i: the pre-collision f_opp(i)
    static_for<1, Q>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if ((mask >> i) & 1u) f[OPP(i)] = sv[(size_t)i * ns + k];
    });
  }
  store_node<S, MODE>(p, x, y, z, idx, false, f);
}

// Outflow capture for an in-place AA step: before any node of the step is
// updated, each outflow face node's incoming populations are evaluated at its
// first interior node (as the outflow entry would) and stored in p.oflow, slot
// ((ordinal * 5 + j) * oflow_n + plane_slot).  Outflow entries carry their
// ordinal among the outflow entries in U[0].
template <class S, int MODE>
__global__ void __launch_bounds__(256) outflow_capture_kernel(const StepParams p) {
  using C = typename StoreOps<S>::C;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  for (int k = 0; k < p.nprog; k++) {
    const ProgEntry& e = p.prog[k];
    if (e.kind != PK_OUTFLOW) continue;
    static_for<0, 6>([&](auto pc) {
      constexpr int pl = decltype(pc)::value;
      constexpr int dim = pl < 2 ? 2 : (pl < 4 ? 1 : 0);
      constexpr int inward = (pl & 1) ? -1 : 1;
      if (e.plane != pl) return;
      const uint32_t na = dim == 0 ? (uint32_t)ny : (uint32_t)nx;
      const uint32_t nb = dim == 2 ? (uint32_t)ny : (uint32_t)nz;
      if (t >= na * nb) return;
      const int a = (int)(t % na), b = (int)(t / na);
      const int c = (pl & 1) ? (dim == 0 ? nx - 1 : (dim == 1 ? ny - 1 : nz - 1)) : 0;
      const int x = dim == 0 ? c : a;
      const int y = dim == 0 ? a : (dim == 1 ? c : b);
      const int z = dim == 2 ? c : b;
      const int ax = x + (dim == 0 ? inward : 0), ay = y + (dim == 1 ? inward : 0), az = z + (dim == 2 ? inward : 0);
      const uint32_t aidx = node_index(p, ax, ay, az);
      const bool awall = p.flags[aidx] != 0;
      C* cap = (C*)p.oflow + (size_t)e.U[0] * 5 * (size_t)p.oflow_n + t;
      int j = 0;
      static_for<1, Q>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr int cd = dim == 0 ? CX(i) : (dim == 1 ? CY(i) : CZ(i));
        if constexpr (cd == inward) cap[(size_t)(j++) * p.oflow_n] = pull_one<S, MODE, i>(p, ax, ay, az, aidx, awall);
      });
    });
  }
}

}  // namespace hlbm

namespace hlbm {

// ---------------------------------------------------------------------------
// Batched domains: many independent subdomains of uneven sizes advance in one
// persistent launch per step.  Each domain owns a slice of two population
// arenas (A/B) and of a flag arena; a tile table lists (domain, first node)
// pairs of 256 consecutive nodes.  Per-domain parameters replace the uniform
// kernel arguments of the single-domain path; the per-node pipeline is the
// same code (gather_node / finish_node), so a domain's populations are the
// same alone and in any mix.
// ---------------------------------------------------------------------------
struct BatchDomain {
  int32_t nx, ny, nz;
  int32_t per[3];
  int32_t coll;          // COLL_BGK / COLL_MRT
  int32_t wall_bb;
  int32_t cav_on[2];
  int32_t active_steps;  // steps this domain takes in the current chunk
  int32_t parity;        // arena holding f* at the start of the chunk
  int32_t mrt_custom;
  uint32_t t0;           // the domain's time steps done when the chunk starts (rounding seed)
  uint64_t f_off;        // element offset of population plane 0 in an arena
  uint64_t stride;       // elements between population planes
  uint64_t flag_off;     // byte offset in the flag arena
  double inv_nx;         // 1/nx and 1/(nx*ny): node coordinates without integer division
  double inv_nxy;
  int32_t has_force;
  int32_t any_wall;      // the solve has obstacles (else no flag / mask is read)
  float k1;              // FP32 BGK constants (StepParams::k1, kw)
  float kw[3];
  float f3[3];           // 3 F (FP32-compute forms)
  float sig[3];          // shifted storage: the solve's rest state (RestShift)
  double tau;
  double inv_tau;
  double ef[Q];
  double cav[2][5];
  float cavf[2][5];
  double mrt[10];
  float mrtf[10];        // StepParams::mrtf
};

// A tile is BATCH_TILE_NODES consecutive nodes of one domain, walked by a
// 256-thread workgroup in passes of 256 nodes: the domain's parameters load
// once per tile and serve every pass.
constexpr uint32_t BATCH_TILE_NODES = 1024;
constexpr uint32_t BATCH_PASS_NODES = 256;

// Work units (one workgroup's share between parameter loads) per tile of the
// batch kernel for a storage element size and collision: the 16-bit BGK kernel
// walks a whole tile per unit, the others a 256-node pass.
HLBM_HD constexpr uint32_t batch_units_per_tile(int esize, int coll) {
  return (esize == 2 && coll == COLL_BGK) ? 1u : BATCH_TILE_NODES / BATCH_PASS_NODES;
}

struct BatchTile {
  uint32_t dom;
  uint32_t first;  // first node (linear index) of the tile
};

struct BatchParams {
  const BatchDomain* doms;
  const BatchTile* tiles;
  const uint32_t* ntiles;  // device-resident counts: [BGK tiles, all tiles]; BGK tiles come first
  void* arena[2];
  const uint8_t* flags;
  const uint32_t* bbm;     // per-node bounce-back masks, at the same offsets as the flags
  int32_t step;            // step index inside the chunk
  int32_t xcd_split;       // each XCD walks its own contiguous run of tiles
  int32_t sr;              // 16-bit storage rounds stochastically
};

template <class S>
__device__ __forceinline__ StepParams batch_step_params(const BatchParams& bp, const BatchDomain& d, int parity) {
  StepParams p;
  p.nx = d.nx;
  p.ny = d.ny;
  p.nz = d.nz;
  p.per[0] = d.per[0];
  p.per[1] = d.per[1];
  p.per[2] = d.per[2];
  p.N = (uint32_t)d.nx * (uint32_t)d.ny * (uint32_t)d.nz;
  p.stride = d.stride;
  // batch solves keep the default layout, x fastest
  p.lay.st[0] = 1u;
  p.lay.st[1] = (uint32_t)d.nx;
  p.lay.st[2] = (uint32_t)d.nx * (uint32_t)d.ny;
  p.lay.ord[0] = 0;
  p.lay.ord[1] = 1;
  p.lay.ord[2] = 2;
  p.ln[0] = d.nx;
  p.ln[1] = d.ny;
  p.ln[2] = d.nz;
  // both arena bases are in registers: a select, not an indexed load
  const S* a0 = (const S*)bp.arena[0];
  const S* a1 = (const S*)bp.arena[1];
  p.fin = (const void*)((parity ? a1 : a0) + d.f_off);
  p.fout = (void*)(const_cast<S*>(parity ? a0 : a1) + d.f_off);
  p.flags = bp.flags + d.flag_off;
  p.m0 = nullptr;
  p.m1 = nullptr;
  p.nset = 0;
  p.write_macro = 0;
  p.wall_bb = d.wall_bb;
  p.plane_bb = 1;
  p.cav_on[0] = d.cav_on[0];
  p.cav_on[1] = d.cav_on[1];
  p.cav_gather = 1;
  p.nprog = 0;
  p.prog_planes = 0;
  p.oflow = nullptr;
  p.oflow_n = 0;
  p.pair32 = 1;  // checked when the solve was submitted
  p.force_wide = 0;
  p.sr = bp.sr;
  p.seed = mix32((d.t0 + (uint32_t)bp.step) * 0x9E3779B9u + 1u);
  p.any_wall = 1;
  p.bbm = bp.bbm + d.flag_off;
  p.mrt_custom = d.mrt_custom;
  for (int k = 0; k < 5; k++) {
    p.cav[0][k] = d.cav[0][k];
    p.cav[1][k] = d.cav[1][k];
    p.cavf[0][k] = d.cavf[0][k];
    p.cavf[1][k] = d.cavf[1][k];
  }
  p.tau = d.tau;
  p.inv_tau = d.inv_tau;
  p.has_force = d.has_force;
  for (int k = 0; k < 3; k++) p.sig[k] = d.sig[k];
  for (int k = 0; k < 3; k++) p.f3[k] = d.f3[k];
  p.k1 = d.k1;
  for (int k = 0; k < 3; k++) p.kw[k] = d.kw[k];
  for (int i = 0; i < Q; i++) p.ef[i] = d.ef[i];
  for (int k = 0; k < 10; k++) p.mrt[k] = d.mrt[k];
  for (int k = 0; k < 10; k++) p.mrtf[k] = d.mrtf[k];
  return p;
}

// One node of a batch tile.  WALLS: the solve has obstacles, so the node's
// bounce-back mask is read; a solve without obstacles (closed cavities, empty
// channels) compiles the mask out, and only its faces bounce back.
template <class S, int COLL, bool WALLS>
__device__ __forceinline__ void batch_node(const BatchParams& bp, const BatchDomain& d, const StepParams& p,
                                           uint32_t n) {
  using C = typename StoreOps<S>::C;
  // floor((n + 1/2) / d) from the rounded reciprocal is exact: the quotient's
  // distance to an integer, >= 1/(2d), exceeds the relative rounding error 2^-52
  const uint32_t nxy = (uint32_t)p.nx * (uint32_t)p.ny;
  const uint32_t zu = (uint32_t)(((double)n + 0.5) * d.inv_nxy);
  const uint32_t rem = n - zu * nxy;
  const uint32_t yu = (uint32_t)(((double)rem + 0.5) * d.inv_nx);
  const int z = (int)zu, y = (int)yu, x = (int)(rem - yu * (uint32_t)p.nx);
  uint32_t m = 0u;
  if constexpr (WALLS) m = d.any_wall ? bp.bbm[d.flag_off + n] : 0u;
  const bool wall = WALLS && (m & 1u) != 0u;
  const uint32_t wbits = WALLS ? (d.wall_bb ? m : (m & 1u)) : 0u;
  const uint32_t bo = n * (uint32_t)sizeof(S);
  const NodeNbrs a = node_nbrs<S>(p, x, y, z);
  C f[Q];
  gather_select<S, MODE_FUSED>(p, bo, a, wbits, f);
  // (solves with obstacles keep the scale in the store: scaling their wall
  // lanes in the collision would push the kernel past 64 vector registers)
  constexpr bool RAW = !WALLS && raw_store<S, COLL, MODE_FUSED>();
  collide_node<S, COLL, MODE_FUSED, RAW>(p, n, wall, f);
  store_bulk<S, MODE_FUSED, RAW>(p, bo, n, a, wbits, f);  // batch solves: default layout
}

// One launch per collision model and time step: the BGK launch walks the
// BGK tiles, the MRT launch the MRT tiles, so each kernel keeps only its own
// collision in registers.  Tiles are BATCH_TILE_NODES consecutive nodes of one
// domain, so the domain's parameters are workgroup-uniform.
//
// The 16-bit BGK kernel loads a tile's parameters once and walks its
// BATCH_TILE_NODES / 256 passes in a loop; it also compiles a second node path
// for solves without obstacles (no mask read) and asks for 8 waves per SIMD,
// which the register allocator meets without spilling vector registers.  The
// wider kernels and the MRT kernels take one 256-node pass per parameter load
// (units of a quarter tile) and one node path: keeping the parameters live
// across passes, or a second path, costs them waves per SIMD.  The host sizes
// each launch to about one workgroup per unit; a workgroup that finds more
// walks them grid-stride (and the BGK kernels then request the next unit's
// tile entry while the current one's parameters load).
template <class S, int COLL>
__global__ void __launch_bounds__(256, (sizeof(S) == 2 && COLL == COLL_BGK) ? 8 : 1)
    batch_kernel(const BatchParams bp) {
  constexpr bool PREFETCH = COLL == COLL_BGK;
  constexpr uint32_t UNITS = batch_units_per_tile((int)sizeof(S), COLL);     // work units per tile
  constexpr uint32_t INNER = BATCH_TILE_NODES / BATCH_PASS_NODES / UNITS;  // passes per parameter load
  uint32_t t0 = (COLL == COLL_BGK ? 0u : bp.ntiles[0]) * UNITS;
  uint32_t t1 = (COLL == COLL_BGK ? bp.ntiles[0] : bp.ntiles[1]) * UNITS;
  uint32_t first = blockIdx.x, stride = gridDim.x;
  if (bp.xcd_split && (gridDim.x & 7u) == 0u) {
    // workgroups go to the 8 XCDs round-robin: XCD k (= blockIdx % 8) takes the
    // k-th eighth of the units, so neighbouring tiles share that XCD's L2
    const uint32_t k = blockIdx.x & 7u, n = t1 - t0, per = (n + 7u) / 8u;
    const uint32_t a = t0 + min(n, k * per), e = t0 + min(n, (k + 1u) * per);
    t0 = a;
    t1 = e;
    first = blockIdx.x >> 3;
    stride = gridDim.x >> 3;
  }
  uint32_t u = t0 + first;
  BatchTile next = BatchTile{0u, 0u};
  if (PREFETCH && u < t1) next = bp.tiles[u / UNITS];
  for (; u < t1; u += stride) {
    BatchTile tl;
    if constexpr (PREFETCH) {
      tl = next;
      if (u + stride < t1) next = bp.tiles[(u + stride) / UNITS];
    } else {
      tl = bp.tiles[u / UNITS];
    }
    const BatchDomain& d = bp.doms[tl.dom];
    if (bp.step >= d.active_steps) continue;
    const int parity = (d.parity + bp.step) & 1;
    const StepParams p = batch_step_params<S>(bp, d, parity);
    const uint32_t lo = tl.first + (u % UNITS) * (INNER * BATCH_PASS_NODES);
    const uint32_t end = min(p.N, lo + INNER * BATCH_PASS_NODES);
    constexpr bool TWO_PATHS = sizeof(S) == 2 && COLL == COLL_BGK;
    if constexpr (INNER == 1) {
      const uint32_t n = lo + threadIdx.x;
      if (n < end) {
        if (TWO_PATHS && !d.any_wall)
          batch_node<S, COLL, false>(bp, d, p, n);
        else
          batch_node<S, COLL, true>(bp, d, p, n);
      }
    } else {
#pragma nounroll
      for (uint32_t n = lo + threadIdx.x; n < end; n += BATCH_PASS_NODES) {
        if (TWO_PATHS && !d.any_wall)
          batch_node<S, COLL, false>(bp, d, p, n);
        else
          batch_node<S, COLL, true>(bp, d, p, n);
      }
    }
  }
}

}  // namespace hlbm
