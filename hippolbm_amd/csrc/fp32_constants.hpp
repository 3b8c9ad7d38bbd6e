// fp32_constants.hpp -- host-side constants of the FP32-compute forms (no
// device code; tests/test_fp32_constants.py compiles it with g++).
#pragma once
#include <cmath>

namespace hlbm {

// Constants of the FP32 BGK update g' = k1 g + kw_c (dr + rho t_i): kw_1 is
// the float nearest 1/(18 tau) for which kw_0 = 6 kw_1 and k1 = 1 - 18 kw_1
// are floats too (kw_2 = kw_1 / 2 always is).  In exact arithmetic the update
// then keeps mass and momentum (k1 + kw_0 + 6 kw_1 + 12 kw_2 = k1 + 6 kw_1 +
// 24 kw_2 = 1), so only the rounding of each operation remains; rounding the
// four constants independently leaves a systematic momentum change of up to
// 1e-7 per step, which a force-driven channel amplifies by H^2 / nu.
inline void bgk_fp32_constants(double tau, float& k1, float kw[3]) {
  const double target = 1.0 / (18.0 * tau);
  const float base = (float)target;
  float up = base, dn = base, best = base;
  bool found = false;
  for (int k = 0; k < 4096 && !found; k++) {
    for (int side = 0; side < 2 && !found; side++) {
      const double d = side ? (double)dn : (double)up;
      if ((double)(float)(1.0 - 18.0 * d) == 1.0 - 18.0 * d && (double)(float)(6.0 * d) == 6.0 * d) {
        best = (float)d;
        found = true;
      }
    }
    up = std::nextafter(up, 2.0f * base + 1.0f);
    dn = std::nextafter(dn, 0.0f);
  }
  kw[1] = best;  // tau outside any practical range falls back to independent rounding
  kw[2] = best * 0.5f;
  kw[0] = (float)(6.0 * (double)best);
  k1 = (float)(1.0 - 18.0 * (double)best);
}

// FP32 MRT rates s1 s2 s4 s9 s10 s13 s16 weps wepsj wxx: the custom ones, or
// hippoLBM's defaults (mrt.hpp:43-47) with s9 = s13 = 1/tau.
inline void mrt_fp32_rates(int custom, const double* mrt, double inv_tau, float out[10]) {
  if (custom) {
    for (int k = 0; k < 10; k++) out[k] = (float)mrt[k];
    return;
  }
  const double d[10] = {1.19, 1.4, 1.2, inv_tau, 1.4, inv_tau, 1.98, 0.0, -475.0 / 63.0, 0.0};
  for (int k = 0; k < 10; k++) out[k] = (float)d[k];
}

}  // namespace hlbm
