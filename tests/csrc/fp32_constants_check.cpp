// Host check of fp32_constants.hpp: for tau over (0.5, 2] the FP32 BGK
// constants satisfy k1 + kw0 + 6 kw1 + 12 kw2 = 1 (mass) and
// k1 + 6 kw1 + 24 kw2 = 1 (momentum) exactly, with kw_c within 1e-6 (a few
// float ulps) of w_c / tau; beyond that range they still approximate it.
// Prints the largest deviations, exits non-zero on a failure.
#include <cmath>
#include <cstdio>
#include <initializer_list>

#include "../../hippolbm_amd/csrc/fp32_constants.hpp"

int main() {
  int bad = 0;
  double worst = 0;
  for (int k = 1; k <= 150000; k++) {
    const double tau = 0.5 + k * (1.5 / 150000.0);
    float k1, kw[3];
    hlbm::bgk_fp32_constants(tau, k1, kw);
    const long double mass = (long double)k1 + kw[0] + 6.0L * kw[1] + 12.0L * kw[2];
    const long double mom = (long double)k1 + 6.0L * kw[1] + 24.0L * kw[2];
    const double dev[3] = {std::fabs(kw[0] * 3.0 * tau - 1.0), std::fabs(kw[1] * 18.0 * tau - 1.0),
                           std::fabs(kw[2] * 36.0 * tau - 1.0)};
    for (double d : dev) worst = std::fmax(worst, d);
    if (mass != 1.0L || mom != 1.0L || dev[0] > 1e-6 || dev[1] > 1e-6 || dev[2] > 1e-6) bad++;
  }
  double far = 0;
  for (double tau : {3.0, 10.0, 100.0}) {
    float k1, kw[3];
    hlbm::bgk_fp32_constants(tau, k1, kw);
    far = std::fmax(far, std::fabs(kw[1] * 18.0 * tau - 1.0));
    far = std::fmax(far, std::fabs((1.0 - k1) * tau - 1.0));
  }
  float r[10];
  const double dflt[10] = {1.19, 1.4, 1.2, 1.0 / 0.7, 1.4, 1.0 / 0.7, 1.98, 0.0, -475.0 / 63.0, 0.0};
  hlbm::mrt_fp32_rates(0, nullptr, 1.0 / 0.7, r);
  for (int k = 0; k < 10; k++)
    if (r[k] != (float)dflt[k]) bad++;
  std::printf("failures %d worst %.3e far %.3e\n", bad, worst, far);
  return (bad == 0 && far < 1e-4) ? 0 : 1;
}
