// Host check of mrt_pairs.hpp: the pair forms of the moment transform and of
// its inverse against the basis tables the FP64 kernels use (MrtBasis::M and
// the exact inverse NUM/DEN), in double precision on random vectors.
// Prints the largest relative deviations and exits non-zero above 1e-13.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../hippolbm_amd/csrc/mrt_pairs.hpp"

using hlbm::MrtBasis;

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  double worst_m = 0, worst_f = 0, worst_rt = 0;
  for (int t = 0; t < 20000; t++) {
    double g[19], m[19], mref[19], f[19], fref[19], back[19];
    for (int i = 0; i < 19; i++) g[i] = U(rng) * (t % 3 == 0 ? 1e-3 : 1.0);
    hlbm::mrt_moments_pairs(g, m);
    for (int k = 0; k < 19; k++) {
      long double s = 0;
      for (int i = 0; i < 19; i++) s += (long double)MrtBasis::M[k][i] * g[i];
      mref[k] = (double)s;
    }
    double scale = 0;
    for (int k = 0; k < 19; k++) scale = std::fmax(scale, std::fabs(mref[k]));
    for (int k = 0; k < 19; k++) worst_m = std::fmax(worst_m, std::fabs(m[k] - mref[k]) / scale);
    // inverse of arbitrary moments against the exact inverse table
    for (int k = 0; k < 19; k++) m[k] = U(rng);
    hlbm::mrt_inverse_pairs(m, f);
    for (int i = 0; i < 19; i++) {
      long double s = 0;
      for (int k = 0; k < 19; k++) s += (long double)MrtBasis::NUM[i][k] / MrtBasis::DEN[i][k] * m[k];
      fref[i] = (double)s;
    }
    scale = 0;
    for (int i = 0; i < 19; i++) scale = std::fmax(scale, std::fabs(fref[i]));
    for (int i = 0; i < 19; i++) worst_f = std::fmax(worst_f, std::fabs(f[i] - fref[i]) / scale);
    // round trip
    hlbm::mrt_moments_pairs(g, m);
    hlbm::mrt_inverse_pairs(m, back);
    scale = 0;
    for (int i = 0; i < 19; i++) scale = std::fmax(scale, std::fabs(g[i]));
    for (int i = 0; i < 19; i++) worst_rt = std::fmax(worst_rt, std::fabs(back[i] - g[i]) / scale);
  }
  std::printf("moments %.3e inverse %.3e round-trip %.3e\n", worst_m, worst_f, worst_rt);
  return (worst_m < 1e-13 && worst_f < 1e-13 && worst_rt < 1e-13) ? 0 : 1;
}
