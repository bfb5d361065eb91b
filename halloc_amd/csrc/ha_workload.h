// ha_workload.h -- workload arithmetic shared by the device drivers and the
// host tests: the Poisson(8) inversion table of the random-graph degrees.
#pragma once

#include <cstdint>

namespace ha {

constexpr uint32_t POISSON8_N = 40;  // P(X >= 40) < 1e-14 for X ~ Poisson(8)

// t[k] = floor(2^24 * P(X <= k)), X ~ Poisson(8); a degree is the number of
// t[k] at or below a uniform 24-bit integer
struct PoissonCdf {
  uint32_t t[POISSON8_N];
};
constexpr PoissonCdf poisson8_cdf() {
  PoissonCdf c{};
  double p = 3.3546262790251185e-4;  // P(X = 0) = exp(-8)
  double cum = 0.0;
  for (uint32_t k = 0; k < POISSON8_N; ++k) {
    cum += p;
    c.t[k] = cum >= 1.0 ? 0x1000000u : (uint32_t)(cum * 16777216.0);
    p = p * 8.0 / (double)(k + 1);
  }
  return c;
}
constexpr PoissonCdf kPoisson8 = poisson8_cdf();

}  // namespace ha
