// Prints the Poisson(8) inversion table of halloc_amd/csrc/ha_workload.h, one
// threshold per line; tests/test_native.py compares it with scipy's CDF.
#include <cstdio>

#include "../../halloc_amd/csrc/ha_workload.h"

int main() {
  for (unsigned k = 0; k < ha::POISSON8_N; ++k) std::printf("%u\n", ha::kPoisson8.t[k]);
  return 0;
}
