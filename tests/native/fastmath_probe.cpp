// Host build of the kernels' FP64 helpers for unit testing (tests/test_fastmath.py).
#include "../../awflcloud_amd/csrc/awfl_fastmath.hpp"
extern "C" void probe_pow_gamma(const double* x, double* y, long n) {
  for (long i = 0; i < n; i++) y[i] = awfl::pow_gamma(x[i]);
}
extern "C" void probe_rcp(const double* x, double* y, long n) {
  for (long i = 0; i < n; i++) y[i] = awfl::rcp(x[i]);
}
