// Host build of the kernels' FP64 helpers for unit testing (tests/test_fastmath.py).
#include "../../awflcloud_amd/csrc/awfl_fastmath.hpp"
extern "C" void probe_pow_gamma(const double* x, double* y, long n) {
  for (long i = 0; i < n; i++) y[i] = awfl::pow_gamma(x[i]);
}
extern "C" void probe_rcp(const double* x, double* y, long n) {
  for (long i = 0; i < n; i++) y[i] = awfl::rcp(x[i]);
}
extern "C" void probe_pow_gamma_near(const double* x, const double* B, const double* Bh, const double* Bl,
                                     double* y, long n) {
  for (long i = 0; i < n; i++) y[i] = awfl::pow_gamma_near(x[i], (x[i] - B[i]) / B[i], Bh[i], Bl[i]);
}
