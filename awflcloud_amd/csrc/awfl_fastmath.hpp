// awfl_fastmath.hpp -- FP64 building blocks of the sweep kernels.
//
// * rcp(x): 1/x for positive (or negative) normal x from the hardware
//   reciprocal estimate refined by two Newton-Raphson steps (|err| <~ 1 ulp).
// * pow_gamma(x): x^GAMMA for positive normal x, GAMMA the FP64 constant
//   CP/CV of the equation of state p = C0 (rho theta)^GAMMA.  Computed as
//   exp(GAMMA * log(x)) with log in double-double (atanh series after
//   reduction of the mantissa to [sqrt(1/2), sqrt(2))), the product with GAMMA
//   in double-double, and exp by reduction modulo ln2 and a degree-15 Taylor
//   polynomial.  Max error ~0.85 ulp for x in [1e-3, 1e8] (tested against
//   200-bit mpmath).  About 70 VALU instructions against ~215 for the general
//   library pow (which handles negative bases, zeros, infinities, NaNs and
//   arbitrary exponents that never occur here).
//
// The same source compiles for the host, where it is unit-tested.
//
// Parity builds (test builds of the FP64 sweeps that attribute their
// differences from the CPU oracle; AWFL_PARITY is a bit mask, 0 in the
// library): AWFL_PAR_DIV makes rcp() an IEEE division and sqrt_halfrsqrt()
// the correctly rounded sqrt and a division; AWFL_PAR_POW makes
// pow_gamma_near() the general pow_gamma(); AWFL_PAR_LIBPOW makes pow_gamma()
// the device library's pow; AWFL_PAR_BRIDGE forms the WENO bridge polynomial
// by its three statements instead of the folded map; AWFL_PAR_ORDER runs the
// cell and face bodies in the reference's operation order (awfl_cell_ref.hpp).
#pragma once
#define AWFL_PAR_DIV 1
#define AWFL_PAR_POW 2
#define AWFL_PAR_LIBPOW 4
#define AWFL_PAR_BRIDGE 8
#define AWFL_PAR_ORDER 16
#ifndef AWFL_PARITY
#define AWFL_PARITY 0
#endif
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
#define __host__
#define __device__
#define __forceinline__ inline
#endif

#include <cmath>

namespace awfl {
using std::fma;
using std::frexp;
using std::ldexp;
using std::rint;
using std::fabs;


constexpr double POW_GAMMA = 1.40027894002789400278940027894;

__host__ __device__ __forceinline__ double rcp(double x) {
#if AWFL_PARITY & AWFL_PAR_DIV
  return 1.0 / x;
#endif
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
#else
  double r = 1.0 / x;
#endif
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  return r;
}

// sqrt(x) and 1/(2 sqrt(x)) for a positive normal x: two Goldschmidt steps from
// the hardware reciprocal square root, then one Newton correction of the root
// (the refinement the library sqrt uses, without its scaling for denormal,
// zero and infinite arguments, which a squared sound speed never is)
__host__ __device__ __forceinline__ void sqrt_halfrsqrt(double x, double& s, double& hr) {
#if AWFL_PARITY & AWFL_PAR_DIV
  s = std::sqrt(x);
  hr = 0.5 / s;
#elif defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rsq(x);
  double g = x * r, h = 0.5 * r;
  double e = fma(-g, h, 0.5);
  h = fma(h, e, h);
  g = fma(g, e, g);
  e = fma(-g, h, 0.5);
  h = fma(h, e, h);
  g = fma(g, e, g);
  const double d = fma(-g, g, x);
  s = fma(d, h, g);
  hr = h;
#else
  s = std::sqrt(x);
  hr = 0.5 / s;
#endif
}

// ln 2 split so that k*LN2_HI is exact for |k| < 2^21
constexpr double LN2_HI = 0x1.62e42fefa3800p-1;
constexpr double LN2_LO = 0x1.ef35793c76730p-45;
constexpr double INV_LN2 = 0x1.71547652b82fep+0;
constexpr double SQRT_HALF = 0x1.6a09e667f3bcdp-1;

// 2/(2k+1), k = 1..11: atanh(s) = s + s^3 (1/3 + s^2/5 + ...), times 2
constexpr double LOGC[11] = {2.0 / 3, 2.0 / 5, 2.0 / 7, 2.0 / 9, 2.0 / 11, 2.0 / 13,
                             2.0 / 15, 2.0 / 17, 2.0 / 19, 2.0 / 21, 2.0 / 23};
// 1/n!, n = 2..15
constexpr double EXPC[14] = {1.0 / 2,
                             1.0 / 6,
                             1.0 / 24,
                             1.0 / 120,
                             1.0 / 720,
                             1.0 / 5040,
                             1.0 / 40320,
                             1.0 / 362880,
                             1.0 / 3628800,
                             1.0 / 39916800,
                             1.0 / 479001600,
                             1.0 / 6227020800.0,
                             1.0 / 87178291200.0,
                             1.0 / 1307674368000.0};

__host__ __device__ __forceinline__ double pow_gamma(double x) {
#if defined(__clang__)
  // no contraction of a*b+c: every rounding is the one written, so the device
  // and host builds give the same result
#pragma clang fp contract(off)
#endif
#if AWFL_PARITY & AWFL_PAR_LIBPOW
  return std::pow(x, POW_GAMMA);
#endif
  // x = m 2^e, m in [sqrt(1/2), sqrt(2))
  int e;
  double m = frexp(x, &e);
  if (m < SQRT_HALF) {
    m *= 2.0;
    e -= 1;
  }
  // s = (m-1)/(m+1) as sh + sl
  double num = m - 1.0;                 // exact
  double uh = m + 1.0;
  double bb = uh - 1.0;
  double ul = (1.0 - (uh - bb)) + (m - bb);   // two-sum error of 1 + m
  double ru = rcp(uh);
  double sh = num * ru;
  double rem = fma(-sh, uh, num) - sh * ul;
  double sl = rem * ru;
  // log(m) = 2 sh + (2 sl + 2 s^3 P(s^2))
  double s2 = sh * sh;
  double P = LOGC[10];
#pragma unroll
  for (int i = 9; i >= 0; i--) P = fma(P, s2, LOGC[i]);
  double tail = sh * s2 * P;
  double de = (double)e;
  double t1 = de * LN2_HI;              // exact
  double lh = 2.0 * sh;
  double Lh = t1 + lh;
  double b2 = Lh - t1;
  double Ll = ((t1 - (Lh - b2)) + (lh - b2)) + (fma(de, LN2_LO, 2.0 * sl) + tail);
  double L1 = Lh + Ll;
  double L2 = Ll - (L1 - Lh);
  // y = GAMMA * L (double-double)
  double yh = POW_GAMMA * L1;
  double yl = fma(POW_GAMMA, L1, -yh) + POW_GAMMA * L2;
  // exp(yh + yl) = 2^k exp(r)
  double k = rint(yh * INV_LN2);
  double rh = fma(-k, LN2_HI, yh);      // exact
  double t = fma(-k, LN2_LO, yl);
  double r = rh + t;
  double rl = (rh - r) + t;
  double Q = EXPC[13];
#pragma unroll
  for (int i = 12; i >= 0; i--) Q = fma(Q, r, EXPC[i]);
  double w = fma(r * r, Q, rl);
  double er = 1.0 + (r + w);
  return ldexp(er, (int)k);
}

// (1+e)^GAMMA by its binomial series, degree 12: truncation < 3e-19 relative
// for |e| <= 1/16.  Coefficients c_n = binom(GAMMA, n), c_0 = 1.
struct BinomGamma {
  double c[13];
  constexpr BinomGamma() : c() {
    c[0] = 1.0;
    for (int n = 1; n <= 12; n++) c[n] = c[n - 1] * (POW_GAMMA - (n - 1)) / n;
  }
};
constexpr BinomGamma BINOM_GAMMA{};
constexpr double NEAR_MAX = 0.0625;

// x^GAMMA for x = B (1 + e) near a reference value B whose power is known as
// a double-double Bh + Bl = B^GAMMA: B^GAMMA (1 + s) with s = (1+e)^GAMMA - 1
// from the series, added as Bh + (Bh s + Bl (1+s)) so only the last addition
// rounds at full size (~0.6 ulp).  Outside |e| <= 1/16 the general
// pow_gamma(x) is used.
// out-of-line general path, so the five inlined uses of the fast path do not
// each carry its code and constants
__host__ __device__ __attribute__((noinline)) inline double pow_gamma_far(double x) { return pow_gamma(x); }

__host__ __device__ __forceinline__ double pow_gamma_near(double x, double e, double Bh, double Bl) {
#if AWFL_PARITY & AWFL_PAR_POW
  return pow_gamma_far(x);
#endif
  if (fabs(e) <= NEAR_MAX) {
    double q = BINOM_GAMMA.c[12];
#pragma unroll
    for (int n = 11; n >= 1; n--) q = fma(q, e, BINOM_GAMMA.c[n]);
    double sv = q * e;
    double t = fma(Bl, sv, Bl);
    return Bh + fma(Bh, sv, t);
  }
  return pow_gamma_far(x);
}

}  // namespace awfl
