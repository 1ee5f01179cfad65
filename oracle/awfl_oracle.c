/* awfl_oracle.c -- CPU FP64 restatement of the reference ADER-WENO step.
 *
 * TEST INFRASTRUCTURE ONLY (the correctness oracle and the timed CPU
 * baseline).  Nothing in the product links or calls this file.
 *
 * Build: gcc -O3 -fopenmp (no -ffast-math, no -march, so no FMA contraction on
 * x86-64, as the reference CPU build: build/fhqwhgads_cpu.env:6).
 *
 * Every loop below follows the reference loop nest cited beside it, in the
 * same floating-point operation order.  "S:" abbreviates
 * Spatial_euler3d_cons_expl_cart_fv_Agrid.h.
 */
#include "awfl_oracle.h"
#include "../awflcloud_amd/csrc/awfl_tables.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* phys_params.h:4-10 */
static const double GRAV = 9.81;
static const double CP = 1004.;
static const double RD = 287.;
static const double P0 = 1.0e5;
static const double C0 = 27.5629410929725921310572974482;
static const double GAMMA = 1.40027894002789400278940027894;

enum { idR = 0, idU = 1, idV = 2, idW = 3, idT = 4 };   /* S:69-73 */
enum { BC_PERIODIC = 0, BC_WALL = 1 };                   /* S:75-76 */
enum { NUMSTATE = 5 };

#define N9 OR_MAXN
typedef double arr2[N9][N9];

struct oracle_ctx {
  oracle_config cfg;
  int ord, ngll, nader, hs;
  int nx, ny, nz, sim2d;
  double dx, dy, dz;
  double dtInit;
  int dimSwitch;
  /* Spatial::init tables (S:274-315) */
  double wenoRecon[N9][N9][N9];
  double c2g[N9][N9];        /* coefs_to_gll (ord x ngll)     */
  double s2g[N9][N9];        /* sten_to_gll (ord x ngll)      */
  double deriv[N9][N9];      /* derivMatrix (ngll x ngll)     */
  double idl[N9];
  double sigma;
  double gllPts_ord[N9], gllWts_ord[N9], gllPts_ngll[N9], gllWts_ngll[N9];
  int tv_lo_n, tv_hi_n;
  const int *tv_lo_j, *tv_lo_k, *tv_hi_j, *tv_hi_k;
  const double *tv_lo_c, *tv_hi_c;
  /* hydrostatic background (S:45-52) */
  double *hyDensCells, *hyPressureCells, *hyThetaCells, *hyDensThetaCells;
  double *hyDensGLL, *hyPressureGLL, *hyThetaGLL, *hyDensThetaGLL;
  /* scratch of the unfused sweep (S:327-330) and the tendency (Temporal_ader.h:36) */
  double *stateLimits, *stateFlux, *tend;
  /* tracers (S:331-339; numTracers, tracerPos) */
  int ntr;
  int tpos[OR_MAXTR];
  double *tracerLimits, *tracerFlux, *tracerTend, *tracerFac;
};

/* ------------------------------------------------------------------------ */
/* indexing                                                                 */
/* ------------------------------------------------------------------------ */
#define NXH (c->nx + 2 * c->hs)
#define NYH (c->ny + 2 * c->hs)
#define NZH (c->nz + 2 * c->hs)
#define ST(l, k, j, i) state[(((size_t)(l) * NZH + (k)) * NYH + (j)) * NXH + (i)]
#define TD(l, k, j, i) tend[(((size_t)(l) * c->nz + (k)) * c->ny + (j)) * c->nx + (i)]
#define SL(l, s, k, j, i) \
  c->stateLimits[((((size_t)(l) * 2 + (s)) * (c->nz + 1) + (k)) * (c->ny + 1) + (j)) * (c->nx + 1) + (i)]
#define SF(l, k, j, i) c->stateFlux[(((size_t)(l) * (c->nz + 1) + (k)) * (c->ny + 1) + (j)) * (c->nx + 1) + (i)]
#define HYGLL(a, k, kk) (a)[(size_t)(k) * c->ngll + (kk)]
#define TRC(t, k, j, i) tracers[(((size_t)(t) * NZH + (k)) * NYH + (j)) * NXH + (i)]
#define TTD(t, k, j, i) ttend[(((size_t)(t) * c->nz + (k)) * c->ny + (j)) * c->nx + (i)]
#define TLM(t, s, k, j, i) \
  c->tracerLimits[((((size_t)(t) * 2 + (s)) * (c->nz + 1) + (k)) * (c->ny + 1) + (j)) * (c->nx + 1) + (i)]
#define TFX(t, k, j, i) c->tracerFlux[(((size_t)(t) * (c->nz + 1) + (k)) * (c->ny + 1) + (j)) * (c->nx + 1) + (i)]

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* Sensitivity probe (tests only): when g_pow_jitter != 0 every pow() of the
 * step is moved by one ulp up or down, chosen by a hash of its arguments, to
 * measure how far ulp-level libm differences alone move the result. */
static uint64_t g_pow_jitter = 0;
static uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
/* Power mode 1 (attribution of the HIP kernels' differences, never the
 * default): x^GAMMA by the double-double exp(GAMMA log x) that the HIP parity
 * builds use (pow_gamma of awfl_fastmath.hpp with IEEE division), restated
 * operation for operation, so that both sides share one power function:
 * log of the mantissa in [sqrt(1/2), sqrt(2)) by the atanh series
 * log m = 2 atanh((m-1)/(m+1)), the product with GAMMA in double-double, exp
 * by Cody-Waite reduction modulo ln 2 and a degree-15 Taylor polynomial. */
static int g_pow_mode = 0;
void oracle_set_pow_mode(int mode) { g_pow_mode = mode; }
#ifndef __cplusplus   /* the op-counting C++ build counts the library pow */
static double dd_pow_gamma(double x) {
  static const double LN2_HI = 0x1.62e42fefa3800p-1, LN2_LO = 0x1.ef35793c76730p-45;
  static const double INV_LN2 = 0x1.71547652b82fep+0, SQRT_HALF = 0x1.6a09e667f3bcdp-1;
  static const double LOGC[11] = {2.0 / 3, 2.0 / 5, 2.0 / 7, 2.0 / 9, 2.0 / 11, 2.0 / 13,
                                  2.0 / 15, 2.0 / 17, 2.0 / 19, 2.0 / 21, 2.0 / 23};
  static const double EXPC[14] = {1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040, 1.0 / 40320,
                                  1.0 / 362880, 1.0 / 3628800, 1.0 / 39916800, 1.0 / 479001600,
                                  1.0 / 6227020800.0, 1.0 / 87178291200.0, 1.0 / 1307674368000.0};
  int e;
  double m = frexp(x, &e);
  if (m < SQRT_HALF) { m *= 2.0; e -= 1; }
  double num = m - 1.0, uh = m + 1.0, bb = uh - 1.0;
  double ul = (1.0 - (uh - bb)) + (m - bb);
  double ru = 1.0 / uh;
  double sh = num * ru;
  double rem = fma(-sh, uh, num) - sh * ul;
  double sl = rem * ru;
  double s2 = sh * sh, P = LOGC[10];
  for (int i = 9; i >= 0; i--) P = fma(P, s2, LOGC[i]);
  double tail = sh * s2 * P;
  double de = (double)e, t1 = de * LN2_HI, lh = 2.0 * sh;
  double Lh = t1 + lh, b2 = Lh - t1;
  double Ll = ((t1 - (Lh - b2)) + (lh - b2)) + (fma(de, LN2_LO, 2.0 * sl) + tail);
  double L1 = Lh + Ll, L2 = Ll - (L1 - Lh);
  double yh = GAMMA * L1;
  double yl = fma(GAMMA, L1, -yh) + GAMMA * L2;
  double k = rint(yh * INV_LN2);
  double rh = fma(-k, LN2_HI, yh), t = fma(-k, LN2_LO, yl);
  double r = rh + t, rl = (rh - r) + t, Q = EXPC[13];
  for (int i = 12; i >= 0; i--) Q = fma(Q, r, EXPC[i]);
  double w = fma(r * r, Q, rl);
  return ldexp(1.0 + (r + w), (int)k);
}
#endif
static double spow(double x, double y) {
  double r = pow(x, y);
#ifndef __cplusplus
  if (g_pow_mode == 1 && y == GAMMA) r = dd_pow_gamma(x);
#endif
  if (g_pow_jitter) {
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    uint64_t h = mix64(b ^ g_pow_jitter);
    r = nextafter(r, (h & 1) ? INFINITY : -INFINITY);
  }
  return r;
}
void oracle_set_pow_jitter(uint64_t seed) { g_pow_jitter = seed; }
double oracle_eos_pow(double x) { return spow(x, GAMMA); }

/* ------------------------------------------------------------------------ */
/* tables                                                                   */
/* ------------------------------------------------------------------------ */
#define TVSEL(N)                                                        \
  case N: *n = AWFL_TV_##N##_NTERMS; *j = awfl_tv_##N##_j; *k = awfl_tv_##N##_k; \
    *cc = awfl_tv_##N##_c; return 0;
static int tv_table(int n_, int* n, const int** j, const int** k, const double** cc) {
  switch (n_) { TVSEL(2) TVSEL(3) TVSEL(4) TVSEL(5) TVSEL(6) TVSEL(7) TVSEL(8) TVSEL(9) }
  return -1;
}

static const double* gll_pts(int n) {
  switch (n) {
    case 1: return awfl_gll_pts_1; case 2: return awfl_gll_pts_2; case 3: return awfl_gll_pts_3;
    case 4: return awfl_gll_pts_4; case 5: return awfl_gll_pts_5; case 6: return awfl_gll_pts_6;
    case 7: return awfl_gll_pts_7; case 8: return awfl_gll_pts_8; case 9: return awfl_gll_pts_9;
  }
  return 0;
}
static const double* gll_wts(int n) {
  switch (n) {
    case 1: return awfl_gll_wts_1; case 2: return awfl_gll_wts_2; case 3: return awfl_gll_wts_3;
    case 4: return awfl_gll_wts_4; case 5: return awfl_gll_wts_5; case 6: return awfl_gll_wts_6;
    case 7: return awfl_gll_wts_7; case 8: return awfl_gll_wts_8; case 9: return awfl_gll_wts_9;
  }
  return 0;
}

/* copy an [n][m] generated table into a [9][9] array */
static void put2(double dst[N9][N9], const double* src, int n, int m) {
  for (int a = 0; a < n; a++)
    for (int b = 0; b < m; b++) dst[a][b] = src[a * m + b];
}

static const double* mat_sq(const char* which, int n) {
#define MSQ(name)                                                                      \
  switch (n) {                                                                         \
    case 2: return &awfl_##name##_2[0][0]; case 3: return &awfl_##name##_3[0][0];       \
    case 4: return &awfl_##name##_4[0][0]; case 5: return &awfl_##name##_5[0][0];       \
    case 6: return &awfl_##name##_6[0][0]; case 7: return &awfl_##name##_7[0][0];       \
    case 8: return &awfl_##name##_8[0][0]; case 9: return &awfl_##name##_9[0][0];       \
  }                                                                                    \
  return 0;
  if (!strcmp(which, "g2c")) { MSQ(gll_to_coefs) }
  if (!strcmp(which, "c2g")) { MSQ(coefs_to_gll) }
  if (!strcmp(which, "c2d")) { MSQ(coefs_to_deriv) }
  return 0;
#undef MSQ
}

static const double* c2g_lower(int ord, int m) {
#define CL(o)                                                                              \
  if (ord == o) switch (m) {                                                               \
      case 1: return &awfl_coefs_to_gll_lower_##o##_1[0][0];                               \
      case 2: return &awfl_coefs_to_gll_lower_##o##_2[0][0];                               \
      case 3: return &awfl_coefs_to_gll_lower_##o##_3[0][0];                               \
    }
  CL(3) CL(5) CL(7) CL(9)
#undef CL
  if (ord >= 5 && m == 4) return ord == 5 ? &awfl_coefs_to_gll_lower_5_4[0][0]
                               : ord == 7 ? &awfl_coefs_to_gll_lower_7_4[0][0]
                                          : &awfl_coefs_to_gll_lower_9_4[0][0];
  if (ord >= 5 && m == 5) return ord == 5 ? &awfl_coefs_to_gll_lower_5_5[0][0]
                               : ord == 7 ? &awfl_coefs_to_gll_lower_7_5[0][0]
                                          : &awfl_coefs_to_gll_lower_9_5[0][0];
  if (ord >= 7 && m == 6) return ord == 7 ? &awfl_coefs_to_gll_lower_7_6[0][0] : &awfl_coefs_to_gll_lower_9_6[0][0];
  if (ord >= 7 && m == 7) return ord == 7 ? &awfl_coefs_to_gll_lower_7_7[0][0] : &awfl_coefs_to_gll_lower_9_7[0][0];
  if (ord == 9 && m == 8) return &awfl_coefs_to_gll_lower_9_8[0][0];
  if (ord == 9 && m == 9) return &awfl_coefs_to_gll_lower_9_9[0][0];
  return 0;
}

static const double* weno_recon_tab(int ord) {
  switch (ord) {
    case 3: return &awfl_weno_recon_3[0][0][0];
    case 5: return &awfl_weno_recon_5[0][0][0];
    case 7: return &awfl_weno_recon_7[0][0][0];
    case 9: return &awfl_weno_recon_9[0][0][0];
  }
  return 0;
}
static const double* s2c_tab(int ord) {
  switch (ord) {
    case 3: return &awfl_sten_to_coefs_3[0][0];
    case 5: return &awfl_sten_to_coefs_5[0][0];
    case 7: return &awfl_sten_to_coefs_7[0][0];
    case 9: return &awfl_sten_to_coefs_9[0][0];
  }
  return 0;
}

/* yakl::intrinsics::matmul_cr: ret(i,j) = sum_k left(k,j) * right(i,k) */
static void matmul_cr(double ret[N9][N9], double left[N9][N9], double right[N9][N9], int n_in, int n_mid,
                      int n_out) {
  double tmp[N9][N9];
  for (int i = 0; i < n_in; i++)
    for (int j = 0; j < n_out; j++) {
      double t = 0;
      for (int k = 0; k < n_mid; k++) t += left[k][j] * right[i][k];
      tmp[i][j] = t;
    }
  for (int i = 0; i < n_in; i++)
    for (int j = 0; j < n_out; j++) ret[i][j] = tmp[i][j];
}

/* weno::wenoSetIdealSigma + convexify (WenoLimiter.h:20-25, 28-90) */
static void weno_set_ideal_sigma(int ord, double* idl, double* sigma) {
  int n = (ord - 1) / 2 + 2;
  if (ord == 3) {
    *sigma = 0.0343557947899881; idl[0] = 1.; idl[1] = 1.; idl[2] = 1224.61619926508;
  } else if (ord == 5) {
    *sigma = 0.73564225445964; idl[0] = 1.; idl[1] = 73.564225445964; idl[2] = 1.; idl[3] = 1584.89319246111;
  } else if (ord == 7) {
    *sigma = 0.125594321575479; idl[0] = 1.; idl[1] = 7.35642254459641; idl[2] = 7.35642254459641;
    idl[3] = 1.; idl[4] = 794.328234724281;
  } else if (ord == 9) {
    *sigma = 0.0288539981181442; idl[0] = 1.; idl[1] = 2.15766927997459; idl[2] = 2.40224886796286;
    idl[3] = 2.15766927997459; idl[4] = 1.; idl[5] = 1136.12697719888;
  }
  double sum = 0.;
  const double eps = 1.0e-20;
  for (int i = 0; i < n; i++) sum += idl[i];
  for (int i = 0; i < n; i++) idl[i] /= (sum + eps);
}

/* ------------------------------------------------------------------------ */
/* profiles (Profiles.h:8-73)                                               */
/* ------------------------------------------------------------------------ */
static double initConstTheta_density(double t0, double z) {
  double exner = 1. - GRAV * z / (CP * t0);
  double p = pow(exner, CP / RD) * P0;
  double rt = pow(p / C0, 1. / GAMMA);
  return rt / t0;
}
static double initConstTheta_pressure(double t0, double z) {
  double r = initConstTheta_density(t0, z);
  return C0 * pow(r * t0, GAMMA);
}
static double initConstBVF_pot_temp(double t0, double bvf, double z) { return t0 * exp(bvf * bvf * z / GRAV); }
static double initConstBVF_density(double t0, double bvf, double z) {
  double t = initConstBVF_pot_temp(t0, bvf, z);
  double exner = 1. - GRAV * GRAV / (CP * bvf * bvf) * (t - t0) / (t * t0);
  double p = pow(exner, CP / RD) * P0;
  double rt = pow(p / C0, 1. / GAMMA);
  return rt / t;
}
static double initConstBVF_pressure(double t0, double bvf, double z) {
  double t = initConstBVF_pot_temp(t0, bvf, z);
  double r = initConstBVF_density(t0, bvf, z);
  return C0 * pow(r * t, GAMMA);
}
/* Point values of the two background families at height z (for tests):
 * family 0 constant theta (t0 = 300 K, S:378-383), 1 constant Brunt-Vaisala
 * frequency (t0 = 300 K, bvf = 0.01 1/s, S:384-391). */
void oracle_profile(int family, double z, double* th, double* rh, double* ph) {
  if (family == 0) {
    *th = 300;
    *rh = initConstTheta_density(300, z);
    *ph = initConstTheta_pressure(300, z);
  } else {
    *th = initConstBVF_pot_temp(300, 0.01, z);
    *rh = initConstBVF_density(300, 0.01, z);
    *ph = initConstBVF_pressure(300, 0.01, z);
  }
}
static double ellipsoid_linear(double x, double y, double z, double x0, double y0, double z0, double xrad,
                               double yrad, double zrad, double amp) {
  double xn = (x - x0) / xrad, yn = (y - y0) / yrad, zn = (z - z0) / zrad;
  double dist = sqrt(xn * xn + yn * yn + zn * zn);
  return amp * dmax(1. - dist, 0.);
}
static double ellipsoid_cosine(double x, double y, double z, double x0, double y0, double z0, double xrad,
                               double yrad, double zrad, double amp, double pwr) {
  double val = 0;
  double xn = (x - x0) / xrad, yn = (y - y0) / yrad, zn = (z - z0) / zrad;
  double dist = sqrt(xn * xn + yn * yn + zn * zn);
  if (dist <= 1.) val = amp * pow((cos(M_PI * dist) + 1) / 2, pwr);
  return val;
}

/* ------------------------------------------------------------------------ */
/* context                                                                  */
/* ------------------------------------------------------------------------ */
static void background(oracle_ctx* c);
static int nthreads_of(const oracle_ctx* c) {
#ifdef _OPENMP
  return c->cfg.nthreads > 0 ? c->cfg.nthreads : omp_get_max_threads();
#else
  (void)c;
  return 1;
#endif
}

oracle_ctx* oracle_create(const oracle_config* cfg) {
  int ord = cfg->ord, ngll = cfg->ngll;
  if (!(ord == 3 || ord == 5 || ord == 7 || ord == 9) || ngll < 2 || ngll > ord) return 0;
  if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return 0;
  oracle_ctx* c = (oracle_ctx*)calloc(1, sizeof(oracle_ctx));
  c->cfg = *cfg;
  c->ord = ord; c->ngll = ngll; c->nader = ngll; c->hs = (ord - 1) / 2;
  c->nx = cfg->nx; c->ny = cfg->ny; c->nz = cfg->nz;
  c->sim2d = cfg->ny == 1;                              /* S:217 */
  c->dx = cfg->xlen / cfg->nx; c->dy = cfg->ylen / cfg->ny; c->dz = cfg->zlen / cfg->nz;  /* S:268-270 */
  c->dtInit = 0; c->dimSwitch = 1;                      /* S:193-194 */
  /* S:274-296 */
  const double* wr = weno_recon_tab(ord);
  int hs2 = c->hs + 2;
  for (int p = 0; p < hs2; p++)
    for (int s = 0; s < ord; s++)
      for (int q = 0; q < ord; q++) c->wenoRecon[p][s][q] = wr[(p * ord + s) * ord + q];
  {
    double s2c[N9][N9], c2gl[N9][N9];
    put2(s2c, s2c_tab(ord), ord, ord);
    put2(c2gl, c2g_lower(ord, ngll), ord, ngll);
    put2(c->c2g, c2g_lower(ord, ngll), ord, ngll);
    matmul_cr(c->s2g, c2gl, s2c, ord, ord, ngll);
  }
  /* S:298-309 */
  {
    double g2c[N9][N9], c2d[N9][N9], c2g[N9][N9], inner[N9][N9];
    put2(g2c, mat_sq("g2c", ngll), ngll, ngll);
    put2(c2d, mat_sq("c2d", ngll), ngll, ngll);
    put2(c2g, mat_sq("c2g", ngll), ngll, ngll);
    matmul_cr(inner, c2d, g2c, ngll, ngll, ngll);
    matmul_cr(c->deriv, c2g, inner, ngll, ngll, ngll);
  }
  /* S:310-315 */
  memcpy(c->gllPts_ord, gll_pts(ord), sizeof(double) * ord);
  memcpy(c->gllWts_ord, gll_wts(ord), sizeof(double) * ord);
  memcpy(c->gllPts_ngll, gll_pts(ngll), sizeof(double) * ngll);
  memcpy(c->gllWts_ngll, gll_wts(ngll), sizeof(double) * ngll);
  weno_set_ideal_sigma(ord, c->idl, &c->sigma);
  tv_table(c->hs + 1, &c->tv_lo_n, &c->tv_lo_j, &c->tv_lo_k, &c->tv_lo_c);
  tv_table(ord, &c->tv_hi_n, &c->tv_hi_j, &c->tv_hi_k, &c->tv_hi_c);
  /* S:327-339 */
  size_t nlim = (size_t)NUMSTATE * 2 * (c->nz + 1) * (c->ny + 1) * (c->nx + 1);
  size_t nflx = (size_t)NUMSTATE * (c->nz + 1) * (c->ny + 1) * (c->nx + 1);
  size_t ntend = (size_t)NUMSTATE * c->nz * c->ny * c->nx;
  c->stateLimits = (double*)calloc(nlim, sizeof(double));
  c->stateFlux = (double*)calloc(nflx, sizeof(double));
  c->tend = (double*)calloc(ntend, sizeof(double));
  int nzh = c->nz + 2 * c->hs;
  c->hyDensCells = (double*)calloc(nzh, sizeof(double));
  c->hyPressureCells = (double*)calloc(nzh, sizeof(double));
  c->hyThetaCells = (double*)calloc(nzh, sizeof(double));
  c->hyDensThetaCells = (double*)calloc(nzh, sizeof(double));
  c->hyDensGLL = (double*)calloc((size_t)c->nz * ngll, sizeof(double));
  c->hyPressureGLL = (double*)calloc((size_t)c->nz * ngll, sizeof(double));
  c->hyThetaGLL = (double*)calloc((size_t)c->nz * ngll, sizeof(double));
  c->hyDensThetaGLL = (double*)calloc((size_t)c->nz * ngll, sizeof(double));
  if (!c->stateLimits || !c->stateFlux || !c->tend) { oracle_destroy(c); return 0; }
  background(c);
  return c;
}

void oracle_destroy(oracle_ctx* c) {
  if (!c) return;
  free(c->stateLimits); free(c->stateFlux); free(c->tend);
  free(c->tracerLimits); free(c->tracerFlux); free(c->tracerTend); free(c->tracerFac);
  free(c->hyDensCells); free(c->hyPressureCells); free(c->hyThetaCells); free(c->hyDensThetaCells);
  free(c->hyDensGLL); free(c->hyPressureGLL); free(c->hyThetaGLL); free(c->hyDensThetaGLL);
  free(c);
}

int oracle_hs(const oracle_ctx* c) { return c->hs; }
int oracle_get_dim_switch(const oracle_ctx* c) { return c->dimSwitch; }
void oracle_set_dim_switch(oracle_ctx* c, int v) { c->dimSwitch = v ? 1 : 0; }
void oracle_set_dt_cache(oracle_ctx* c, double dt) { c->dtInit = dt; }

int oracle_get_table(const oracle_ctx* c, int which, double* out) {
  int n = 0, ord = c->ord, ngll = c->ngll;
  switch (which) {
    case 0:
      for (int p = 0; p < ord; p++)
        for (int s = 0; s < ord; s++)
          for (int q = 0; q < ord; q++) out[n++] = c->wenoRecon[p][s][q];
      return n;
    case 1: for (int a = 0; a < ord; a++) for (int b = 0; b < ngll; b++) out[n++] = c->c2g[a][b]; return n;
    case 2: for (int a = 0; a < ord; a++) for (int b = 0; b < ngll; b++) out[n++] = c->s2g[a][b]; return n;
    case 3: for (int a = 0; a < ngll; a++) for (int b = 0; b < ngll; b++) out[n++] = c->deriv[a][b]; return n;
    case 4: for (int a = 0; a < c->hs + 2; a++) out[n++] = c->idl[a]; out[n++] = c->sigma; return n;
    case 5: for (int a = 0; a < ord; a++) out[n++] = c->gllPts_ord[a]; return n;
    case 6: for (int a = 0; a < ord; a++) out[n++] = c->gllWts_ord[a]; return n;
    case 7: for (int a = 0; a < ngll; a++) out[n++] = c->gllPts_ngll[a]; return n;
    case 8: for (int a = 0; a < ngll; a++) out[n++] = c->gllWts_ngll[a]; return n;
  }
  return -1;
}

int oracle_get_background(const oracle_ctx* c, int which, double* out) {
  int nzh = c->nz + 2 * c->hs, ng = c->nz * c->ngll;
  const double* src[8] = {c->hyDensCells, c->hyPressureCells, c->hyThetaCells, c->hyDensThetaCells,
                          c->hyDensGLL, c->hyPressureGLL, c->hyThetaGLL, c->hyDensThetaGLL};
  if (which < 0 || which > 7) return -1;
  int n = which < 4 ? nzh : ng;
  memcpy(out, src[which], sizeof(double) * n);
  return n;
}

/* S:370-425 */
static void background(oracle_ctx* c) {
  int hs = c->hs, ord = c->ord, ngll = c->ngll;
  double dz = c->dz;
  int igw = c->cfg.data_spec == 4;
  for (int k = 0; k < c->nz + 2 * hs; k++) {
    c->hyDensCells[k] = 0; c->hyPressureCells[k] = 0; c->hyThetaCells[k] = 0; c->hyDensThetaCells[k] = 0;
    for (int kk = 0; kk < ord; kk++) {
      double zloc = (k - hs + 0.5) * dz + c->gllPts_ord[kk] * dz;
      double th, rh, ph;
      if (!igw) {
        th = 300;
        rh = initConstTheta_density(th, zloc);
        ph = initConstTheta_pressure(th, zloc);
      } else {
        const double bvf = 0.01, t0 = 300;
        th = initConstBVF_pot_temp(t0, bvf, zloc);
        rh = initConstBVF_density(t0, bvf, zloc);
        ph = initConstBVF_pressure(t0, bvf, zloc);
      }
      double wt = c->gllWts_ord[kk];
      c->hyDensCells[k] += rh * wt;
      c->hyThetaCells[k] += th * wt;
      c->hyDensThetaCells[k] += rh * th * wt;
      c->hyPressureCells[k] += ph * wt;
    }
  }
  for (int k = 0; k < c->nz; k++) {
    for (int kk = 0; kk < ngll; kk++) {
      double zloc = (k + 0.5) * dz + c->gllPts_ngll[kk] * dz;
      double th, rh, ph;
      if (!igw) {
        th = 300;
        rh = initConstTheta_density(th, zloc);
        ph = initConstTheta_pressure(th, zloc);
      } else {
        const double bvf = 0.01, t0 = 300;
        th = initConstBVF_pot_temp(t0, bvf, zloc);
        rh = initConstBVF_density(t0, bvf, zloc);
        ph = initConstBVF_pressure(t0, bvf, zloc);
      }
      HYGLL(c->hyDensGLL, k, kk) = rh;
      HYGLL(c->hyThetaGLL, k, kk) = th;
      HYGLL(c->hyDensThetaGLL, k, kk) = rh * th;
      HYGLL(c->hyPressureGLL, k, kk) = ph;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* initial data                                                             */
/* ------------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
double oracle_uniform_pm1(uint64_t seed, uint64_t idx) {
  uint64_t h = splitmix64(splitmix64(seed) ^ idx);
  return (double)(h >> 11) * 0x1.0p-52 - 1.0;
}

/* initTracers (S:505-560): tracer 0 uniform (rh * 1), 1 a replica of the
 * thermal's theta bubble (ellipsoid_linear, amplitude 2, at (xlen/2, ylen/2,
 * 2000) with radii 2000), 2 a block (k in [nz/10, 3nz/10), i and, in 3-D, j
 * in [4n/10, 6n/10)); every value times the constant-theta (300 K) density
 * rh, integrated with the ord^3 GLL quadrature.  Tracers beyond the third
 * stay zero. */
void oracle_init_tracers(const oracle_ctx* c, double* tracers) {
  int hs = c->hs, ord = c->ord, nx = c->nx, ny = c->ny, nz = c->nz;
  double dx = c->dx, dy = c->dy, dz = c->dz, xlen = c->cfg.xlen, ylen = c->cfg.ylen;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int l = 0; l < c->ntr; l++) {
          TRC(l, hs + k, hs + j, hs + i) = 0;
          for (int kk = 0; kk < ord; kk++)
            for (int jj = 0; jj < ord; jj++)
              for (int ii = 0; ii < ord; ii++) {
                double zloc = (k + 0.5) * dz + c->gllPts_ord[kk] * dz;
                double yloc = c->sim2d ? ylen / 2 : (j + 0.5) * dy + c->gllPts_ord[jj] * dy;
                double xloc = (i + 0.5) * dx + c->gllPts_ord[ii] * dx;
                double wt = c->gllWts_ord[kk] * c->gllWts_ord[jj] * c->gllWts_ord[ii];
                double th = 300;
                double rh = initConstTheta_density(th, zloc);
                if (l == 0) {
                  TRC(l, hs + k, hs + j, hs + i) += rh * 1 * wt;
                } else if (l == 1) {
                  double tval = ellipsoid_linear(xloc, yloc, zloc, xlen / 2, ylen / 2, 2000, 2000, 2000, 2000, 2);
                  TRC(l, hs + k, hs + j, hs + i) += rh * tval * wt;
                } else if (l == 2) {
                  int inside = k >= 1 * nz / 10 && k < 3 * nz / 10 && i >= 4 * nx / 10 && i < 6 * nx / 10;
                  if (!c->sim2d) inside = inside && j >= 4 * ny / 10 && j < 6 * ny / 10;
                  double tval = inside ? 1 : 0;
                  TRC(l, hs + k, hs + j, hs + i) += rh * tval * wt;
                }
              }
        }
}

/* S:428-485 with a general list of ellipsoids, per-cell noise and a
 * band-limited field (the synthetic inputs of the benchmarks). */
void oracle_init_state(const oracle_ctx* c, const oracle_init_spec* sp, double* state) {
  int hs = c->hs, ord = c->ord;
  double dx = c->dx, dy = c->dy, dz = c->dz;
  double xlen = c->cfg.xlen, ylen = c->cfg.ylen, zlen = c->cfg.zlen;
  int igw = c->cfg.data_spec == 4;
  int nthr = nthreads_of(c);
#pragma omp parallel for collapse(2) schedule(static) num_threads(nthr)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        for (int l = 0; l < NUMSTATE; l++) ST(l, hs + k, hs + j, hs + i) = 0;
        double cell_tp = 0;
        if (sp->noise_amp != 0)
          cell_tp += sp->noise_amp * oracle_uniform_pm1(sp->noise_seed, ((uint64_t)k * c->ny + j) * c->nx + i);
        if (sp->nmodes > 0) {
          double xc = (i + 0.5) * dx, yc = (j + 0.5) * dy, zc = (k + 0.5) * dz;
          double acc = 0;
          for (int m = 0; m < sp->nmodes; m++) {
            const double* md = sp->modes + 5 * m;
            acc += md[3] * cos(2 * M_PI * (md[0] * xc / xlen + md[1] * yc / ylen + md[2] * zc / zlen) + md[4]);
          }
          cell_tp += acc;
        }
        for (int kk = 0; kk < ord; kk++)
          for (int jj = 0; jj < ord; jj++)
            for (int ii = 0; ii < ord; ii++) {
              double zloc = (k + 0.5) * dz + c->gllPts_ord[kk] * dz;
              double yloc;
              if (c->sim2d) yloc = sp->y2d != 0 ? sp->y2d : ylen / 2;
              else yloc = (j + 0.5) * dy + c->gllPts_ord[jj] * dy;
              double xloc = (i + 0.5) * dx + c->gllPts_ord[ii] * dx;
              double wt = c->gllWts_ord[kk] * c->gllWts_ord[jj] * c->gllWts_ord[ii];
              double th, rh;
              if (!igw) {
                th = 300;
                rh = initConstTheta_density(th, zloc);
              } else {
                const double t0 = 300, bvf = 0.01;
                rh = initConstBVF_density(t0, bvf, zloc);
                th = initConstBVF_pot_temp(t0, bvf, zloc);
              }
              double tp = 0;
              for (int b = 0; b < sp->nbub; b++) {
                if (sp->shape[b] == 0)
                  tp += ellipsoid_linear(xloc, yloc, zloc, sp->x0[b], sp->y0[b], sp->z0[b], sp->xr[b], sp->yr[b],
                                         sp->zr[b], sp->amp[b]);
                else if (sp->shape[b] == 2) {   /* Profiles.h:76-82 */
                  double xn = (xloc - sp->x0[b]) / sp->xr[b], yn = (yloc - sp->y0[b]) / sp->yr[b];
                  tp += sp->amp[b] * sin(M_PI * zloc / zlen) / sqrt(1 + xn * xn + yn * yn);
                } else
                  tp += ellipsoid_cosine(xloc, yloc, zloc, sp->x0[b], sp->y0[b], sp->z0[b], sp->xr[b], sp->yr[b],
                                         sp->zr[b], sp->amp[b], sp->pwr[b]);
              }
              tp += cell_tp;
              double t = th + tp;
              ST(idT, hs + k, hs + j, hs + i) += (rh * t - rh * th) * wt;
              if (sp->u0 != 0) ST(idU, hs + k, hs + j, hs + i) += sp->u0 * rh * wt;
              if (sp->v0 != 0 && !c->sim2d) ST(idV, hs + k, hs + j, hs + i) += sp->v0 * rh * wt;
            }
        if (sp->perturb_amp != 0) {
          uint64_t h = splitmix64(splitmix64(sp->perturb_seed) ^ (((uint64_t)k * c->ny + j) * c->nx + i));
          double rval = (double)(h >> 11) * 0x1.0p-53;
          ST(idT, hs + k, hs + j, hs + i) *= 1. + rval * sp->perturb_amp;
        }
      }
}

/* ------------------------------------------------------------------------ */
/* computeTimeStep (S:159-187)                                              */
/* ------------------------------------------------------------------------ */
double oracle_compute_time_step(oracle_ctx* c, const double* state, double cfl) {
  if (c->dtInit <= 0) {
    int hs = c->hs;
    double mn = INFINITY;
    for (int k = 0; k < c->nz; k++)
      for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nx; i++) {
          double r = ST(idR, hs + k, hs + j, hs + i) + c->hyDensCells[hs + k];
          double u = ST(idU, hs + k, hs + j, hs + i) / r;
          double v = ST(idV, hs + k, hs + j, hs + i) / r;
          double w = ST(idW, hs + k, hs + j, hs + i) / r;
          double t = (ST(idT, hs + k, hs + j, hs + i) + c->hyDensThetaCells[hs + k]) / r;
          double p = C0 * spow(r * t, GAMMA);
          double cs = sqrt(GAMMA * p / r);
          double udt = cfl * c->dx / dmax(fabs(u - cs), fabs(u + cs));
          double vdt = cfl * c->dy / dmax(fabs(v - cs), fabs(v + cs));
          double wdt = cfl * c->dz / dmax(fabs(w - cs), fabs(w + cs));
          double d = dmin(dmin(udt, vdt), wdt);
          if (d < mn) mn = d;
        }
    c->dtInit = mn;
  }
  return c->dtInit;
}

/* ------------------------------------------------------------------------ */
/* WENO (WenoLimiter.h:10-172)                                              */
/* ------------------------------------------------------------------------ */
static double coefs_to_tv(int n, const int* tj, const int* tk, const double* tc, const double* a) {
  /* generated expression: c0*(a*a) + c1*(a*a) + c2*a_j*a_k + ..., summed left to right */
  double r = 0;
  for (int t = 0; t < n; t++) {
    double term = (tj[t] == tk[t]) ? tc[t] * (a[tj[t]] * a[tk[t]]) : tc[t] * a[tj[t]] * a[tk[t]];
    r = (t == 0) ? term : r + term;
  }
  return r;
}

static void map_weights(int n, const double* idl, double* wts) {
  for (int i = 0; i < n; i++)
    wts[i] = wts[i] * (idl[i] + idl[i] * idl[i] - 3. * idl[i] * wts[i] + wts[i] * wts[i]) /
             (idl[i] * idl[i] + wts[i] * (1. - 2. * idl[i]));
}

static void convexify(int n, double* wts) {
  double sum = 0.;
  const double eps = 1.0e-20;
  for (int i = 0; i < n; i++) sum += wts[i];
  for (int i = 0; i < n; i++) wts[i] /= (sum + eps);
}

void oracle_weno_coefs(const oracle_ctx* c, const double* u, double* aw) {
  int ord = c->ord, hs = c->hs;
  double tv[N9], wts[N9], a[N9][N9], lotmp[N9], hitmp[N9], lo_avg;
  const double eps = 1.0e-20;
  for (int j = 0; j < hs + 2; j++)
    for (int i = 0; i < ord; i++) a[j][i] = 0.;
  for (int i = 0; i < hs + 1; i++)
    for (int ii = 0; ii < hs + 1; ii++)
      for (int s = 0; s < hs + 1; s++) a[i][ii] += c->wenoRecon[i][s][ii] * u[i + s];
  for (int ii = 0; ii < ord; ii++)
    for (int s = 0; s < ord; s++) a[hs + 1][ii] += c->wenoRecon[hs + 1][s][ii] * u[s];
  for (int i = 0; i < hs + 1; i++)
    for (int ii = 0; ii < hs + 1; ii++) a[hs + 1][ii] -= c->idl[i] * a[i][ii];
  for (int ii = 0; ii < ord; ii++) a[hs + 1][ii] /= c->idl[hs + 1];
  for (int i = 0; i < hs + 1; i++) {
    for (int ii = 0; ii < hs + 1; ii++) lotmp[ii] = a[i][ii];
    tv[i] = coefs_to_tv(c->tv_lo_n, c->tv_lo_j, c->tv_lo_k, c->tv_lo_c, lotmp);
  }
  for (int ii = 0; ii < ord; ii++) hitmp[ii] = a[hs + 1][ii];
  tv[hs + 1] = coefs_to_tv(c->tv_hi_n, c->tv_hi_j, c->tv_hi_k, c->tv_hi_c, hitmp);
  lo_avg = 0.;
  for (int i = 0; i < hs + 1; i++) lo_avg += tv[i];
  lo_avg /= hs + 1;
  tv[hs + 1] = lo_avg + (tv[hs + 1] - lo_avg) * c->sigma;
  for (int i = 0; i < hs + 2; i++) wts[i] = c->idl[i] / (tv[i] * tv[i] + eps);
  convexify(hs + 2, wts);
  map_weights(hs + 2, c->idl, wts);
  convexify(hs + 2, wts);
  for (int i = 0; i < ord; i++) aw[i] = 0.;
  for (int i = 0; i < hs + 2; i++)
    for (int ii = 0; ii < ord; ii++) aw[ii] += wts[i] * a[i][ii];
}

/* reconstruct_gll_values (S:1800-1833); writes DTs(0, 0..ngll-1) */
static void reconstruct(const oracle_ctx* c, const double* stencil, double* dts0, int doweno) {
  int ord = c->ord, ngll = c->ngll;
  if (doweno) {
    double wc[N9];
    oracle_weno_coefs(c, stencil, wc);
    for (int ii = 0; ii < ngll; ii++) {
      double tmp = 0;
      for (int s = 0; s < ord; s++) tmp += c->c2g[s][ii] * wc[s];
      dts0[ii] = tmp;
    }
  } else {
    for (int ii = 0; ii < ngll; ii++) {
      double tmp = 0;
      for (int s = 0; s < ord; s++) tmp += c->s2g[s][ii] * stencil[s];
      dts0[ii] = tmp;
    }
  }
}

void oracle_reconstruct(const oracle_ctx* c, const double* u, int doweno, double* gll) {
  reconstruct(c, u, gll, doweno);
}

/* ------------------------------------------------------------------------ */
/* ADER differential transforms (S:1837-2091)                               */
/* q = the momentum along the sweep (ru for x, rv for y, rw for z); the four */
/* nonlinear fluxes f1..f4 are q*u, q*v, q*w, q*t in variable order, with    */
/* the sweep-normal one carrying the pressure.                              */
/* ------------------------------------------------------------------------ */
static void dt_xy(const oracle_ctx* c, int dir, arr2 r, arr2 ru, arr2 rv, arr2 rw, arr2 rt, arr2 f1, arr2 f2,
                  arr2 f3, arr2 f4, arr2 rt_gamma, double dx) {
  int nAder = c->nader, ngll = c->ngll;
  for (int kt = 1; kt < nAder; kt++)
    for (int ii = 0; ii < ngll; ii++) {
      f1[kt][ii] = 0; f2[kt][ii] = 0; f3[kt][ii] = 0; f4[kt][ii] = 0; rt_gamma[kt][ii] = 0;
    }
  for (int kt = 0; kt < nAder - 1; kt++) {
    for (int ii = 0; ii < ngll; ii++) {
      double d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0;
      if (dir == 0) {
        /* X: S:1863-1882; f1..f4 = ruu, ruv, ruw, rut */
        for (int s = 0; s < ngll; s++) {
          d1 += c->deriv[s][ii] * (ru[kt][s]);
          if (kt == 0) d2 += c->deriv[s][ii] * (f1[kt][s] + C0 * rt_gamma[kt][s]);
          else d2 += c->deriv[s][ii] * (f1[kt][s] + C0 * rt_gamma[kt][s] / 2);
          d3 += c->deriv[s][ii] * (f2[kt][s]);
          d4 += c->deriv[s][ii] * (f3[kt][s]);
          d5 += c->deriv[s][ii] * (f4[kt][s]);
        }
        r[kt + 1][ii] = -d1 / dx / (kt + 1.);
        ru[kt + 1][ii] = -d2 / dx / (kt + 1.);
        rv[kt + 1][ii] = -d3 / dx / (kt + 1.);
        rw[kt + 1][ii] = -d4 / dx / (kt + 1.);
        rt[kt + 1][ii] = -d5 / dx / (kt + 1.);
      } else {
        /* Y: S:1946-1965; f1..f4 = rvu, rvv, rvw, rvt */
        for (int s = 0; s < ngll; s++) {
          d1 += c->deriv[s][ii] * rv[kt][s];
          d2 += c->deriv[s][ii] * f1[kt][s];
          if (kt == 0) d3 += c->deriv[s][ii] * (f2[kt][s] + C0 * rt_gamma[kt][s]);
          else d3 += c->deriv[s][ii] * (f2[kt][s] + C0 * rt_gamma[kt][s] / 2);
          d4 += c->deriv[s][ii] * f3[kt][s];
          d5 += c->deriv[s][ii] * f4[kt][s];
        }
        r[kt + 1][ii] = -d1 / dx / (kt + 1);
        ru[kt + 1][ii] = -d2 / dx / (kt + 1);
        rv[kt + 1][ii] = -d3 / dx / (kt + 1);
        rw[kt + 1][ii] = -d4 / dx / (kt + 1);
        rt[kt + 1][ii] = -d5 / dx / (kt + 1);
      }
    }
    for (int ii = 0; ii < ngll; ii++) {
      double t1 = 0, t2 = 0, t3 = 0, t4 = 0;
      arr2* q = dir == 0 ? (arr2*)ru : (arr2*)rv;
      for (int l = 0; l <= kt + 1; l++) {
        t1 += (*q)[l][ii] * ru[kt + 1 - l][ii] - r[l][ii] * f1[kt + 1 - l][ii];
        t2 += (*q)[l][ii] * rv[kt + 1 - l][ii] - r[l][ii] * f2[kt + 1 - l][ii];
        t3 += (*q)[l][ii] * rw[kt + 1 - l][ii] - r[l][ii] * f3[kt + 1 - l][ii];
        t4 += (*q)[l][ii] * rt[kt + 1 - l][ii] - r[l][ii] * f4[kt + 1 - l][ii];
      }
      f1[kt + 1][ii] = t1 / r[0][ii];
      f2[kt + 1][ii] = t2 / r[0][ii];
      f3[kt + 1][ii] = t3 / r[0][ii];
      f4[kt + 1][ii] = t4 / r[0][ii];
      double tg = 0;
      for (int l = 0; l <= kt; l++)
        tg += (kt + 1. - l) * (GAMMA * rt_gamma[l][ii] * rt[kt + 1 - l][ii] - rt[l][ii] * rt_gamma[kt + 1 - l][ii]);
      rt_gamma[kt + 1][ii] = (GAMMA * rt_gamma[0][ii] * rt[kt + 1][ii] + tg / (kt + 1.)) / rt[0][ii];
    }
  }
  for (int kt = 1; kt < nAder; kt++)
    for (int ii = 0; ii < ngll; ii++) rt_gamma[kt][ii] /= 2;
}

/* Z: S:2004-2091; f1..f4 = rwu, rwv, rww, rwt */
static void dt_z(const oracle_ctx* c, arr2 r, arr2 ru, arr2 rv, arr2 rw, arr2 rt, arr2 f1, arr2 f2, arr2 f3,
                 arr2 f4, arr2 rt_gamma, arr2 source, int k, double dz) {
  int nAder = c->nader, ngll = c->ngll, nz = c->nz, bc_z = c->cfg.bc_z;
  for (int kt = 1; kt < nAder; kt++)
    for (int ii = 0; ii < ngll; ii++) {
      f1[kt][ii] = 0; f2[kt][ii] = 0; f3[kt][ii] = 0; f4[kt][ii] = 0; rt_gamma[kt][ii] = 0;
    }
  for (int kt = 0; kt < nAder - 1; kt++) {
    for (int ii = 0; ii < ngll; ii++) {
      double d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0;
      for (int s = 0; s < ngll; s++) {
        d1 += c->deriv[s][ii] * rw[kt][s];
        d2 += c->deriv[s][ii] * f1[kt][s];
        d3 += c->deriv[s][ii] * f2[kt][s];
        if (kt == 0) d4 += c->deriv[s][ii] * (f3[kt][s] + C0 * rt_gamma[kt][s] - HYGLL(c->hyPressureGLL, k, s));
        else d4 += c->deriv[s][ii] * (f3[kt][s] + C0 * rt_gamma[kt][s] / 2);
        d5 += c->deriv[s][ii] * f4[kt][s];
      }
      r[kt + 1][ii] = -d1 / dz / (kt + 1);
      ru[kt + 1][ii] = -d2 / dz / (kt + 1);
      rv[kt + 1][ii] = -d3 / dz / (kt + 1);
      rw[kt + 1][ii] = -d4 / dz / (kt + 1) + source[kt][ii] / (kt + 1);
      rt[kt + 1][ii] = -d5 / dz / (kt + 1);
      if (bc_z == BC_WALL) {
        if (k == nz - 1) rw[kt + 1][ngll - 1] = 0;
        if (k == 0) rw[kt + 1][0] = 0;
      }
      source[kt + 1][ii] = -r[kt + 1][ii] * GRAV;
    }
    for (int ii = 0; ii < ngll; ii++) {
      double t1 = 0, t2 = 0, t3 = 0, t4 = 0;
      for (int l = 0; l <= kt + 1; l++) {
        t1 += rw[l][ii] * ru[kt + 1 - l][ii] - r[l][ii] * f1[kt + 1 - l][ii];
        t2 += rw[l][ii] * rv[kt + 1 - l][ii] - r[l][ii] * f2[kt + 1 - l][ii];
        t3 += rw[l][ii] * rw[kt + 1 - l][ii] - r[l][ii] * f3[kt + 1 - l][ii];
        t4 += rw[l][ii] * rt[kt + 1 - l][ii] - r[l][ii] * f4[kt + 1 - l][ii];
      }
      f1[kt + 1][ii] = t1 / r[0][ii];
      f2[kt + 1][ii] = t2 / r[0][ii];
      f3[kt + 1][ii] = t3 / r[0][ii];
      f4[kt + 1][ii] = t4 / r[0][ii];
      double tg = 0;
      for (int l = 0; l <= kt; l++)
        tg += (kt + 1 - l) * (GAMMA * rt_gamma[l][ii] * rt[kt + 1 - l][ii] - rt[l][ii] * rt_gamma[kt + 1 - l][ii]);
      rt_gamma[kt + 1][ii] = (GAMMA * rt_gamma[0][ii] * rt[kt + 1][ii] + tg / (kt + 1)) / rt[0][ii];
    }
  }
  for (int kt = 1; kt < nAder; kt++)
    for (int ii = 0; ii < ngll; ii++) rt_gamma[kt][ii] /= 2;
}

void oracle_diff_transform(const oracle_ctx* c, int dir, double* arrs, double dx, int k_cell) {
  arr2 a[11];
  int nv = dir == 2 ? 11 : 10;
  for (int v = 0; v < nv; v++)
    for (int kt = 0; kt < c->nader; kt++)
      for (int ii = 0; ii < c->ngll; ii++) a[v][kt][ii] = arrs[(v * c->nader + kt) * c->ngll + ii];
  if (dir == 2) dt_z(c, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], k_cell, dx);
  else dt_xy(c, dir, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], dx);
  for (int v = 0; v < nv; v++)
    for (int kt = 0; kt < c->nader; kt++)
      for (int ii = 0; ii < c->ngll; ii++) arrs[(v * c->nader + kt) * c->ngll + ii] = a[v][kt][ii];
}

/* compute_timeAvg_edges (S:2168-2177), copy form */
static void tavg_edges_copy(const oracle_ctx* c, arr2 dts, double* tavg, double dt) {
  int ngll = c->ngll;
  tavg[0] = dts[0][0];
  tavg[ngll - 1] = dts[0][ngll - 1];
  double dtmult = dt;
  for (int kt = 1; kt < c->nader; kt++) {
    tavg[0] += dts[kt][0] * dtmult / (kt + 1.);
    tavg[ngll - 1] += dts[kt][ngll - 1] * dtmult / (kt + 1.);
    dtmult *= dt;
  }
}
/* compute_timeAvg_edges (S:2142-2149), in-place form */
static void tavg_edges(const oracle_ctx* c, arr2 dts, double dt) {
  int ngll = c->ngll;
  double dtmult = dt;
  for (int kt = 1; kt < c->nader; kt++) {
    dts[0][0] += dts[kt][0] * dtmult / (kt + 1.);
    dts[0][ngll - 1] += dts[kt][ngll - 1] * dtmult / (kt + 1.);
    dtmult *= dt;
  }
}
/* compute_timeAvg (S:2130-2138), in-place, all points */
static void tavg_all(const oracle_ctx* c, arr2 dts, double dt) {
  double dtmult = dt;
  for (int kt = 1; kt < c->nader; kt++) {
    for (int ii = 0; ii < c->ngll; ii++) dts[0][ii] += dts[kt][ii] * dtmult / (kt + 1.);
    dtmult *= dt;
  }
}

/* ------------------------------------------------------------------------ */
/* Riemann fluxes                                                           */
/* ------------------------------------------------------------------------ */
static void riemann_x(const double* L, const double* R, double* f) {   /* S:842-887 */
  double q1_L = L[0], q1_R = R[0], q2_L = L[1], q2_R = R[1], q3_L = L[2], q3_R = R[2];
  double q4_L = L[3], q4_R = R[3], q5_L = L[4], q5_R = R[4];
  double r = 0.5 * (q1_L + q1_R);
  double u = 0.5 * (q2_L / q1_L + q2_R / q1_R);
  double v = 0.5 * (q3_L / q1_L + q3_R / q1_R);
  double w = 0.5 * (q4_L / q1_L + q4_R / q1_R);
  double t = 0.5 * (q5_L / q1_L + q5_R / q1_R);
  double p = C0 * spow(r * t, GAMMA);
  double cs2 = GAMMA * p / r;
  double cs = sqrt(cs2);
  double w1, w2, w3;
  if (u > 0) { w1 = q1_L - q5_L / t; w2 = q3_L - v * q5_L / t; w3 = q4_L - w * q5_L / t; }
  else { w1 = q1_R - q5_R / t; w2 = q3_R - v * q5_R / t; w3 = q4_R - w * q5_R / t; }
  double w5 = u * q1_R / (2 * cs) - q2_R / (2 * cs) + q5_R / (2 * t);
  double w6 = -u * q1_L / (2 * cs) + q2_L / (2 * cs) + q5_L / (2 * t);
  double q1 = w1 + w5 + w6;
  double q2 = u * w1 + (u - cs) * w5 + (u + cs) * w6;
  double q3 = w2 + v * w5 + v * w6;
  double q4 = w3 + w * w5 + w * w6;
  double q5 = t * w5 + t * w6;
  f[0] = q2;
  f[1] = q2 * q2 / q1 + C0 * spow(q5, GAMMA);
  f[2] = q2 * q3 / q1;
  f[3] = q2 * q4 / q1;
  f[4] = q2 * q5 / q1;
}

static void riemann_y(const double* L, const double* R, double* f) {   /* S:1202-1245 (averages of conservatives) */
  double q1_L = L[0], q1_R = R[0], q2_L = L[1], q2_R = R[1], q3_L = L[2], q3_R = R[2];
  double q4_L = L[3], q4_R = R[3], q5_L = L[4], q5_R = R[4];
  double r = 0.5 * (q1_L + q1_R);
  double u = 0.5 * (q2_L + q2_R);
  double v = 0.5 * (q3_L + q3_R);
  double w = 0.5 * (q4_L + q4_R);
  double t = 0.5 * (q5_L + q5_R);
  double p = C0 * spow(r * t, GAMMA);
  double cs2 = GAMMA * p / r;
  double cs = sqrt(cs2);
  double w1, w2, w3;
  if (v > 0) { w1 = q1_L - q5_L / t; w2 = q2_L - u * q5_L / t; w3 = q4_L - w * q5_L / t; }
  else { w1 = q1_R - q5_R / t; w2 = q2_R - u * q5_R / t; w3 = q4_R - w * q5_R / t; }
  double w5 = v * q1_R / (2 * cs) - q3_R / (2 * cs) + q5_R / (2 * t);
  double w6 = -v * q1_L / (2 * cs) + q3_L / (2 * cs) + q5_L / (2 * t);
  double q1 = w1 + w5 + w6;
  double q2 = w2 + u * w5 + u * w6;
  double q3 = v * w1 + (v - cs) * w5 + (v + cs) * w6;
  double q4 = w3 + w * w5 + w * w6;
  double q5 = t * w5 + t * w6;
  f[0] = q3;
  f[1] = q3 * q2 / q1;
  f[2] = q3 * q3 / q1 + C0 * spow(q5, GAMMA);
  f[3] = q3 * q4 / q1;
  f[4] = q3 * q5 / q1;
}

static void riemann_z(const oracle_ctx* c, const double* L, const double* R, int k, double* f) { /* S:1583-1629 */
  double q1_L = L[0], q1_R = R[0], q2_L = L[1], q2_R = R[1], q3_L = L[2], q3_R = R[2];
  double q4_L = L[3], q4_R = R[3], q5_L = L[4], q5_R = R[4];
  double r = 0.5 * (q1_L + q1_R);
  double u = 0.5 * (q2_L / q1_L + q2_R / q1_R);
  double v = 0.5 * (q3_L / q1_L + q3_R / q1_R);
  double w = 0.5 * (q4_L / q1_L + q4_R / q1_R);
  double t = 0.5 * (q5_L / q1_L + q5_R / q1_R);
  double p = C0 * spow(r * t, GAMMA);
  double cs2 = GAMMA * p / r;
  double cs = sqrt(cs2);
  double w1, w2, w3;
  if (w > 0) { w1 = q1_L - q5_L / t; w2 = q2_L - u * q5_L / t; w3 = q3_L - v * q5_L / t; }
  else { w1 = q1_R - q5_R / t; w2 = q2_R - u * q5_R / t; w3 = q3_R - v * q5_R / t; }
  double w5 = w * q1_R / (2 * cs) - q4_R / (2 * cs) + q5_R / (2 * t);
  double w6 = -w * q1_L / (2 * cs) + q4_L / (2 * cs) + q5_L / (2 * t);
  double q1 = w1 + w5 + w6;
  double q2 = w2 + u * w5 + u * w6;
  double q3 = w3 + v * w5 + v * w6;
  double q4 = w * w1 + (w - cs) * w5 + (w + cs) * w6;
  double q5 = t * w5 + t * w6;
  f[0] = q4;
  f[1] = q4 * q2 / q1;
  f[2] = q4 * q3 / q1;
  f[3] = q4 * q4 / q1 + C0 * spow(q5, GAMMA);
  f[4] = q4 * q5 / q1;
  if (k < c->nz) f[3] -= HYGLL(c->hyPressureGLL, k, 0);
  if (k == c->nz) f[3] -= HYGLL(c->hyPressureGLL, c->nz - 1, c->ngll - 1);
}

void oracle_riemann(const oracle_ctx* c, int dir, const double* qL, const double* qR, int kface, double* f) {
  if (dir == 0) riemann_x(qL, qR, f);
  else if (dir == 1) riemann_y(qL, qR, f);
  else riemann_z(c, qL, qR, kface, f);
}

/* ------------------------------------------------------------------------ */
/* computeTendenciesX (S:598-957)                                           */
/* ------------------------------------------------------------------------ */
#define OMP_FOR2 _Pragma("omp parallel for collapse(2) schedule(static) num_threads(nthr)")

/* ------------------------------------------------------------------------ */
/* Tracers.  The reference compiles every tracer block out (#define         */
/* NO_TRACERS, S:10); this restates the source text of those blocks.  Three */
/* departures, each where the text cannot run as written:                  */
/*  T1 positivity limiter (S:904-935, 1262-1294, 1646-1678): every face's    */
/*     reduction factor is formed from the unlimited fluxes, then applied   */
/*     (two passes).  The reference scales tracerFlux in place while the    */
/*     neighbouring faces read it, an order-dependent result.               */
/*  T2 the y limiter's v < 0 branch reads tracerFlux(tr,k,j,ind_j) and      */
/*     (tr,k,j,ind_j+1) (S:1283-1284), faces of the x index; here it reads  */
/*     the y faces ind_j and ind_j+1 as its v > 0 branch and the x limiter  */
/*     do.                                                                  */
/*  T3 the z limiter wraps ind_k like x and y (S:1657-1666 would index      */
/*     k = -1 and nz+1 on a periodic z axis).                               */
/* ------------------------------------------------------------------------ */
int oracle_set_tracers(oracle_ctx* c, int ntr, const int* pos) {
  if (ntr < 0 || ntr > OR_MAXTR) return -1;
  free(c->tracerLimits); free(c->tracerFlux); free(c->tracerTend); free(c->tracerFac);
  c->tracerLimits = c->tracerFlux = c->tracerTend = c->tracerFac = 0;
  c->ntr = ntr;
  for (int t = 0; t < OR_MAXTR; t++) c->tpos[t] = (t < ntr && pos) ? (pos[t] != 0) : 0;
  if (ntr == 0) return 0;
  size_t nf = (size_t)ntr * (c->nz + 1) * (c->ny + 1) * (c->nx + 1);
  c->tracerLimits = (double*)calloc(2 * nf, sizeof(double));
  c->tracerFlux = (double*)calloc(nf, sizeof(double));
  c->tracerFac = (double*)calloc(nf, sizeof(double));
  c->tracerTend = (double*)calloc((size_t)ntr * c->nz * c->ny * c->nx, sizeof(double));
  if (!c->tracerLimits || !c->tracerFlux || !c->tracerFac || !c->tracerTend) return -1;
  return 0;
}
int oracle_num_tracers(const oracle_ctx* c) { return c->ntr; }

/* tracer mass -> mixing ratio inside the domain (S:622-627, 983-989, 1342-1349) */
static void tracer_divide(oracle_ctx* c, const double* state, double* tracers) {
  int hs = c->hs;
  for (int t = 0; t < c->ntr; t++)
    for (int k = 0; k < c->nz; k++)
      for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nx; i++)
          TRC(t, hs + k, hs + j, hs + i) /= (ST(idR, hs + k, hs + j, hs + i) + c->hyDensCells[hs + k]);
}

/* mixing ratio -> tracer mass (S:952-954, 1306-1308, 1692-1694) */
static void tracer_multiply(oracle_ctx* c, const double* state, double* tracers) {
  int hs = c->hs;
  for (int t = 0; t < c->ntr; t++)
    for (int k = 0; k < c->nz; k++)
      for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nx; i++)
          TRC(t, hs + k, hs + j, hs + i) *= (ST(idR, hs + k, hs + j, hs + i) + c->hyDensCells[hs + k]);
}

/* halos of the mixing ratio along dir: periodic copy, or the end cell's value
 * at a wall (S:636-640, 655-660, 1000-1004, 1018-1022, 1357-1361, 1375-1379) */
static void tracer_halo(oracle_ctx* c, double* tracers, int dir) {
  int hs = c->hs, nx = c->nx, ny = c->ny, nz = c->nz;
  int bc = dir == 0 ? c->cfg.bc_x : dir == 1 ? c->cfg.bc_y : c->cfg.bc_z;
  for (int t = 0; t < c->ntr; t++)
    for (int a = 0; a < (dir == 2 ? ny : nz); a++)
      for (int b = 0; b < (dir == 0 ? ny : nx); b++)
        for (int h = 0; h < hs; h++) {
          if (dir == 0) {
            int k = a, j = b;
            TRC(t, hs + k, hs + j, h) = bc == BC_PERIODIC ? TRC(t, hs + k, hs + j, nx + h) : TRC(t, hs + k, hs + j, hs);
            TRC(t, hs + k, hs + j, hs + nx + h) =
                bc == BC_PERIODIC ? TRC(t, hs + k, hs + j, hs + h) : TRC(t, hs + k, hs + j, hs + nx - 1);
          } else if (dir == 1) {
            int k = a, i = b;
            TRC(t, hs + k, h, hs + i) = bc == BC_PERIODIC ? TRC(t, hs + k, ny + h, hs + i) : TRC(t, hs + k, hs, hs + i);
            TRC(t, hs + k, hs + ny + h, hs + i) =
                bc == BC_PERIODIC ? TRC(t, hs + k, hs + h, hs + i) : TRC(t, hs + k, hs + ny - 1, hs + i);
          } else {
            int j = a, i = b;
            TRC(t, h, hs + j, hs + i) = bc == BC_PERIODIC ? TRC(t, nz + h, hs + j, hs + i) : TRC(t, hs, hs + j, hs + i);
            TRC(t, hs + nz + h, hs + j, hs + i) =
                bc == BC_PERIODIC ? TRC(t, hs + h, hs + j, hs + i) : TRC(t, hs + nz - 1, hs + j, hs + i);
          }
        }
}

/* diffTransformTracer (S:2095-2126): rho*tracer rt and its flux rut from the
 * density r and the normal momentum ru at every level */
static void dt_tracer(const oracle_ctx* c, arr2 r, arr2 ru, arr2 rt, arr2 rut, double dx) {
  int nAder = c->nader, ngll = c->ngll;
  for (int kt = 1; kt < nAder; kt++)
    for (int ii = 0; ii < ngll; ii++) rut[kt][ii] = 0;
  for (int kt = 0; kt < nAder - 1; kt++) {
    for (int ii = 0; ii < ngll; ii++) {
      double df_dx = 0;
      for (int s = 0; s < ngll; s++) df_dx += c->deriv[s][ii] * rut[kt][s];
      rt[kt + 1][ii] = -df_dx / dx / (kt + 1.);
    }
    for (int ii = 0; ii < ngll; ii++) {
      double tot = 0;
      for (int ir = 0; ir <= kt + 1; ir++) tot += ru[ir][ii] * rt[kt + 1 - ir][ii] - r[ir][ii] * rut[kt + 1 - ir][ii];
      rut[kt + 1][ii] = tot / r[0][ii];
    }
  }
}

/* Edge values of rho*tracer for one cell (S:763-808, 1126-1170, 1503-1548):
 * stencil st of the mixing ratio, density r and normal momentum rn at every
 * transform level of the cell. */
static void tracer_edges(const oracle_ctx* c, const double* st, arr2 r, arr2 rn, double d, double dt, int tr,
                         double* left, double* right) {
  int ngll = c->ngll;
  arr2 rt, rnt;
  reconstruct(c, st, rt[0], c->cfg.weno_scalars);
  for (int ii = 0; ii < ngll; ii++) rt[0][ii] *= r[0][ii];
  if (c->tpos[tr])
    for (int ii = 0; ii < ngll; ii++) rt[0][ii] = dmax(0., rt[0][ii]);
  for (int ii = 0; ii < ngll; ii++) rnt[0][ii] = rt[0][ii] * rn[0][ii] / r[0][ii];
  if (c->nader > 1) dt_tracer(c, r, rn, rt, rnt, d);
  tavg_edges(c, rt, dt);
  if (c->tpos[tr]) {
    rt[0][0] = dmax(0., rt[0][0]);
    rt[0][ngll - 1] = dmax(0., rt[0][ngll - 1]);
  }
  *left = rt[0][0];
  *right = rt[0][ngll - 1];
}

/* stateLimits-style boundary values of the tracer edges (S:826-835, 1186-1195, 1567-1576) */
static void tracer_limits_bc(oracle_ctx* c, int dir) {
  int nx = c->nx, ny = c->ny, nz = c->nz;
  int bc = dir == 0 ? c->cfg.bc_x : dir == 1 ? c->cfg.bc_y : c->cfg.bc_z;
  for (int t = 0; t < c->ntr; t++)
    for (int a = 0; a < (dir == 2 ? ny : nz); a++)
      for (int b = 0; b < (dir == 0 ? ny : nx); b++) {
        int k0, j0, i0, k1, j1, i1;   /* face 0 and face n of the line */
        if (dir == 0) { k0 = k1 = a; j0 = j1 = b; i0 = 0; i1 = nx; }
        else if (dir == 1) { k0 = k1 = a; i0 = i1 = b; j0 = 0; j1 = ny; }
        else { j0 = j1 = a; i0 = i1 = b; k0 = 0; k1 = nz; }
        if (bc == BC_PERIODIC) {
          TLM(t, 0, k0, j0, i0) = TLM(t, 0, k1, j1, i1);
          TLM(t, 1, k1, j1, i1) = TLM(t, 1, k0, j0, i0);
        } else {
          TLM(t, 0, k0, j0, i0) = TLM(t, 1, k0, j0, i0);
          TLM(t, 1, k1, j1, i1) = TLM(t, 0, k1, j1, i1);
        }
      }
}

/* upwind tracer fluxes at one face (S:889-901, 1248-1259, 1632-1643): the
 * face's mass flux times rho*tracer / rho of the upwind edge; un is the
 * normal velocity the state flux upwinds with */
static void tracer_face(oracle_ctx* c, int k, int j, int i, double un, double massFlux, double q1_L, double q1_R) {
  for (int t = 0; t < c->ntr; t++) {
    if (un > 0) TFX(t, k, j, i) = massFlux * TLM(t, 0, k, j, i) / q1_L;
    else TFX(t, k, j, i) = massFlux * TLM(t, 1, k, j, i) / q1_R;
  }
}

/* positivity limiter along dir (S:904-935, 1262-1294, 1646-1678; T1-T3) */
static void tracer_limit(oracle_ctx* c, const double* state, const double* tracers, int dir, double d, double dt) {
  const double eps = 1.e-10;
  int hs = c->hs, nx = c->nx, ny = c->ny, nz = c->nz, NM = 1 + dir;
  int n = dir == 0 ? nx : dir == 1 ? ny : nz;
  int ek = dir == 2, ej = dir == 1, ei = dir == 0;
  for (int t = 0; t < c->ntr; t++) {
    if (!c->tpos[t]) continue;
    /* factor of every cell from its outflow through the unlimited fluxes */
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          double f1 = dmin(TFX(t, k, j, i), 0.);
          double f2 = dmax(TFX(t, k + ek, j + ej, i + ei), 0.);
          double fluxOut = dt * (f2 - f1) / d;
          double dens = ST(idR, hs + k, hs + j, hs + i) + c->hyDensCells[hs + k];
          c->tracerFac[(((size_t)t * (nz + 1) + k) * (ny + 1) + j) * (nx + 1) + i] =
              dmin(1., TRC(t, hs + k, hs + j, hs + i) * dens / (fluxOut + eps));
        }
    /* every face takes the factor of its upwind cell */
    for (int k = 0; k < nz + ek; k++)
      for (int j = 0; j < ny + ej; j++)
        for (int i = 0; i < nx + ei; i++) {
          double u = 0.5 * (SL(NM, 0, k, j, i) + SL(NM, 1, k, j, i));
          int f = dir == 0 ? i : dir == 1 ? j : k, a;
          if (u > 0) a = f == 0 ? n - 1 : f - 1;
          else if (u < 0) a = f == n ? 0 : f;
          else continue;
          int ka = dir == 2 ? a : k, ja = dir == 1 ? a : j, ia = dir == 0 ? a : i;
          TFX(t, k, j, i) *= c->tracerFac[(((size_t)t * (nz + 1) + ka) * (ny + 1) + ja) * (nx + 1) + ia];
        }
  }
}

/* tracer tendency along dir (S:948-955, 1302-1309, 1688-1695) */
static void tracer_tend(oracle_ctx* c, double* ttend, int dir, double d) {
  int ek = dir == 2, ej = dir == 1, ei = dir == 0;
  for (int t = 0; t < c->ntr; t++)
    for (int k = 0; k < c->nz; k++)
      for (int j = 0; j < c->ny; j++)
        for (int i = 0; i < c->nx; i++)
          TTD(t, k, j, i) = -(TFX(t, k + ek, j + ej, i + ei) - TFX(t, k, j, i)) / d;
}



static void tend_x(oracle_ctx* c, double* state, double* tend, double* tracers, double* ttend, double dt) {
  int hs = c->hs, nx = c->nx, ny = c->ny, nz = c->nz, ord = c->ord, ngll = c->ngll, nthr = nthreads_of(c);
  int ws = c->cfg.weno_scalars, wwd = c->cfg.weno_winds, bc = c->cfg.bc_x;
  double dx = c->dx;
  const int ntr = tracers ? c->ntr : 0;
  if (ntr) tracer_divide(c, state, tracers);   /* S:622-627 */
  /* S:631-662 */
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int ii = 0; ii < hs; ii++)
        for (int l = 0; l < NUMSTATE; l++) {
          if (bc == BC_PERIODIC) {
            ST(l, hs + k, hs + j, ii) = ST(l, hs + k, hs + j, nx + ii);
            ST(l, hs + k, hs + j, hs + nx + ii) = ST(l, hs + k, hs + j, hs + ii);
          } else if (l == idU) {
            ST(l, hs + k, hs + j, ii) = 0;
            ST(l, hs + k, hs + j, hs + nx + ii) = 0;
          } else {
            ST(l, hs + k, hs + j, ii) = ST(l, hs + k, hs + j, hs);
            ST(l, hs + k, hs + j, hs + nx + ii) = ST(l, hs + k, hs + j, hs + nx - 1);
          }
        }
  if (ntr) tracer_halo(c, tracers, 0);
  /* S:665-811 */
  OMP_FOR2
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        arr2 r, ru, rv, rw, rt, ruu, ruv, ruw, rut, rtg;
        double st[N9];
        for (int ii = 0; ii < ord; ii++) st[ii] = ST(idR, hs + k, hs + j, i + ii);
        reconstruct(c, st, r[0], ws);
        for (int ii = 0; ii < ngll; ii++) r[0][ii] += c->hyDensCells[hs + k];
        for (int ii = 0; ii < ord; ii++) st[ii] = ST(idU, hs + k, hs + j, i + ii);
        reconstruct(c, st, ru[0], wwd);
        for (int ii = 0; ii < ord; ii++) st[ii] = ST(idV, hs + k, hs + j, i + ii);
        reconstruct(c, st, rv[0], wwd);
        for (int ii = 0; ii < ord; ii++) st[ii] = ST(idW, hs + k, hs + j, i + ii);
        reconstruct(c, st, rw[0], wwd);
        for (int ii = 0; ii < ord; ii++) st[ii] = ST(idT, hs + k, hs + j, i + ii);
        reconstruct(c, st, rt[0], ws);
        for (int ii = 0; ii < ngll; ii++) rt[0][ii] += c->hyDensThetaCells[hs + k];
        for (int ii = 0; ii < ngll; ii++) {
          double rr = r[0][ii];
          double u = ru[0][ii] / rr, v = rv[0][ii] / rr, w = rw[0][ii] / rr, t = rt[0][ii] / rr;
          ruu[0][ii] = rr * u * u;
          ruv[0][ii] = rr * u * v;
          ruw[0][ii] = rr * u * w;
          rut[0][ii] = rr * u * t;
          rtg[0][ii] = spow(rr * t, GAMMA);
        }
        if (c->nader > 1) dt_xy(c, 0, r, ru, rv, rw, rt, ruu, ruv, ruw, rut, rtg, dx);
        double r_tavg[N9], ru_tavg[N9];
        tavg_edges_copy(c, r, r_tavg, dt);
        tavg_edges_copy(c, ru, ru_tavg, dt);
        tavg_edges(c, rv, dt);
        tavg_edges(c, rw, dt);
        tavg_edges(c, rt, dt);
        SL(idR, 1, k, j, i) = r_tavg[0];
        SL(idU, 1, k, j, i) = ru_tavg[0];
        SL(idV, 1, k, j, i) = rv[0][0];
        SL(idW, 1, k, j, i) = rw[0][0];
        SL(idT, 1, k, j, i) = rt[0][0];
        SL(idR, 0, k, j, i + 1) = r_tavg[ngll - 1];
        SL(idU, 0, k, j, i + 1) = ru_tavg[ngll - 1];
        SL(idV, 0, k, j, i + 1) = rv[0][ngll - 1];
        SL(idW, 0, k, j, i + 1) = rw[0][ngll - 1];
        SL(idT, 0, k, j, i + 1) = rt[0][ngll - 1];
        for (int t = 0; t < ntr; t++) {            /* S:760-809 */
          double tst[N9], tl, tr_;
          for (int ii = 0; ii < ord; ii++) tst[ii] = TRC(t, hs + k, hs + j, i + ii);
          tracer_edges(c, tst, r, ru, dx, dt, t, &tl, &tr_);
          TLM(t, 1, k, j, i) = tl;
          TLM(t, 0, k, j, i + 1) = tr_;
        }
      }
  /* S:816-837 */
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int l = 0; l < NUMSTATE; l++) {
        if (bc == BC_PERIODIC) {
          SL(l, 0, k, j, 0) = SL(l, 0, k, j, nx);
          SL(l, 1, k, j, nx) = SL(l, 1, k, j, 0);
        } else {
          SL(l, 0, k, j, 0) = SL(l, 1, k, j, 0);
          SL(l, 1, k, j, nx) = SL(l, 0, k, j, nx);
        }
      }
  if (ntr) tracer_limits_bc(c, 0);
  /* S:842-902 */
  OMP_FOR2
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx + 1; i++) {
        double L[5], R[5], f[5];
        for (int l = 0; l < 5; l++) { L[l] = SL(l, 0, k, j, i); R[l] = SL(l, 1, k, j, i); }
        riemann_x(L, R, f);
        for (int l = 0; l < 5; l++) SF(l, k, j, i) = f[l];
        if (ntr) tracer_face(c, k, j, i, 0.5 * (L[1] / L[0] + R[1] / R[0]), f[0], L[0], R[0]);
      }
  if (ntr) {
    tracer_limit(c, state, tracers, 0, dx, dt);
    tracer_tend(c, ttend, 0, dx);
    tracer_multiply(c, state, tracers);
  }
  /* S:940-956 */
  OMP_FOR2
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int l = 0; l < NUMSTATE; l++) {
          if (c->sim2d && l == idV) TD(l, k, j, i) = 0;
          else TD(l, k, j, i) = -(SF(l, k, j, i + 1) - SF(l, k, j, i)) / dx;
        }
}

/* ------------------------------------------------------------------------ */
/* computeTendenciesY (S:961-1311)                                          */
/* ------------------------------------------------------------------------ */
static void tend_y(oracle_ctx* c, double* state, double* tend, double* tracers, double* ttend, double dt) {
  int hs = c->hs, nx = c->nx, ny = c->ny, nz = c->nz, ord = c->ord, ngll = c->ngll, nthr = nthreads_of(c);
  int ws = c->cfg.weno_scalars, wwd = c->cfg.weno_winds, bc = c->cfg.bc_y;
  double dy = c->dy;
  const int ntr = tracers ? c->ntr : 0;
  if (ntr) tracer_divide(c, state, tracers);   /* S:983-989 */
  /* S:994-1025 */
  for (int k = 0; k < nz; k++)
    for (int i = 0; i < nx; i++)
      for (int jj = 0; jj < hs; jj++)
        for (int l = 0; l < NUMSTATE; l++) {
          if (bc == BC_PERIODIC) {
            ST(l, hs + k, jj, hs + i) = ST(l, hs + k, ny + jj, hs + i);
            ST(l, hs + k, hs + ny + jj, hs + i) = ST(l, hs + k, hs + jj, hs + i);
          } else if (l == idV) {
            ST(l, hs + k, jj, hs + i) = 0;
            ST(l, hs + k, hs + ny + jj, hs + i) = 0;
          } else {
            ST(l, hs + k, jj, hs + i) = ST(l, hs + k, hs, hs + i);
            ST(l, hs + k, hs + ny + jj, hs + i) = ST(l, hs + k, hs + ny - 1, hs + i);
          }
        }
  if (ntr) tracer_halo(c, tracers, 1);
  /* S:1028-1171 */
  OMP_FOR2
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        arr2 r, ru, rv, rw, rt, rvu, rvv, rvw, rvt, rtg;
        double st[N9];
        for (int jj = 0; jj < ord; jj++) st[jj] = ST(idR, hs + k, j + jj, hs + i);
        reconstruct(c, st, r[0], ws);
        for (int jj = 0; jj < ngll; jj++) r[0][jj] += c->hyDensCells[hs + k];
        for (int jj = 0; jj < ord; jj++) st[jj] = ST(idU, hs + k, j + jj, hs + i);
        reconstruct(c, st, ru[0], wwd);
        for (int jj = 0; jj < ord; jj++) st[jj] = ST(idV, hs + k, j + jj, hs + i);
        reconstruct(c, st, rv[0], wwd);
        for (int jj = 0; jj < ord; jj++) st[jj] = ST(idW, hs + k, j + jj, hs + i);
        reconstruct(c, st, rw[0], wwd);
        for (int jj = 0; jj < ord; jj++) st[jj] = ST(idT, hs + k, j + jj, hs + i);
        reconstruct(c, st, rt[0], ws);
        for (int jj = 0; jj < ngll; jj++) rt[0][jj] += c->hyDensThetaCells[hs + k];
        for (int jj = 0; jj < ngll; jj++) {
          double rr = r[0][jj];
          double u = ru[0][jj] / rr, v = rv[0][jj] / rr, w = rw[0][jj] / rr, t = rt[0][jj] / rr;
          rvu[0][jj] = rr * v * u;
          rvv[0][jj] = rr * v * v;
          rvw[0][jj] = rr * v * w;
          rvt[0][jj] = rr * v * t;
          rtg[0][jj] = spow(rr * t, GAMMA);
        }
        if (c->nader > 1) dt_xy(c, 1, r, ru, rv, rw, rt, rvu, rvv, rvw, rvt, rtg, dy);
        double r_tavg[N9], rv_tavg[N9];
        tavg_edges_copy(c, r, r_tavg, dt);
        tavg_edges(c, ru, dt);
        tavg_edges_copy(c, rv, rv_tavg, dt);
        tavg_edges(c, rw, dt);
        tavg_edges(c, rt, dt);
        SL(idR, 1, k, j, i) = r_tavg[0];
        SL(idU, 1, k, j, i) = ru[0][0];
        SL(idV, 1, k, j, i) = rv_tavg[0];
        SL(idW, 1, k, j, i) = rw[0][0];
        SL(idT, 1, k, j, i) = rt[0][0];
        SL(idR, 0, k, j + 1, i) = r_tavg[ngll - 1];
        SL(idU, 0, k, j + 1, i) = ru[0][ngll - 1];
        SL(idV, 0, k, j + 1, i) = rv_tavg[ngll - 1];
        SL(idW, 0, k, j + 1, i) = rw[0][ngll - 1];
        SL(idT, 0, k, j + 1, i) = rt[0][ngll - 1];
        for (int t = 0; t < ntr; t++) {            /* S:1123-1171 */
          double tst[N9], tl, tr_;
          for (int jj = 0; jj < ord; jj++) tst[jj] = TRC(t, hs + k, j + jj, hs + i);
          tracer_edges(c, tst, r, rv, dy, dt, t, &tl, &tr_);
          TLM(t, 1, k, j, i) = tl;
          TLM(t, 0, k, j + 1, i) = tr_;
        }
      }
  /* S:1176-1197 */
  for (int k = 0; k < nz; k++)
    for (int i = 0; i < nx; i++)
      for (int l = 0; l < NUMSTATE; l++) {
        if (bc == BC_PERIODIC) {
          SL(l, 0, k, 0, i) = SL(l, 0, k, ny, i);
          SL(l, 1, k, ny, i) = SL(l, 1, k, 0, i);
        } else {
          SL(l, 0, k, 0, i) = SL(l, 1, k, 0, i);
          SL(l, 1, k, ny, i) = SL(l, 0, k, ny, i);
        }
      }
  if (ntr) tracer_limits_bc(c, 1);
  /* S:1202-1260 */
  OMP_FOR2
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny + 1; j++)
      for (int i = 0; i < nx; i++) {
        double L[5], R[5], f[5];
        for (int l = 0; l < 5; l++) { L[l] = SL(l, 0, k, j, i); R[l] = SL(l, 1, k, j, i); }
        riemann_y(L, R, f);
        for (int l = 0; l < 5; l++) SF(l, k, j, i) = f[l];
        if (ntr) tracer_face(c, k, j, i, 0.5 * (L[2] + R[2]), f[0], L[0], R[0]);   /* v of S:1211 */
      }
  /* S:1298-1310 */
  OMP_FOR2
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int l = 0; l < NUMSTATE; l++) TD(l, k, j, i) = -(SF(l, k, j + 1, i) - SF(l, k, j, i)) / dy;
  if (ntr) {
    tracer_limit(c, state, tracers, 1, dy, dt);
    tracer_tend(c, ttend, 1, dy);
    tracer_multiply(c, state, tracers);
  }
}

/* ------------------------------------------------------------------------ */
/* computeTendenciesZ (S:1315-1697)                                         */
/* ------------------------------------------------------------------------ */
static void tend_z(oracle_ctx* c, double* state, double* tend, double* tracers, double* ttend, double dt) {
  int hs = c->hs, nx = c->nx, ny = c->ny, nz = c->nz, ord = c->ord, ngll = c->ngll, nthr = nthreads_of(c);
  int ws = c->cfg.weno_scalars, wwd = c->cfg.weno_winds, bc = c->cfg.bc_z;
  double dz = c->dz;
  const int ntr = tracers ? c->ntr : 0;
  if (ntr) tracer_divide(c, state, tracers);   /* S:1342-1349 */
  /* S:1351-1382 */
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      for (int kk = 0; kk < hs; kk++)
        for (int l = 0; l < NUMSTATE; l++) {
          if (bc == BC_PERIODIC) {
            ST(l, kk, hs + j, hs + i) = ST(l, nz + kk, hs + j, hs + i);
            ST(l, hs + nz + kk, hs + j, hs + i) = ST(l, hs + kk, hs + j, hs + i);
          } else if (l == idW) {
            ST(l, kk, hs + j, hs + i) = 0;
            ST(l, hs + nz + kk, hs + j, hs + i) = 0;
          } else {
            ST(l, kk, hs + j, hs + i) = ST(l, hs, hs + j, hs + i);
            ST(l, hs + nz + kk, hs + j, hs + i) = ST(l, hs + nz - 1, hs + j, hs + i);
          }
        }
  if (ntr) tracer_halo(c, tracers, 2);
  /* S:1385-1552 */
  OMP_FOR2
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        arr2 r, ru, rv, rw, rt, rwu, rwv, rww, rwt, rtg, src;
        double st[N9];
        for (int kk = 0; kk < ord; kk++) st[kk] = ST(idR, k + kk, hs + j, hs + i);
        reconstruct(c, st, r[0], ws);
        for (int kk = 0; kk < ngll; kk++) r[0][kk] += HYGLL(c->hyDensGLL, k, kk);
        for (int kk = 0; kk < ord; kk++) st[kk] = ST(idU, k + kk, hs + j, hs + i);
        reconstruct(c, st, ru[0], wwd);
        for (int kk = 0; kk < ord; kk++) st[kk] = ST(idV, k + kk, hs + j, hs + i);
        reconstruct(c, st, rv[0], wwd);
        for (int kk = 0; kk < ord; kk++) st[kk] = ST(idW, k + kk, hs + j, hs + i);
        reconstruct(c, st, rw[0], wwd);
        if (bc == BC_WALL) {
          if (k == nz - 1) rw[0][ngll - 1] = 0;
          if (k == 0) rw[0][0] = 0;
        }
        for (int kk = 0; kk < ord; kk++) st[kk] = ST(idT, k + kk, hs + j, hs + i);
        reconstruct(c, st, rt[0], ws);
        for (int kk = 0; kk < ngll; kk++) rt[0][kk] += HYGLL(c->hyDensThetaGLL, k, kk);
        for (int kk = 0; kk < ngll; kk++) {
          double rr = r[0][kk];
          double u = ru[0][kk] / rr, v = rv[0][kk] / rr, w = rw[0][kk] / rr, t = rt[0][kk] / rr;
          rwu[0][kk] = rr * w * u;
          rwv[0][kk] = rr * w * v;
          rww[0][kk] = rr * w * w;
          rwt[0][kk] = rr * w * t;
          rtg[0][kk] = spow(rr * t, GAMMA);
          /* S:1439: scalar assignment to the whole SArray (every entry) */
          double sv = -(r[0][kk] - HYGLL(c->hyDensGLL, k, kk)) * GRAV;
          for (int a = 0; a < c->nader; a++)
            for (int b = 0; b < ngll; b++) src[a][b] = sv;
        }
        if (c->nader > 1) dt_z(c, r, ru, rv, rw, rt, rwu, rwv, rww, rwt, rtg, src, k, dz);
        double r_tavg[N9], rw_tavg[N9];
        tavg_edges_copy(c, r, r_tavg, dt);
        tavg_edges(c, ru, dt);
        tavg_edges(c, rv, dt);
        tavg_edges_copy(c, rw, rw_tavg, dt);
        tavg_edges(c, rt, dt);
        tavg_all(c, src, dt);
        SL(idR, 1, k, j, i) = r_tavg[0];
        SL(idU, 1, k, j, i) = ru[0][0];
        SL(idV, 1, k, j, i) = rv[0][0];
        SL(idW, 1, k, j, i) = rw_tavg[0];
        SL(idT, 1, k, j, i) = rt[0][0];
        SL(idR, 0, k + 1, j, i) = r_tavg[ngll - 1];
        SL(idU, 0, k + 1, j, i) = ru[0][ngll - 1];
        SL(idV, 0, k + 1, j, i) = rv[0][ngll - 1];
        SL(idW, 0, k + 1, j, i) = rw_tavg[ngll - 1];
        SL(idT, 0, k + 1, j, i) = rt[0][ngll - 1];
        for (int t = 0; t < ntr; t++) {            /* S:1500-1548 */
          double tst[N9], tl, tr_;
          for (int kk = 0; kk < ord; kk++) tst[kk] = TRC(t, k + kk, hs + j, hs + i);
          tracer_edges(c, tst, r, rw, dz, dt, t, &tl, &tr_);
          TLM(t, 1, k, j, i) = tl;
          TLM(t, 0, k + 1, j, i) = tr_;
        }
        double src_avg = 0;
        for (int kk = 0; kk < ngll; kk++) src_avg += src[0][kk] * c->gllWts_ngll[kk];
        TD(idR, k, j, i) = 0;
        TD(idU, k, j, i) = 0;
        TD(idV, k, j, i) = 0;
        TD(idW, k, j, i) = src_avg;
        TD(idT, k, j, i) = 0;
      }
  /* S:1557-1578 */
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      for (int l = 0; l < NUMSTATE; l++) {
        if (bc == BC_PERIODIC) {
          SL(l, 0, 0, j, i) = SL(l, 0, nz, j, i);
          SL(l, 1, nz, j, i) = SL(l, 1, 0, j, i);
        } else {
          SL(l, 0, 0, j, i) = SL(l, 1, 0, j, i);
          SL(l, 1, nz, j, i) = SL(l, 0, nz, j, i);
        }
      }
  if (ntr) tracer_limits_bc(c, 2);
  /* S:1583-1644 */
  OMP_FOR2
  for (int k = 0; k < nz + 1; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double L[5], R[5], f[5];
        for (int l = 0; l < 5; l++) { L[l] = SL(l, 0, k, j, i); R[l] = SL(l, 1, k, j, i); }
        riemann_z(c, L, R, k, f);
        for (int l = 0; l < 5; l++) SF(l, k, j, i) = f[l];
        if (ntr) tracer_face(c, k, j, i, 0.5 * (L[3] / L[0] + R[3] / R[0]), f[0], L[0], R[0]);
      }
  if (ntr) {
    tracer_limit(c, state, tracers, 2, dz, dt);
    tracer_tend(c, ttend, 2, dz);
    tracer_multiply(c, state, tracers);
  }
  /* S:1680-1696 */
  OMP_FOR2
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int l = 0; l < NUMSTATE; l++) {
          if (c->sim2d && l == idV) TD(l, k, j, i) = 0;
          else TD(l, k, j, i) += -(SF(l, k + 1, j, i) - SF(l, k, j, i)) / dz;
        }
}

void oracle_tendencies_dir_tr(oracle_ctx* c, double* state, double* tend, double* tracers, double* ttend, double dt,
                              int dir) {
  if (dir == 0) tend_x(c, state, tend, tracers, ttend, dt);
  else if (dir == 1) {
    if (c->sim2d) {                                                                      /* S:572-575 */
      memset(tend, 0, sizeof(double) * NUMSTATE * c->nx * c->ny * c->nz);
      if (tracers && c->ntr) memset(ttend, 0, sizeof(double) * c->ntr * c->nx * c->ny * c->nz);
    } else {
      tend_y(c, state, tend, tracers, ttend, dt);
    }
  } else tend_z(c, state, tend, tracers, ttend, dt);
}
void oracle_tendencies_dir(oracle_ctx* c, double* state, double* tend, double dt, int dir) {
  oracle_tendencies_dir_tr(c, state, tend, 0, 0, dt, dir);
}

/* S:565-594 */
void oracle_compute_tendencies_tr(oracle_ctx* c, double* state, double* tend, double* tracers, double* ttend,
                                  double dt, int split) {
  static const int fwd[3] = {0, 1, 2}, bwd[3] = {2, 1, 0};
  int dir = c->dimSwitch ? fwd[split] : bwd[split];
  oracle_tendencies_dir_tr(c, state, tend, tracers, ttend, dt, dir);
  if (split == 2) c->dimSwitch = !c->dimSwitch;
}
void oracle_compute_tendencies(oracle_ctx* c, double* state, double* tend, double dt, int split) {
  oracle_compute_tendencies_tr(c, state, tend, 0, 0, dt, split);
}

/* Temporal_ader::timeStep (Temporal_ader.h:48-72), tracers (:67-69) when given */
void oracle_time_step_tr(oracle_ctx* c, double* state, double* tracers, double dt) {
  int hs = c->hs, nthr = nthreads_of(c);
  double* tend = c->tend;
  double* ttend = c->tracerTend;
  const int ntr = tracers ? c->ntr : 0;
  for (int spl = 0; spl < 3; spl++) {
    double dtloc = dt;
    oracle_compute_tendencies_tr(c, state, tend, ntr ? tracers : 0, ttend, dtloc, spl);
    OMP_FOR2
    for (int l = 0; l < NUMSTATE; l++)
      for (int k = 0; k < c->nz; k++)
        for (int j = 0; j < c->ny; j++)
          for (int i = 0; i < c->nx; i++) ST(l, hs + k, hs + j, hs + i) += dtloc * TD(l, k, j, i);
    for (int t = 0; t < ntr; t++)
      for (int k = 0; k < c->nz; k++)
        for (int j = 0; j < c->ny; j++)
          for (int i = 0; i < c->nx; i++) TRC(t, hs + k, hs + j, hs + i) += dtloc * TTD(t, k, j, i);
  }
}
void oracle_time_step(oracle_ctx* c, double* state, double dt) { oracle_time_step_tr(c, state, 0, dt); }

double oracle_run_steps(oracle_ctx* c, double* state, double dt, int nsteps) {
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int s = 0; s < nsteps; s++) oracle_time_step(c, state, dt);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
