// opcount.cpp -- op-counting build of the CPU oracle (test infrastructure only).
//
// Compiles awfl_oracle.c as C++ with every `double` replaced by a scalar that
// counts its floating-point operations, runs one split sweep of each
// direction (oracle_tendencies_dir: halo fill, reconstruction, differential
// transform, time averages, edge boundary, Riemann flux, tendency -- S:598-1697
// with S = Spatial_euler3d_cons_expl_cart_fv_Agrid.h) and the tendency
// application of Temporal_ader.h:63-65, and prints the operations per cell.
// The oracle follows the reference's operation order statement by statement
// (and, like the reference CPU build, contracts nothing into FMAs), so these
// are the reference's own operation counts.
//
//   make -C oracle opcount && oracle/opcount [nx ny nz [ord ngll]]
//
// Counted: add/sub, mul, div, sqrt, pow, other libm calls (cos, sin, exp,
// fabs, fmax, fmin).  Flops = add + mul + div + sqrt + pow (each 1), the
// convention of the survey's tally with an FMA counted as 2.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <cmath>
#include <type_traits>

namespace opc {
unsigned long long n_add, n_mul, n_div, n_sqrt, n_pow, n_other;
void reset() { n_add = n_mul = n_div = n_sqrt = n_pow = n_other = 0; }
}  // namespace opc

struct CD {
  double v;
  CD() = default;
  constexpr CD(const CD&) = default;
  template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
  constexpr CD(T x) : v((double)x) {}
  template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
  constexpr explicit operator T() const { return (T)v; }
  CD& operator+=(CD o) { ++opc::n_add; v += o.v; return *this; }
  CD& operator-=(CD o) { ++opc::n_add; v -= o.v; return *this; }
  CD& operator*=(CD o) { ++opc::n_mul; v *= o.v; return *this; }
  CD& operator/=(CD o) { ++opc::n_div; v /= o.v; return *this; }
};
static_assert(sizeof(CD) == sizeof(double), "CD must have the layout of a double");
inline CD operator+(CD a, CD b) { ++opc::n_add; return CD(a.v + b.v); }
inline CD operator-(CD a, CD b) { ++opc::n_add; return CD(a.v - b.v); }
inline CD operator*(CD a, CD b) { ++opc::n_mul; return CD(a.v * b.v); }
inline CD operator/(CD a, CD b) { ++opc::n_div; return CD(a.v / b.v); }
inline CD operator-(CD a) { return CD(-a.v); }
inline CD operator+(CD a) { return a; }
inline bool operator<(CD a, CD b) { return a.v < b.v; }
inline bool operator>(CD a, CD b) { return a.v > b.v; }
inline bool operator<=(CD a, CD b) { return a.v <= b.v; }
inline bool operator>=(CD a, CD b) { return a.v >= b.v; }
inline bool operator==(CD a, CD b) { return a.v == b.v; }
inline bool operator!=(CD a, CD b) { return a.v != b.v; }
inline CD sqrt(CD a) { ++opc::n_sqrt; return CD(std::sqrt(a.v)); }
inline CD pow(CD a, CD b) { ++opc::n_pow; return CD(std::pow(a.v, b.v)); }
inline CD fabs(CD a) { ++opc::n_other; return CD(std::fabs(a.v)); }
inline CD fmax(CD a, CD b) { ++opc::n_other; return CD(std::fmax(a.v, b.v)); }
inline CD fmin(CD a, CD b) { ++opc::n_other; return CD(std::fmin(a.v, b.v)); }
inline CD cos(CD a) { ++opc::n_other; return CD(std::cos(a.v)); }
inline CD sin(CD a) { ++opc::n_other; return CD(std::sin(a.v)); }
inline CD exp(CD a) { ++opc::n_other; return CD(std::exp(a.v)); }
inline CD nextafter(CD a, CD b) { return CD(std::nextafter(a.v, b.v)); }

#define double CD
#include "awfl_oracle.c"
#undef double

static void count_one(int ord, int ngll, int nx, int ny, int nz) {
  oracle_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.nx = nx; cfg.ny = ny; cfg.nz = nz;
  cfg.xlen = 20000.0; cfg.ylen = 20000.0; cfg.zlen = 10000.0;
  cfg.bc_x = BC_PERIODIC; cfg.bc_y = BC_PERIODIC; cfg.bc_z = BC_WALL;
  cfg.weno_scalars = 1; cfg.weno_winds = 1;
  cfg.ord = ord; cfg.ngll = ngll; cfg.data_spec = 1; cfg.nthreads = 1;
  oracle_ctx* c = oracle_create(&cfg);
  if (!c) {
    printf("oracle_create failed for ord %d ngll %d\n", ord, ngll);
    exit(1);
  }
  const int hs = oracle_hs(c);
  const size_t ns = (size_t)5 * (nz + 2 * hs) * (ny + 2 * hs) * (nx + 2 * hs);
  const size_t nt = (size_t)5 * nz * ny * nx;
  CD* st = (CD*)calloc(ns, sizeof(CD));
  CD* tend = (CD*)calloc(nt, sizeof(CD));
  // thermal bubble with seeded noise and a uniform wind: every branch of the
  // sweep takes its general path
  oracle_init_spec spec;
  memset(&spec, 0, sizeof spec);
  spec.nbub = 1;
  spec.shape[0] = 1;
  spec.x0[0] = 10000.0; spec.y0[0] = 10000.0; spec.z0[0] = 2000.0;
  spec.xr[0] = 2000.0; spec.yr[0] = 2000.0; spec.zr[0] = 2000.0;
  spec.amp[0] = 2.0; spec.pwr[0] = 1.0;
  spec.noise_amp = 1e-3; spec.noise_seed = 1;
  spec.u0 = 3.0;
  oracle_init_state(c, &spec, st);
  const double ncell = (double)nx * ny * nz;
  const CD dt = oracle_compute_time_step(c, st, 0.8);
  double tot[6] = {0, 0, 0, 0, 0, 0};
  printf("ord %d ngll %d, %dx%dx%d (periodic x/y, wall z): operations per cell\n", ord, ngll, nx, ny, nz);
  printf("  %-10s %9s %9s %9s %6s %6s %7s %9s\n", "sweep", "add/sub", "mul", "div", "sqrt", "pow", "other", "flops");
  for (int dir = 0; dir < 3; dir++) {
    opc::reset();
    oracle_tendencies_dir(c, st, tend, dt, dir);
    const double a = opc::n_add / ncell, m = opc::n_mul / ncell, d = opc::n_div / ncell;
    const double q = opc::n_sqrt / ncell, p = opc::n_pow / ncell, o = opc::n_other / ncell;
    printf("  %-10s %9.1f %9.1f %9.1f %6.2f %6.2f %7.2f %9.1f\n", dir == 0 ? "x" : dir == 1 ? "y" : "z", a, m, d,
           q, p, o, a + m + d + q + p);
    tot[0] += a; tot[1] += m; tot[2] += d; tot[3] += q; tot[4] += p; tot[5] += o;
  }
  // tendency application (Temporal_ader.h:63-65): state += dt * tend, per split
  const double apply = 2.0 * 5.0;
  printf("  %-10s %9.1f %9.1f %9.1f %6.2f %6.2f %7.2f %9.1f\n", "mean sweep", tot[0] / 3, tot[1] / 3, tot[2] / 3,
         tot[3] / 3, tot[4] / 3, tot[5] / 3, (tot[0] + tot[1] + tot[2] + tot[3] + tot[4]) / 3);
  printf("  step (3 sweeps + 3 applications of %g flops): %.1f flops per cell\n", apply,
         tot[0] + tot[1] + tot[2] + tot[3] + tot[4] + 3 * apply);
  free(st);
  free(tend);
  oracle_destroy(c);
}

int main(int argc, char** argv) {
  int nx = 64, ny = 64, nz = 32;
  if (argc >= 4) {
    nx = atoi(argv[1]); ny = atoi(argv[2]); nz = atoi(argv[3]);
  }
  if (argc == 6) {                     // one (ord, ngll) only
    count_one(atoi(argv[4]), atoi(argv[5]), nx, ny, nz);
    return 0;
  }
  count_one(5, 3, nx, ny, nz);
  count_one(9, 5, nx, ny, nz);
  count_one(3, 2, nx, ny, nz);
  return 0;
}
