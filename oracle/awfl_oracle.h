/* awfl_oracle.h -- CPU FP64 restatement of the reference ADER-WENO step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the
 * HIP path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker / the timed CPU baseline.  Nothing
 * in the product (awflcloud_amd/, include/) links or calls it.
 *
 * It restates, loop nest by loop nest and in the same floating-point operation
 * order, the FP64 path of the reference:
 *   Spatial_euler3d_cons_expl_cart_fv_Agrid.h (below "S:")  S:159-187, 274-315,
 *     370-485, 565-594, 598-1697, 1800-2177
 *   WenoLimiter.h:10-172, Temporal_ader.h:48-72, Profiles.h:8-73,
 *   phys_params.h:4-10, const.h:21-29.
 * The transform tables come from the generated awfl_tables.h; the composite
 * operators of Spatial::init (S:274-315) are re-formed here.
 */
#ifndef AWFL_ORACLE_H
#define AWFL_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXN 9
#define OR_MAXTR 8

typedef struct oracle_ctx oracle_ctx;

/* Grid/scheme description (mirrors the YAML keys of S:197-272). */
typedef struct {
  int nx, ny, nz;
  double xlen, ylen, zlen;
  int bc_x, bc_y, bc_z;          /* 0 periodic, 1 wall (S:75-76) */
  int weno_scalars, weno_winds;  /* S:223-224 */
  int ord, ngll;                 /* ORD, NGLL (const.h:21-27); nAder = ngll */
  int data_spec;                 /* background family: 1..3 const theta, 4 igw (S:78-81) */
  int nthreads;                  /* OpenMP threads; 0 = runtime default */
} oracle_config;

oracle_ctx* oracle_create(const oracle_config* cfg);
void oracle_destroy(oracle_ctx* c);
int oracle_hs(const oracle_ctx* c);

/* Tables as formed by Spatial::init, for tests. which: 0 wenoRecon(ord^3),
 * 1 coefs_to_gll(ord*ngll), 2 sten_to_gll(ord*ngll), 3 derivMatrix(ngll^2),
 * 4 idl(hs+2) then sigma, 5 gll pts ord, 6 gll wts ord, 7 gll pts ngll,
 * 8 gll wts ngll.  Writes row-major [in][out] and returns the count. */
int oracle_get_table(const oracle_ctx* c, int which, double* out);

/* Hydrostatic background (S:370-425). which: 0 hyDensCells, 1 hyPressureCells,
 * 2 hyThetaCells, 3 hyDensThetaCells (nz+2hs each), 4 hyDensGLL, 5 hyPressureGLL,
 * 6 hyThetaGLL, 7 hyDensThetaGLL (nz*ngll each). */
int oracle_get_background(const oracle_ctx* c, int which, double* out);

/* dt as computeTimeStep (S:159-187): computed from the state on the first call
 * while the cached value is <= 0, then cached. */
double oracle_compute_time_step(oracle_ctx* c, const double* state, double cfl);
void oracle_set_dt_cache(oracle_ctx* c, double dt);

/* computeTendencies(state, tend, dt, split) (S:565-594): fills state halos,
 * writes tend(5,nz,ny,nx) and toggles dimSwitch after the last split. */
void oracle_compute_tendencies(oracle_ctx* c, double* state, double* tend, double dt, int split);
/* One split's tendency for an explicit direction (0 x, 1 y, 2 z); no dimSwitch change. */
void oracle_tendencies_dir(oracle_ctx* c, double* state, double* tend, double dt, int dir);
/* Temporal_ader::timeStep (Temporal_ader.h:48-72), numTracers = 0. */
void oracle_time_step(oracle_ctx* c, double* state, double dt);
/* nsteps time steps; returns wall seconds of the stepping only. */
double oracle_run_steps(oracle_ctx* c, double* state, double dt, int nsteps);

/* Tracers (S:598-960 and the y/z analogues; compiled out in the reference by
 * NO_TRACERS, S:10): ntr <= OR_MAXTR tracer masses rho*q laid out
 * [ntr][nz+2hs][ny+2hs][nx+2hs], tendencies [ntr][nz][ny][nx]; pos[t] != 0
 * marks tracerPos (clipping and the positivity flux limiter).  Returns 0. */
int oracle_set_tracers(oracle_ctx* c, int ntr, const int* pos);
int oracle_num_tracers(const oracle_ctx* c);
void oracle_compute_tendencies_tr(oracle_ctx* c, double* state, double* tend, double* tracers, double* ttend,
                                  double dt, int split);
void oracle_tendencies_dir_tr(oracle_ctx* c, double* state, double* tend, double* tracers, double* ttend, double dt,
                              int dir);
/* initTracers (S:505-560): the reference's three tracers (uniform, theta
 * replica, block) times the constant-theta density; later tracers zero. */
void oracle_init_tracers(const oracle_ctx* c, double* tracers);
/* Temporal_ader::timeStep with tracers (Temporal_ader.h:48-72). */
void oracle_time_step_tr(oracle_ctx* c, double* state, double* tracers, double dt);

int oracle_get_dim_switch(const oracle_ctx* c);
void oracle_set_dim_switch(oracle_ctx* c, int v);

/* ---- per-function entry points for unit tests ---- */
/* weno::compute_weno_coefs (WenoLimiter.h:93-172) on one stencil u[ord] -> aw[ord]. */
void oracle_weno_coefs(const oracle_ctx* c, const double* u, double* aw);
/* reconstruct_gll_values (S:1800-1833): u[ord] -> gll[ngll]. */
void oracle_reconstruct(const oracle_ctx* c, const double* u, int doweno, double* gll);
/* diffTransformEulerCons{X,Y,Z} (S:1837-2091) on arrays laid out [var][kt][ii]
 * with var = r, ru, rv, rw, rt, f1, f2, f3, f4, rt_gamma (, source for z).
 * Level-0 entries must be set; returns all levels. For z, k_cell and the
 * background pressure row hyPressureGLL(k,:) are taken from the context. */
void oracle_diff_transform(const oracle_ctx* c, int dir, double* arrs, double dx, int k_cell);
/* Riemann flux at one face (S:842-887, 1202-1245, 1583-1629), qL/qR[5] -> f[5].
 * For z, kface selects the background pressure subtraction (S:1628-1629). */
void oracle_riemann(const oracle_ctx* c, int dir, const double* qL, const double* qR, int kface, double* f);

/* ---- initial data (S:428-485 plus the synthetic fields of the benchmarks) ---- */
typedef struct {
  int nbub;                      /* number of ellipsoids (<= 4) */
  int shape[4];                  /* 0 linear cone (Profiles.h:49-58), 1 cosine (Profiles.h:60-73),
                                    2 igw profile (Profiles.h:76-82) with hr = xr = yr */
  double x0[4], y0[4], z0[4], xr[4], yr[4], zr[4], amp[4], pwr[4];
  double noise_amp;              /* uniform per-cell theta' noise in [-amp, amp) */
  uint64_t noise_seed;
  int nmodes;                    /* band-limited theta' field evaluated at cell centres */
  const double* modes;           /* nmodes x (kx, ky, kz, amp, phase) */
  double u0, v0;                 /* uniform background wind (m/s), added as rho_h*u0 */
  double y2d;                    /* y of the quadrature points when ny == 1 (0: ylen/2) */
  double perturb_amp;            /* (rho theta)' *= 1 + amp * U[0,1) (S:488-499, seeded) */
  uint64_t perturb_seed;
} oracle_init_spec;
void oracle_init_state(const oracle_ctx* c, const oracle_init_spec* s, double* state);

/* Background point values at height z: family 0 constant theta
 * (Profiles.h:8-20), 1 constant Brunt-Vaisala frequency (Profiles.h:28-47),
 * both with t0 = 300 K, bvf = 0.01 1/s as initState uses them (S:378-391). */
void oracle_profile(int family, double z, double* th, double* rh, double* ph);

/* Sensitivity probe: seed != 0 moves every pow() of the step by one ulp
 * (direction from a hash of the argument); 0 restores plain pow(). */
void oracle_set_pow_jitter(uint64_t seed);
/* 0: the C library's pow (default); 1: x^GAMMA by the double-double method of
 * the HIP parity builds (attribution of their differences only) */
void oracle_set_pow_mode(int mode);
/* x^GAMMA as the step evaluates it under the current mode and jitter */
double oracle_eos_pow(double x);

/* Counter-based uniform in [-1,1): SplitMix64 of (seed, index). */
double oracle_uniform_pm1(uint64_t seed, uint64_t idx);

#ifdef __cplusplus
}
#endif
#endif
