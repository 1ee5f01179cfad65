/* awfl_dycore.h -- C ABI of the MI355X-native ADER-WENO dynamical-core step.
 *
 * Drop-in boundary for the Spatial/Temporal concept of the reference
 * (Spatial_euler3d_cons_expl_cart_fv_Agrid.h, below "S:"; Temporal_ader.h,
 * below "T:").  Plain C types only: device pointers are `double*` allocated
 * by the caller (hipMalloc or any HIP allocator), sizes are int64.  Every
 * entry point returns 0 on success and a negative AWFL_E_* code on failure;
 * awfl_last_error() gives the message (the reference throws via endrun,
 * const.h:76-79, a C++ adapter rethrows).
 *
 * Threading: one context = one HIP stream; a context is not thread-safe
 * (like the reference's single-threaded Spatial object).  Calls are
 * asynchronous on the context's stream except awfl_compute_time_step and
 * the host-copy helpers, which synchronise (as ParallelMin did, S:182-183).
 *
 * Layout (S:34-38, 129-143): state(l,k,j,i) = [5][nz+2hs][ny+2hs][nx+2hs],
 * x fastest, l = rho', rho*u, rho*v, rho*w, (rho*theta)' (S:69-73); the
 * tendency is [5][nz][ny][nx].  hs = (ord-1)/2.  The halo cells of the
 * caller's state are neither read nor written: boundary values (periodic wrap
 * or solid wall, S:631-662, 994-1025, 1351-1382) are generated on the fly.
 */
#ifndef AWFL_DYCORE_H
#define AWFL_DYCORE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define AWFL_OK 0
#define AWFL_E_ARG -1       /* invalid argument / configuration              */
#define AWFL_E_HIP -2       /* HIP runtime error                             */
#define AWFL_E_STATE -3     /* call out of sequence                          */
#define AWFL_E_UNSUPPORTED -4 /* (ord, ngll) or precision not compiled in    */

#define AWFL_BC_PERIODIC 0  /* S:75 */
#define AWFL_BC_WALL 1      /* S:76 */

#define AWFL_FP64 0         /* plain FP64 state                              */
#define AWFL_FP32_PERT 1    /* FP32 perturbation state + Kahan compensation  */

typedef struct awfl_ctx awfl_ctx;

/* Configuration: the YAML keys of Spatial::init (S:197-272) plus the
 * compile-time ORD/NGLL of const.h:21-27 (nAder = ngll, T:8). */
typedef struct {
  int nx, ny, nz;
  double xlen, ylen, zlen;
  int bc_x, bc_y, bc_z;
  int weno_scalars, weno_winds;
  int ord, ngll;
  int data_spec;        /* 1 thermal, 2 collision, 3 straka (constant theta0=300 K
                           background), 4 igw (constant-N background), S:78-81 */
  int precision;        /* AWFL_FP64 or AWFL_FP32_PERT */
} awfl_config;

/* Synthetic initial data: sum of ellipsoids (Profiles.h:49-73) integrated with
 * the ord^3 GLL quadrature of initState (S:428-485), plus per-cell uniform
 * theta' noise from a counter-based generator and an optional band-limited
 * theta' field evaluated at cell centres. */
typedef struct {
  int nbub;
  int shape[4];                       /* 0 linear cone, 1 cosine^pwr, 2 igw profile
                                         amp sin(pi z/zlen)/sqrt(1+xn^2+yn^2) with
                                         xn = (x-x0)/xr, yn = (y-y0)/yr (Profiles.h:76-82) */
  double x0[4], y0[4], z0[4], xr[4], yr[4], zr[4], amp[4], pwr[4];
  double noise_amp;                   /* theta' noise uniform in [-amp, amp) K */
  uint64_t noise_seed;
  int nmodes;                         /* modes: host array nmodes x (kx,ky,kz,amp,phase) */
  const double* modes;
  double u0, v0;                      /* uniform wind added as rho_h*u0 (igw uses 20) */
  double y2d;                         /* y of every quadrature point when ny == 1
                                         (0: ylen/2, S:437-441; igw uses ylen/3, S:469) */
  double perturb_amp;                 /* (rho theta)' *= 1 + perturb_amp * U[0,1) per cell */
  uint64_t perturb_seed;              /* (S:488-499 with a seeded generator) */
} awfl_init_spec;

/* Spatial::init (S:192-340): builds the transform tables (S:274-315) and the
 * FP64 hydrostatic background (S:370-425) on the device. */
int awfl_create(const awfl_config* cfg, awfl_ctx** out);
/* Spatial::finalize (S:1795) + destructor. */
void awfl_destroy(awfl_ctx* ctx);
/* Message of the last failed call on ctx (or of the last failed awfl_create). */
const char* awfl_last_error(const awfl_ctx* ctx);

/* hs and the dimensions of StateArr / StateTendArr (S:23, 129-143). */
int awfl_layout(const awfl_ctx* ctx, int* hs, int64_t dims_state[4], int64_t dims_tend[4]);
/* Spatial::numSplit (S:153-155): always 3. */
int awfl_num_split(const awfl_ctx* ctx);
/* Stream the context launches on (hipStream_t as void*; NULL = null stream). */
int awfl_set_stream(awfl_ctx* ctx, void* hip_stream);

/* Spatial::initState (S:345-500) for the synthetic specs above. */
int awfl_init_state(awfl_ctx* ctx, double* d_state, const awfl_init_spec* spec);

/* Spatial::computeTimeStep (S:159-187): on the first call (cached dt <= 0)
 * reduces the CFL time step over the state on the device and caches it; later
 * calls return the cached value.  Synchronises the context's stream. */
int awfl_compute_time_step(awfl_ctx* ctx, const double* d_state, double cfl, double* dt);
/* Reset / override the cached time step (checkpoint restart). */
int awfl_set_time_step(awfl_ctx* ctx, double dt);

/* Spatial::computeTendencies (S:565-594): tendency of split `split` (0..2),
 * direction chosen by dimSwitch (x,y,z when set, z,y,x otherwise); toggles
 * dimSwitch after split 2.  d_state is read only. */
int awfl_compute_tendencies(awfl_ctx* ctx, const double* d_state, double* d_tend, double dt, int split);
/* The applySingleTendency contract of Temporal_ader (T:63-65, README.md:19,27,103):
 * state(l,hs+k,hs+j,hs+i) += dt * tend(l,k,j,i) over the interior. */
int awfl_apply_tendencies(awfl_ctx* ctx, double* d_state, const double* d_tend, double dt, int split);

/* Temporal_ader::timeStep (T:48-72) fused: three split sweeps, each one
 * kernel that reconstructs, evolves, fluxes and applies the tendency.  The
 * result is left in d_state.  Uses two context-owned state-sized buffers. */
int awfl_time_step(awfl_ctx* ctx, double* d_state, double dt);
/* awfl_time_step with the upwind branch of every face recorded (mode 1) or
 * imposed (mode 2).  The Riemann flux takes waves 1-3 from the left state
 * when the averaged normal velocity is > 0 (S:863, 1221, 1602); mode 1 runs
 * the step as awfl_time_step does and stores that choice per face, mode 2
 * uses the stored choice instead of the sign test, which makes the step a
 * smooth map near the recorded run (for difference checks of the adjoint).
 * d_branch holds 3 * (nx+1)(ny+1)(nz+1) bytes, one block per split in
 * order.  Diagnostic, no counterpart in the reference; (ord, ngll) = (5, 3)
 * only. */
int awfl_time_step_branches(awfl_ctx* ctx, double* d_state, double dt, unsigned char* d_branch, int mode);
/* Diagnostic timing of awfl_time_step, no counterpart in the reference (the
 * reference driver times whole steps with std::chrono, driver.cpp:49-54):
 * with max_sweeps > 0, the next max_sweeps sweep launches of awfl_time_step
 * are bracketed by HIP events recorded on the context's stream (creates the
 * events; 0 frees them and stops).  awfl_sweep_times waits for the last
 * recorded event and returns, per direction (x, y, z), the mean duration in
 * ms and the number of timed sweeps since the previous call. */
int awfl_sweep_timing(awfl_ctx* ctx, int max_sweeps);
int awfl_sweep_times(awfl_ctx* ctx, double ms[3], int64_t count[3]);
/* awfl_time_step that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and the same result, bit for bit, as awfl_time_step; its one
 * state-sized intermediate lives in a file-scope device array of
 * AWFL_NOALLOC_MAX_BYTES that each device holds from the library's first use
 * on it until the process ends (the x sweep runs in place, so one
 * intermediate suffices).  States larger than that are rejected with
 * AWFL_E_ARG.  All contexts on one device share that array: calls that
 * overlap in time on one device (other streams, other threads) race; calls
 * in sequence do not.  The context's device must be current. */
#define AWFL_NOALLOC_MAX_BYTES (64ll * 1024 * 1024)
int awfl_time_step_noalloc(awfl_ctx* ctx, double* d_state, double dt);
/* One fused split sweep in direction dir (0 x, 1 y, 2 z): d_out = d_in + dt*tend.
 * d_in and d_out must not alias. */
int awfl_sweep(awfl_ctx* ctx, const double* d_in, double* d_out, double dt, int dir);

/* Alternative time integrators over the same spatial operator, behind the
 * Temporal interface (qjrmsCode/TimeIntegrator.cpp; not used by the Agrid
 * driver).  awfl_time_step_strang: x and y sweeps with dt, then z as two
 * sweeps with dt/2 (:32-65), the reverse order on alternate steps (dimSwitch).
 * awfl_time_step_ssprk3: three-stage SSP Runge-Kutta (:68-110) on the
 * semi-discrete operator, the sum of the directional tendencies without the
 * ADER time average.  Both in place on d_state; context-owned buffers. */
int awfl_time_step_strang(awfl_ctx* ctx, double* d_state, double dt);
int awfl_time_step_ssprk3(awfl_ctx* ctx, double* d_state, double dt);

/* dimSwitch of S:93 (checkpoint / restart). */
int awfl_get_dim_switch(const awfl_ctx* ctx);
int awfl_set_dim_switch(awfl_ctx* ctx, int v);

/* Hydrostatic background (S:45-52) copied to host: which = 0 hyDensCells,
 * 1 hyPressureCells, 2 hyThetaCells, 3 hyDensThetaCells (nz+2hs values),
 * 4 hyDensGLL, 5 hyPressureGLL, 6 hyThetaGLL, 7 hyDensThetaGLL (nz*ngll).
 * Returns the count written (>= 0) or an error code. */
int awfl_get_background(const awfl_ctx* ctx, int which, double* h_out, int64_t cap);

/* Equation of state p = C0 (rho*theta)^GAMMA (S:855, 1776) element-wise on
 * the device, with the kernels' own power routine: d_p[i] for d_rt[i] > 0. */
int awfl_eos_pressure(awfl_ctx* ctx, const double* d_rt, double* d_p, int64_t n);

/* The six output fields of Spatial::output (S:1752-1785) on the device, into
 * d_out[6][nz][ny][nx]: dens_pert, u, v, w, pot_temp_pert, pressure_pert. */
int awfl_diagnostics(awfl_ctx* ctx, const double* d_state, double* d_out);

/* Global sums over the interior of each state variable, in FP64 on the device
 * (conservation diagnostics).  h_out[5]. Synchronises. */
int awfl_state_sums(awfl_ctx* ctx, const double* d_state, double* h_out);

/* ---- FP32 perturbation form -------------------------------------------------
 * The state in FP32: two arrays (s, c) in the FP64 state's padded layout, the
 * value held being s - c (c is the Kahan compensation of the updates).  The
 * numerics are those of the FP64 split sweeps (S:1004-1707, 1713-2197 driven
 * by T:42-67) rewritten on deviations from the hydrostatic background, FP32
 * throughout on the device.  Available for (ord, ngll) = (3, 2) and (5, 3);
 * usable on any context whose (ord, ngll) has them (a context created with
 * precision AWFL_FP32_PERT checks that at creation).  No counterpart in the
 * reference, which has only FP64 (const.h:14 typedef double real). */

/* FP64 state -> (s, c) with s = (float)x, c = (float)(s - x), so that s - c
 * carries x to about 2^-48 relative. */
int awfl_state_to_f32(awfl_ctx* ctx, const double* d_state, float* d_s, float* d_c);

/* (s, c) -> FP64 state x = (double)s - (double)c. */
int awfl_state_from_f32(awfl_ctx* ctx, const float* d_s, const float* d_c, double* d_state);

/* One FP32 split sweep out of place: (s_out, c_out) = (s_in, c_in) after
 * the dir sweep with time step dt, the analogue of awfl_sweep. */
int awfl_sweep_f32(awfl_ctx* ctx, const float* d_s_in, const float* d_c_in, float* d_s_out, float* d_c_out,
                   double dt, int dir);

/* One FP32 time step in place on (s, c): the three splits in the order of
 * awfl_time_step, sharing the context's dimSwitch (T:42-67). */
int awfl_time_step_f32(awfl_ctx* ctx, float* d_s, float* d_c, double dt);

/* ---- tangent-linear and discrete adjoint (FP64) -----------------------------
 * Derivatives of the FP64 step with respect to the state, dt held fixed (the
 * reference caches dt after the first computeTimeStep, S:166).  No
 * counterpart in the reference.  Available for (ord, ngll) = (3, 2), (5, 3).
 * Adjoint arrays use the state's padded layout; halo entries of outputs are
 * zero and halo entries of inputs are ignored. */

/* Tangent-linear of awfl_sweep: d_out = sweep(d_in), d_dout = J(d_in) d_din. */
int awfl_sweep_tangent(awfl_ctx* ctx, const double* d_in, const double* d_din, double* d_out, double* d_dout,
                       double dt, int dir);

/* Adjoint of awfl_sweep: d_adj_in = J(d_in)^T d_adj_out (out of place). */
int awfl_sweep_adjoint(awfl_ctx* ctx, const double* d_in, const double* d_adj_out, double* d_adj_in, double dt,
                       int dir);

/* Adjoint of one awfl_time_step taken from d_state_before with the context's
 * dimSwitch equal to dim_switch: d_adj <- (d step / d state)^T d_adj, in
 * place.  Recomputes the step's split inputs. */
int awfl_time_step_adjoint(awfl_ctx* ctx, const double* d_state_before, double* d_adj, double dt, int dim_switch);

/* nsteps awfl_time_step calls in place on d_state, copying the state before
 * steps 0, every, 2*every, ... into consecutive state-sized slots of d_ckpt
 * (ceil(nsteps/every) slots).  Read awfl_get_dim_switch first: the adjoint
 * run needs the value it had before the first step.  The last two segments
 * run as their split sweeps and keep their inputs in the context (device
 * memory for 2*every*3 states, the two buffers awfl_adjoint_ckpt pipelines
 * through), which the next awfl_adjoint_ckpt on the same d_ckpt with the
 * same dt, nsteps, every and dim_switch0 uses instead of recomputing them;
 * the state it produces is the same, bit for bit. */
int awfl_forward_ckpt(awfl_ctx* ctx, double* d_state, double dt, int nsteps, int every, double* d_ckpt);

/* Adjoint of the run of awfl_forward_ckpt: d_adj enters as dJ/d(final
 * state) and leaves as dJ/d(initial state).  Segment by segment, last first,
 * the split inputs of the segment's steps are recomputed from its checkpoint
 * and the split adjoints applied in reverse order.  Two buffers of every*3
 * states alternate: the next segment is recomputed into one on a second
 * stream of the context (default priority, ordered by events) while the
 * adjoints of the current one run on the context's stream.  The inputs of
 * the last two segments are taken as awfl_forward_ckpt left them when it ran
 * on the same checkpoints and arguments, on the same stream, immediately
 * before (the checkpoint contents must be unchanged since).  They serve one
 * adjoint: any later adjoint recomputes them. */
int awfl_adjoint_ckpt(awfl_ctx* ctx, const double* d_ckpt, double dt, int nsteps, int every, int dim_switch0,
                      double* d_adj);

/* ---- tracers ----
 * The reference's TracerArr / TracerTendArr and numTracers / tracerPos
 * (S:331-339, S:565 computeTendencies(state, stateTend, tracers, tracerTend,
 * dt, split), Temporal_ader.h:67-69), whose kernels it compiles out with
 * #define NO_TRACERS (S:10); driver.cpp:27-29 allocates 3 tracers.  Tracer
 * masses rho*q are laid out like the state, [ntr][nz+2hs][ny+2hs][nx+2hs];
 * tracer tendencies [ntr][nz][ny][nx].  FP64 state only. */
#define AWFL_MAX_TRACERS 8

/* Number of tracers (0..AWFL_MAX_TRACERS) and, per tracer, the tracerPos flag
 * (h_positive[t] != 0): clipping of its reconstructed values at zero and the
 * positivity flux limiter (S:904-935). */
int awfl_set_tracers(awfl_ctx* ctx, int num_tracers, const int* h_positive);
int awfl_num_tracers(const awfl_ctx* ctx);

/* Spatial::initTracers (S:505-560): tracer 0 uniform, 1 a replica of the
 * thermal's theta bubble, 2 a block in the domain centre (driver.cpp:27-29
 * names them), each times the constant-theta density; later tracers zero,
 * halos zero. */
int awfl_init_tracers(awfl_ctx* ctx, double* d_tracers);

/* computeTendencies with tracers (S:565-594 with the tracer blocks of
 * S:598-960 and their y/z analogues): state tendency as
 * awfl_compute_tendencies, tracer tendency beside it; toggles dimSwitch after
 * split 2.  d_tracers is read only (the reference divides it by the density
 * and multiplies it back in place). */
int awfl_compute_tendencies_tracers(awfl_ctx* ctx, const double* d_state, const double* d_tracers, double* d_tend,
                                    double* d_tracer_tend, double dt, int split);

/* tracers += dt * tracer_tend over the interior (Temporal_ader.h:67-69). */
int awfl_apply_tracer_tendencies(awfl_ctx* ctx, double* d_tracers, const double* d_tracer_tend, double dt);

/* One split sweep of state and tracers, out of place: the state output is
 * awfl_sweep's (the same statements in another kernel; equal to rounding). */
int awfl_sweep_tracers(awfl_ctx* ctx, const double* d_in, const double* d_tr_in, double* d_out, double* d_tr_out,
                       double dt, int dir);

/* Temporal_ader::timeStep with tracers (Temporal_ader.h:48-72): the three
 * splits of state and tracers in place; flips dimSwitch. */
int awfl_time_step_tracers(awfl_ctx* ctx, double* d_state, double* d_tracers, double dt);

#ifdef __cplusplus
}
#endif
#endif /* AWFL_DYCORE_H */
