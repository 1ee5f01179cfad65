// awfl_splitting.hip -- alternative time integrators over the same FP64
// spatial operator (C ABI awfl_time_step_strang / awfl_time_step_ssprk3):
//   * qjrmsCode/TimeIntegrator.cpp:32-65: x and y sweeps with dt, then z as
//     two sweeps with dt/2 (reverse order on alternate steps), each sweep one
//     fused kernel as in awfl_time_step;
//   * qjrmsCode/TimeIntegrator.cpp:68-110: SSPRK3 on the semi-discrete
//     operator -- the sum of the three directional tendencies with the ADER
//     time average off (dt = 0 in the transform), from the tendency mode of
//     the same sweep kernels.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "awfl_internal.hpp"

namespace {

// state2 = c0 state0 + c1 state1 + cdt tend over the interior, in that order
// (qjrmsCode/TimeIntegrator.cpp:99-110); state2 may be state0 or state1;
// grid (ceil(nx/64), ny, nz)
__global__ __launch_bounds__(64) void k_combine(double* s2, double c0, const double* s0, double c1,
                                                const double* s1, double cdt, const double* __restrict__ tend,
                                                Grid G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nx) return;
  const int j = blockIdx.y, k = blockIdx.z;
  const int64_t o = (int64_t)(G.hs + k) * G.sz + (int64_t)(G.hs + j) * G.sy + G.hs + i;
  const int64_t t = ((int64_t)k * G.ny + j) * G.nx + i;
  const int64_t nc = (int64_t)G.nx * G.ny * G.nz;
  for (int l = 0; l < 5; l++) {
    const double a = c0 * s0[l * G.sl + o] + c1 * s1[l * G.sl + o];
    s2[l * G.sl + o] = a + cdt * tend[l * nc + t];
  }
}

__global__ __launch_bounds__(256) void k_accumulate(double* __restrict__ a, const double* __restrict__ b, int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    a[t] += b[t];
}

int ensure_bufs(awfl_ctx* c) {
  for (int b = 0; b < 2; b++) {
    if (!c->d_buf[b]) HIP_TRY(c, hipMalloc(&c->d_buf[b], sizeof(double) * c->state_elems()));
    if (!c->r_tend[b]) HIP_TRY(c, hipMalloc(&c->r_tend[b], sizeof(double) * c->tend_elems()));
  }
  return AWFL_OK;
}

// semi-discrete tendency of `st` into tend (the three directions summed)
int semi_discrete(awfl_ctx* c, const double* st, double* tend, double* tmp) {
  int rc = launch_sweep_fp64(c, st, tend, 0.0, 0, true);
  if (rc != AWFL_OK) return rc;
  const int64_t n = c->tend_elems();
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  for (int d = 1; d < 3; d++) {
    if (d == 1 && c->sim2d) continue;
    rc = launch_sweep_fp64(c, st, tmp, 0.0, d, true);
    if (rc != AWFL_OK) return rc;
    k_accumulate<<<blocks, 256, 0, c->stream>>>(tend, tmp, n);
    LAUNCH_CHECK(c);
  }
  return AWFL_OK;
}

int combine(awfl_ctx* c, double* s2, double c0, const double* s0, double c1, const double* s1, double cdt,
            const double* tend) {
  dim3 grid((c->cfg.nx + 63) / 64, c->cfg.ny, c->cfg.nz);
  k_combine<<<grid, 64, 0, c->stream>>>(s2, c0, s0, c1, s1, cdt, tend, c->grid());
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

}  // namespace

extern "C" {

int awfl_time_step_strang(awfl_ctx* c, double* d_state, double dt) {
  if (!c) return AWFL_E_ARG;
  if (!d_state) return fail(c, AWFL_E_ARG, "null state");
  int rc = ensure_bufs(c);
  if (rc != AWFL_OK) return rc;
  int dirs[4];
  double dts[4];
  int n = 0;
  const int fwd[4] = {0, 1, 2, 2};
  const double fdt[4] = {dt, dt, 0.5 * dt, 0.5 * dt};
  for (int s = 0; s < 4; s++) {
    const int q = c->dimSwitch ? s : 3 - s;
    if (fwd[q] == 1 && c->sim2d) continue;
    dirs[n] = fwd[q];
    dts[n++] = fdt[q];
  }
  const double* cur = d_state;
  for (int s = 0; s < n; s++) {
    double* dst = (s == n - 1) ? d_state : (cur == c->d_buf[0] ? c->d_buf[1] : c->d_buf[0]);
    rc = launch_sweep_fp64(c, cur, dst, dts[s], dirs[s], false);
    if (rc != AWFL_OK) return rc;
    cur = dst;
  }
  c->dimSwitch = !c->dimSwitch;
  return AWFL_OK;
}

int awfl_time_step_ssprk3(awfl_ctx* c, double* d_state, double dt) {
  if (!c) return AWFL_E_ARG;
  if (!d_state) return fail(c, AWFL_E_ARG, "null state");
  int rc = ensure_bufs(c);
  if (rc != AWFL_OK) return rc;
  double* tmp = c->d_buf[0];
  double* tend = c->r_tend[0];
  double* t2 = c->r_tend[1];
  if ((rc = semi_discrete(c, d_state, tend, t2)) != AWFL_OK) return rc;
  if ((rc = combine(c, tmp, 1.0, d_state, 0.0, d_state, dt, tend)) != AWFL_OK) return rc;
  if ((rc = semi_discrete(c, tmp, tend, t2)) != AWFL_OK) return rc;
  if ((rc = combine(c, tmp, 0.75, d_state, 0.25, tmp, 0.25 * dt, tend)) != AWFL_OK) return rc;
  if ((rc = semi_discrete(c, tmp, tend, t2)) != AWFL_OK) return rc;
  return combine(c, d_state, 1.0 / 3.0, d_state, 2.0 / 3.0, tmp, 2.0 / 3.0 * dt, tend);
}

}  // extern "C"
