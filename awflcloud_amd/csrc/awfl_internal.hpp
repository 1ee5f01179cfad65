// awfl_internal.hpp -- declarations shared by the translation units of the
// library: the kernel grid description, the context behind awfl_ctx*, and
// error helpers.  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/awfl_dycore.h"
#include "awfl_cell.hpp"

namespace awfl_detail {
using awfl::PowRef;

// ---------------------------------------------------------------------------
// grid description passed to kernels by value
// ---------------------------------------------------------------------------
struct Grid {
  int nx, ny, nz, hs;
  int64_t sy, sz, sl;            // strides of the padded state (x stride 1)
  int bc[3];
  int weno_s, weno_w, sim2d;
  double dx, dy, dz;
  const double* hyDensCells;     // nz+2hs
  const double* hyDensThetaCells;
  const double* hyDensGLL;       // nz*ngll
  const double* hyDensThetaGLL;
  const double* hyPressureGLL;
  const double* hyPressureCells; // nz+2hs (diagnostics only)
  const PowRef* prT;             // [nz]   x/y: B = hyDensThetaCells(hs+k)
  const PowRef* prRT;            // [nz]   y faces: B = hyDensCells * hyDensThetaCells
  const PowRef* prG;             // [nz*ngll] z: B = hyDensThetaGLL(k, kk)
  const PowRef* prF;             // [nz+1] z faces: B at the face
};

extern thread_local std::string g_create_error;

// true when the FP32 perturbation-form kernels exist for (ord, ngll)
bool has_f32_variant(int ord, int ngll);

}  // namespace awfl_detail

using awfl_detail::Grid;

// ---------------------------------------------------------------------------
// XCD-aware tile order.  The 8 XCDs of the MI355X receive work-groups
// round-robin by linear id and each has its own L2.  tile_index() renumbers
// the grid so that consecutive tile numbers (neighbouring x-blocks of one
// row, neighbouring rows) run on the same XCD at about the same time and
// share the cache lines that straddle their boundaries.
// ---------------------------------------------------------------------------
constexpr unsigned AWFL_NUM_XCD = 8;

// Threads of a line work-group: the line rounded up to whole waves, at most
// 256.  The adjoint line kernel also uses it in place of blockDim.x, which
// reaches a kernel through a memory load that the work-group's first loop
// would wait for (measured 1-2% faster there; no gain in the forward line
// kernels).
__host__ __device__ constexpr int line_threads(int n) { return n < 256 ? (n + 63) / 64 * 64 : 256; }

__device__ __forceinline__ unsigned tile_index() {
  const unsigned n = gridDim.x * gridDim.y * gridDim.z;
  const unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned xcd = lin % AWFL_NUM_XCD, local = lin / AWFL_NUM_XCD;
  const unsigned q = n / AWFL_NUM_XCD, r = n % AWFL_NUM_XCD;
  return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// tile_index() decoded into the (x, y, z) block coordinates of the grid, x
// fastest, then y (ZFIRST: then z), so that consecutive tiles differ first in
// x and then along the chosen axis
template <bool ZFIRST = false>
__device__ __forceinline__ dim3 tile_coords() {
  unsigned t = tile_index();
  dim3 b;
  b.x = t % gridDim.x;
  t /= gridDim.x;
  if constexpr (ZFIRST) {
    b.z = t % gridDim.z;
    b.y = t / gridDim.z;
  } else {
    b.y = t % gridDim.y;
    b.z = t / gridDim.y;
  }
  return b;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct awfl_ctx {
  awfl_config cfg{};
  const void* var = nullptr;       // FP64 kernel variant (awfl_dycore.hip)
  int hs = 0;
  bool sim2d = false;
  double dx = 0, dy = 0, dz = 0;
  hipStream_t stream = nullptr;
  int dimSwitch = 1;
  double dtInit = 0;
  int device = 0;
  std::vector<double> hy[8];
  double* d_hy = nullptr;
  awfl::PowRef* d_pref = nullptr;         // power references: prT | prRT | prG | prF
  double* d_buf[2] = {nullptr, nullptr};
  double* d_red = nullptr;          // reduction partials
  // FP32 perturbation form (awfl_pert.hip): background tables and ping-pong
  // buffers, created on first use
  void* d_bgf = nullptr;
  float* f_buf[4] = {nullptr, nullptr, nullptr, nullptr};
  // adjoint (awfl_adjoint.hip): per-cell scratch (edge states, their
  // adjoints, stencil adjoints), one adjoint ping-pong state, and the split
  // inputs of a recomputed checkpoint segment
  double* a_scr = nullptr;
  double* a_tmp = nullptr;
  double* a_states = nullptr;
  int64_t a_states_n = 0;
  // a_states holds the split inputs of the last segment of the checkpointed
  // forward run described here (a_rec_ckpt null: of nothing reusable)
  const double* a_rec_ckpt = nullptr;
  double a_rec_dt = 0;
  int a_rec_nsteps = 0, a_rec_every = 0, a_rec_ds0 = 0;
  int a_rec_kept = 0;               // trailing segments kept: the last in a_states, the one before in a_states2
  // the checkpointed adjoint recomputes the next segment's split inputs into
  // a_states2 on a stream of its own (a_s2, default priority) while the
  // current segment's sweep adjoints run on the context's stream; a_ev
  // orders the two
  double* a_states2 = nullptr;
  int64_t a_states2_n = 0;
  hipStream_t a_s2 = nullptr;
  hipEvent_t a_ev[2] = {nullptr, nullptr};
  // alternative integrators (awfl_splitting.hip): two tendency-sized arrays
  double* r_tend[2] = {nullptr, nullptr};
  // tracers (awfl_tracers.hip): count, positivity bit mask, ping-pong arrays
  int ntr = 0;
  unsigned tpos = 0;
  double* t_buf[2] = {nullptr, nullptr};
  int n_red = 0;
  // this device's copy of the file-scope scratch of awfl_time_step_noalloc
  // (null, with the reason in noalloc_err, if its address was not obtained)
  double* noalloc = nullptr;
  std::string noalloc_err;
  // sweep timing of awfl_time_step (awfl_sweep_timing): event pairs recorded
  // on the stream around each sweep launch, with the sweep's direction
  std::vector<hipEvent_t> t_ev;
  std::vector<int> t_dir;
  int t_n = 0;
  std::string err;
  Grid grid() const;
  int64_t state_elems() const {
    return (int64_t)5 * (cfg.nz + 2 * hs) * (cfg.ny + 2 * hs) * (cfg.nx + 2 * hs);
  }
  int64_t tend_elems() const { return (int64_t)5 * cfg.nz * cfg.ny * cfg.nx; }
};

inline Grid awfl_ctx::grid() const {
  Grid G{};
  G.nx = cfg.nx; G.ny = cfg.ny; G.nz = cfg.nz; G.hs = hs;
  int64_t nxh = cfg.nx + 2 * hs, nyh = cfg.ny + 2 * hs, nzh = cfg.nz + 2 * hs;
  G.sy = nxh; G.sz = nxh * nyh; G.sl = nxh * nyh * nzh;
  G.bc[0] = cfg.bc_x; G.bc[1] = cfg.bc_y; G.bc[2] = cfg.bc_z;
  G.weno_s = cfg.weno_scalars; G.weno_w = cfg.weno_winds; G.sim2d = sim2d;
  G.dx = dx; G.dy = dy; G.dz = dz;
  int nzh_i = (int)nzh, ng = cfg.nz * cfg.ngll;
  G.hyDensCells = d_hy;
  G.hyDensThetaCells = d_hy + nzh_i;
  G.hyDensGLL = d_hy + 2 * nzh_i;
  G.hyDensThetaGLL = d_hy + 2 * nzh_i + ng;
  G.hyPressureGLL = d_hy + 2 * nzh_i + 2 * ng;
  G.hyPressureCells = d_hy + 2 * nzh_i + 3 * ng;
  G.prT = d_pref;
  G.prRT = d_pref + cfg.nz;
  G.prG = d_pref + 2 * cfg.nz;
  G.prF = d_pref + 2 * cfg.nz + ng;
  return G;
}

// launch constants of an FP64 sweep along dir with time step dt
inline awfl::SweepK sweep_constants(const awfl_ctx* c, double dt, int dir) {
  awfl::SweepK K{};
  double dsw = dir == 0 ? c->dx : dir == 1 ? c->dy : c->dz;
  K.dt = dt;
  K.dsw = dsw;
  K.rd = 1.0 / dsw;
  double dtmult = dt;
  for (int kt = 0; kt < awfl::KMAX; kt++) {
    K.ndt[kt] = -1.0 / (dsw * (kt + 1.0));
    if (kt >= 1) {
      K.tav[kt] = dtmult / (kt + 1.0);
      dtmult *= dt;
    }
  }
  return K;
}

// one FP64 sweep along dir (awfl_dycore.hip): out = in + dt * tend, or the
// tendency itself (tend = true, out is tendency-sized)
int launch_sweep_fp64(awfl_ctx* c, const double* in, double* out, double dt, int dir, bool tend);

inline int fail(awfl_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else awfl_detail::g_create_error = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                       \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return fail((ctx), AWFL_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCH_CHECK(ctx)                                                                        \
  do {                                                                                           \
    hipError_t e_ = hipGetLastError();                                                           \
    if (e_ != hipSuccess) return fail((ctx), AWFL_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

