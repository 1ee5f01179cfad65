// awfl_dycore.hip -- fused ADER-WENO split sweeps for gfx950 and the C ABI of
// include/awfl_dycore.h.
//
// Design (one GPU, FP64):
//  * One kernel per split sweep.  It reads the state, generates periodic or
//    solid-wall boundary values on the fly (no halo pass), reconstructs every
//    cell, evolves it with the ADER differential transform, time-averages the
//    edge states, computes the upwind flux at every face and writes
//    out = in + dt * tendency (or the tendency itself).  No intermediate
//    (edge states, fluxes) touches HBM.
//  * x sweep (contiguous axis): one workgroup per x-line; the line and its
//    boundary cells are staged in LDS with coalesced loads, edge states and
//    face fluxes are exchanged through LDS.
//  * y and z sweeps (strided axes): one lane per x-column walks a segment of
//    its column along the sweep axis with a register window of ord cells, so
//    every wavefront load/store is 64 consecutive doubles along x and no LDS
//    or barrier is needed.
//  * Sweeps run out of place; a fused time step ping-pongs through two
//    context-owned buffers and ends in the caller's state buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/awfl_dycore.h"
#include "awfl_cell.hpp"
#include "awfl_profiles.hpp"
#include "awfl_internal.hpp"
#if AWFL_PARITY & AWFL_PAR_ORDER
#include "awfl_cell_ref.hpp"
#endif
#include "awfl_sweep_line.hpp"

// Cell body of the sweeps: the production numerics, or (reference-order
// parity builds) the reference's operation order.
namespace cellns {
#if AWFL_PARITY & AWFL_PAR_ORDER
using awfl::ref::cell_edges;
#else
using awfl::cell_edges;
#endif
}  // namespace cellns

// Flux divergence -(F_high - F_low)/d of one cell along the sweep.
__device__ __forceinline__ double flux_div(double F1, double F0, const awfl::SweepK& K) {
#if AWFL_PARITY & AWFL_PAR_ORDER
  return -(F1 - F0) / K.dsw;
#else
  return -(F1 - F0) * K.rd;
#endif
}

using namespace awfl;

namespace {

__device__ __forceinline__ int64_t tend_index(const Grid& G, int l, int k, int j, int i) {
  return (((int64_t)l * G.nz + k) * G.ny + j) * G.nx + i;
}

// ---------------------------------------------------------------------------
// x sweep: one workgroup per x-line
// ---------------------------------------------------------------------------
#ifndef AWFL_X_WAVES
#define AWFL_X_WAVES 1
#endif
#ifndef AWFL_COL_WAVES
#define AWFL_COL_WAVES 1
#endif

// BR: 0 production; 1 also records each face's upwind branch (un > 0) in
// K.br; 2 takes the branch from K.br instead of the sign test (frozen-branch
// forward for difference checks of the adjoint).  Branch bytes: x face f of
// line (k, j) at (k ny + j)(nx+1) + f; y face f of (k, i) at (k (ny+1) + f) nx
// + i; z face f of (j, i) at (f ny + j) nx + i.
template <int DIR, int BR>
__device__ __forceinline__ void face_flux(const double (&qL)[5], const double (&qR)[5], const FaceRef& fr,
                                          double (&F)[5], unsigned char* bp) {
  if constexpr (BR == 0) {
#if AWFL_PARITY & AWFL_PAR_ORDER
    (void)fr;
    awfl::ref::riemann<DIR>(qL, qR, F);
#else
    riemann<DIR>(qL, qR, fr, F);
#endif
  } else if constexpr (BR == 1) {
    double un;
    riemann<DIR>(qL, qR, fr, F, un);
    *bp = un > 0 ? 1 : 0;
  } else {
    double un;
    riemann<DIR, true>(qL, qR, fr, F, un, *bp != 0);
  }
}

// INPLACE: in == out is allowed.  Every global read of the line happens in
// the staging loop, before the first barrier; the face fluxes get LDS of
// their own, so the update takes each cell's old value from the staged line
// instead of reading it again.  (The line belongs to this work-group alone.)
template <int ORD, int NGLL, bool TEND, int BR = 0, bool INPLACE = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(AWFL_X_WAVES))) void k_sweep_x(const double* __restrict__ in, double* __restrict__ out, Grid G,
                                                 SweepK K) {
  constexpr int hs = (ORD - 1) / 2;
  extern __shared__ double sm[];
  const int nx = G.nx;
  const int W = nx + 2 * hs;
  const unsigned line = tile_index();
  const int j = line % G.ny;
  const int k = line / G.ny;
  double* s_st = sm;              // [5][W]   state of the line with boundary cells
  double* s_L = sm + 5 * W;       // [5][nx]  left edge states
  double* s_R = s_L + 5 * nx;     // [5][nx]  right edge states
  // [5][nx+1] face fluxes: over s_st once the edges are formed, or after s_R
  double* s_F = INPLACE ? s_R + 5 * nx : sm;
  const int64_t row = (int64_t)(hs + k) * G.sz + (int64_t)(hs + j) * G.sy + hs;
  const bool per = G.bc[0] == AWFL_BC_PERIODIC;

  for (int c = threadIdx.x; c < W; c += blockDim.x) {
    int ci = c - hs, src;
    bool zn = false;
    if (per) {
      src = ci % nx;
      if (src < 0) src += nx;
    } else if (ci < 0) {
      src = 0; zn = true;
    } else if (ci >= nx) {
      src = nx - 1; zn = true;
    } else {
      src = ci;
    }
#pragma unroll
    for (int l = 0; l < 5; l++) {
      double v = in[l * G.sl + row + src];
      s_st[l * W + c] = (zn && l == idU) ? 0.0 : v;
    }
  }
  __syncthreads();

  CellBg<NGLL> bg;
  bg.r = G.hyDensCells[hs + k];
  bg.rt = G.hyDensThetaCells[hs + k];
  bg.pt = G.prT[k];
  bg.wall_bot = bg.wall_top = false;
  FaceRef fr;
  fr.rt = bg.pt;
  fr.q5 = bg.pt;
  for (int c = threadIdx.x; c < nx; c += blockDim.x) {
    double st[5][ORD];
    sfor<5>([&](auto l) { sfor<ORD>([&](auto s) { st[l][s] = s_st[l * W + c + s]; }); });
    double eL[5], eR[5], sa;
    cellns::cell_edges<ORD, NGLL, 0>(st, bg, G.weno_s, G.weno_w, K, eL, eR, sa);
#pragma unroll
    for (int l = 0; l < 5; l++) {
      s_L[l * nx + c] = eL[l];
      s_R[l * nx + c] = eR[l];
    }
  }
  __syncthreads();

  // periodic: face nx has the inputs of face 0 (S:819-820), so it is not recomputed
  const int nfaces = per ? nx : nx + 1;
  for (int f = threadIdx.x; f < nfaces; f += blockDim.x) {
    int a, b;                 // qL from the right edge of cell a, qR from the left edge of cell b
    bool aL = false, bR = false;
    if (per) {
      a = f == 0 ? nx - 1 : f - 1;
      b = f == nx ? 0 : f;
    } else if (f == 0) {      // wall: both sides from the first cell's left edge (S:822)
      a = 0; b = 0; aL = true;
    } else if (f == nx) {     // wall: both sides from the last cell's right edge (S:823)
      a = nx - 1; b = nx - 1; bR = true;
    } else {
      a = f - 1; b = f;
    }
    double qL[5], qR[5], F[5];
#pragma unroll
    for (int l = 0; l < 5; l++) {
      qL[l] = aL ? s_L[l * nx + a] : s_R[l * nx + a];
      qR[l] = bR ? s_R[l * nx + b] : s_L[l * nx + b];
    }
    face_flux<0, BR>(qL, qR, fr, F, BR ? K.br + ((int64_t)k * G.ny + j) * (nx + 1) + f : nullptr);
#pragma unroll
    for (int l = 0; l < 5; l++) s_F[l * (nx + 1) + f] = F[l];
  }
  __syncthreads();

  for (int c = threadIdx.x; c < nx; c += blockDim.x) {
#pragma unroll
    for (int l = 0; l < 5; l++) {
      const int cp = (per && c + 1 == nx) ? 0 : c + 1;
      double td = flux_div(s_F[l * (nx + 1) + cp], s_F[l * (nx + 1) + c], K);
      if (G.sim2d && l == idV) td = 0.0;
      if constexpr (TEND) out[tend_index(G, l, k, j, c)] = td;
      else if constexpr (INPLACE) out[l * G.sl + row + c] = s_st[l * W + hs + c] + K.dt * td;
      else out[l * G.sl + row + c] = in[l * G.sl + row + c] + K.dt * td;
    }
  }
}

// ---------------------------------------------------------------------------
// y / z sweeps: one lane per x-column walking a segment of its column
// ---------------------------------------------------------------------------
template <int ORD, int NGLL, int DIR, bool TEND, int BR = 0>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(AWFL_COL_WAVES))) void k_sweep_col(const double* __restrict__ in, double* __restrict__ out, Grid G,
                                                  SweepK K, int seglen) {
  constexpr int hs = (ORD - 1) / 2;
  constexpr int NM = 1 + DIR;
  const dim3 tb = tile_coords();
  const int i = tb.x * blockDim.x + threadIdx.x;
  if (i >= G.nx) return;              // no barrier in this kernel
  const int o = tb.y;                 // y sweep: k; z sweep: j
  const int n = DIR == 1 ? G.ny : G.nz;
  const int64_t sa = DIR == 1 ? G.sy : G.sz;
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  const int64_t base = DIR == 1 ? (int64_t)(hs + o) * G.sz + (int64_t)hs * G.sy + hs + i
                                : (int64_t)hs * G.sz + (int64_t)(hs + o) * G.sy + hs + i;
  const int s0 = tb.z * seglen;
  const int s1 = min(n, s0 + seglen);

  auto ld = [&](int l, int a) -> double {
    int src;
    bool zn = false;
    if (per) {
      if (n > hs) {                    // the walk reaches at most hs + 1 cells past either end
        src = a < 0 ? a + n : (a >= n ? a - n : a);
      } else {
        src = a % n;
        if (src < 0) src += n;
      }
    } else if (a < 0) {
      src = 0; zn = true;
    } else if (a >= n) {
      src = n - 1; zn = true;
    } else {
      src = a;
    }
    double v = in[l * G.sl + base + (int64_t)src * sa];
    return (zn && l == NM) ? 0.0 : v;
  };
  auto bg_of = [&](int kc, CellBg<NGLL>& bg) {
    if constexpr (DIR == 1) {
      bg.r = G.hyDensCells[hs + o];
      bg.rt = G.hyDensThetaCells[hs + o];
      bg.pt = G.prT[o];
      bg.wall_bot = bg.wall_top = false;
    } else {
      sfor<NGLL>([&](auto ii) {
        bg.rG[ii] = G.hyDensGLL[kc * NGLL + ii];
        bg.rtG[ii] = G.hyDensThetaGLL[kc * NGLL + ii];
        bg.pG[ii] = G.hyPressureGLL[kc * NGLL + ii];
        bg.ptG[ii] = G.prG[kc * NGLL + ii];
      });
      bg.wall_bot = !per && kc == 0;
      bg.wall_top = !per && kc == n - 1;
    }
  };
  auto face_ref = [&](int f) {
    FaceRef fr;
    if constexpr (DIR == 1) {
      fr.rt = G.prRT[o];
      fr.q5 = G.prT[o];
    } else {
      fr.rt = G.prF[f];
      fr.q5 = fr.rt;
    }
    return fr;
  };
  auto branch = [&](int f) -> unsigned char* {
    if constexpr (BR == 0) return nullptr;
    return K.br + (DIR == 1 ? ((int64_t)o * (G.ny + 1) + f) * G.nx + i : ((int64_t)f * G.ny + o) * G.nx + i);
  };
  auto face_bg = [&](int f, double (&F)[5]) {   // S:1628-1629
    if constexpr (DIR == 2) {
      if (f < G.nz) F[idW] -= G.hyPressureGLL[f * NGLL + 0];
      else F[idW] -= G.hyPressureGLL[(G.nz - 1) * NGLL + NGLL - 1];
    }
  };
  // result of one cell (new state, or its tendency) into registers, and its store
  auto update = [&](const double (&F0)[5], const double (&F1)[5], double src, const double (&cell)[5],
                    double (&res)[5]) {
#pragma unroll
    for (int l = 0; l < 5; l++) {
      double td = flux_div(F1[l], F0[l], K);
      if (DIR == 2 && l == idW) td = src + td;
      if (DIR == 2 && G.sim2d && l == idV) td = 0.0;
      res[l] = TEND ? td : cell[l] + K.dt * td;
    }
  };
  auto put = [&](int c, const double (&res)[5]) {
#pragma unroll
    for (int l = 0; l < 5; l++) {
      if constexpr (TEND) {
        if constexpr (DIR == 1) out[tend_index(G, l, o, c, i)] = res[l];
        else out[tend_index(G, l, c, o, i)] = res[l];
      } else {
        out[l * G.sl + base + (int64_t)c * sa] = res[l];
      }
    }
  };

  double w[5][ORD];
  sfor<5>([&](auto l) { sfor<ORD>([&](auto s) { w[l][s] = ld(l, s0 - 1 - hs + s); }); });
  double prevR[5], prevF[5], prevCell[5], prevSrc = 0.0;
  if (per || s0 > 0) {                 // right edge of the cell below the segment
    CellBg<NGLL> bg;
    bg_of(((s0 - 1) % n + n) % n, bg);
    double eL[5], sa_;
    cellns::cell_edges<ORD, NGLL, DIR>(w, bg, G.weno_s, G.weno_w, K, eL, prevR, sa_);
  }
  // software pipeline: the stencil cell of the next iteration is loaded one
  // iteration ahead, and the result of an iteration is stored at the top of
  // the next one, ahead of that iteration's loads: a wait for the loads then
  // never waits for stores issued just before it
  double nxt[5], res[5];
  sfor<5>([&](auto l) {
    nxt[l] = ld(l, s0 + hs);
    prevF[l] = prevCell[l] = res[l] = 0.0;
  });
  for (int v = s0; v < s1; v++) {
    if (v > s0 + 1) put(v - 2, res);
    sfor<5>([&](auto l) {
      sfor<ORD - 1>([&](auto s) { w[l][s] = w[l][s + 1]; });
      w[l][ORD - 1] = nxt[l];
      nxt[l] = ld(l, v + 1 + hs);
    });
    CellBg<NGLL> bg;
    bg_of(v, bg);
    double eL[5], eR[5], src;
    cellns::cell_edges<ORD, NGLL, DIR>(w, bg, G.weno_s, G.weno_w, K, eL, eR, src);
    double F[5];
    FaceRef fr = face_ref(v);
    if (!per && v == 0) face_flux<DIR, BR>(eL, eL, fr, F, branch(v));   // wall: both sides from the first cell
    else face_flux<DIR, BR>(prevR, eL, fr, F, branch(v));
    face_bg(v, F);
    update(prevF, F, prevSrc, prevCell, res);   // cell v-1 (for v == s0 a value never stored)
#pragma unroll
    for (int l = 0; l < 5; l++) {
      prevF[l] = F[l];
      prevR[l] = eR[l];
      prevCell[l] = w[l][hs];
    }
    prevSrc = src;
  }
  if (s1 > s0 + 1) put(s1 - 2, res);
  // top face of the segment
  double F[5];
  FaceRef fr = face_ref(s1);
  if (!per && s1 == n) {
    face_flux<DIR, BR>(prevR, prevR, fr, F, branch(s1));   // wall: both sides from the last cell
  } else {
    sfor<5>([&](auto l) {
      sfor<ORD - 1>([&](auto s) { w[l][s] = w[l][s + 1]; });
      w[l][ORD - 1] = nxt[l];          // cell s1 + hs, loaded by the last iteration
    });
    CellBg<NGLL> bg;
    bg_of(s1 % n, bg);
    double eL[5], eR[5], sa_;
    cellns::cell_edges<ORD, NGLL, DIR>(w, bg, G.weno_s, G.weno_w, K, eL, eR, sa_);
    face_flux<DIR, BR>(prevR, eL, fr, F, branch(s1));
  }
  face_bg(s1, F);
  update(prevF, F, prevSrc, prevCell, res);
  put(s1 - 1, res);
}

// ---------------------------------------------------------------------------
// small kernels: tendency application, CFL time step, sums, initial data, fill
// ---------------------------------------------------------------------------
__global__ void k_apply(double* __restrict__ st, const double* __restrict__ tend, Grid G, double dt) {
  int64_t n = (int64_t)5 * G.nz * G.ny * G.nx;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    int i = t % G.nx;
    int64_t r = t / G.nx;
    int j = r % G.ny;
    r /= G.ny;
    int k = r % G.nz;
    int l = r / G.nz;
    st[l * G.sl + (int64_t)(G.hs + k) * G.sz + (int64_t)(G.hs + j) * G.sy + G.hs + i] += dt * tend[t];
  }
}

__global__ void k_eos(const double* __restrict__ rt, double* __restrict__ p, int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    p[t] = C0 * pow_gamma(rt[t]);
}

// S:1752-1785: diagnostics dens_pert, u, v, w, pot_temp_pert, pressure_pert
// of the interior, out[6][nz][ny][nx]
__global__ void k_diag(const double* __restrict__ st, Grid G, double* __restrict__ out) {
  const int64_t n = (int64_t)G.nx * G.ny * G.nz;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % G.nx);
    const int j = (int)((t / G.nx) % G.ny);
    const int k = (int)(t / ((int64_t)G.nx * G.ny));
    const int64_t c = (int64_t)(G.hs + k) * G.sz + (int64_t)(G.hs + j) * G.sy + G.hs + i;
    const double rb = G.hyDensCells[G.hs + k], rtb = G.hyDensThetaCells[G.hs + k];
    const double r = st[idR * G.sl + c] + rb;
    const double th = (st[idT * G.sl + c] + rtb) / r;
    out[t] = st[idR * G.sl + c];
    out[n + t] = st[idU * G.sl + c] / r;
    out[2 * n + t] = st[idV * G.sl + c] / r;
    out[3 * n + t] = st[idW * G.sl + c] / r;
    out[4 * n + t] = th - rtb / rb;
    out[5 * n + t] = C0 * pow_gamma(r * th) - G.hyPressureCells[G.hs + k];
  }
}

__global__ void k_fill(double* __restrict__ p, int64_t n, double v) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}

constexpr int RED_THREADS = 256;

__device__ __forceinline__ double block_min(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  return sh[0];
}

// S:169-183: per-cell CFL time step, block minima
__global__ __launch_bounds__(RED_THREADS) void k_dt_partial(const double* __restrict__ st, Grid G, double cfl,
                                                            double* __restrict__ partial) {
  __shared__ double sh[RED_THREADS];
  int64_t n = (int64_t)G.nz * G.ny * G.nx;
  double m = INFINITY;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    int i = t % G.nx;
    int64_t rr = t / G.nx;
    int j = rr % G.ny;
    int k = rr / G.ny;
    int64_t c = (int64_t)(G.hs + k) * G.sz + (int64_t)(G.hs + j) * G.sy + G.hs + i;
    double r = st[0 * G.sl + c] + G.hyDensCells[G.hs + k];
    double u = st[1 * G.sl + c] / r;
    double v = st[2 * G.sl + c] / r;
    double w = st[3 * G.sl + c] / r;
    double th = (st[4 * G.sl + c] + G.hyDensThetaCells[G.hs + k]) / r;
    double p = C0 * pow(r * th, GAMMA);
    double cs = sqrt(GAMMA * p / r);
    double udt = cfl * G.dx / fmax(fabs(u - cs), fabs(u + cs));
    double vdt = cfl * G.dy / fmax(fabs(v - cs), fabs(v + cs));
    double wdt = cfl * G.dz / fmax(fabs(w - cs), fabs(w + cs));
    m = fmin(m, fmin(fmin(udt, vdt), wdt));
  }
  double b = block_min(m, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

__global__ __launch_bounds__(RED_THREADS) void k_min_final(const double* __restrict__ partial, int n,
                                                           double* __restrict__ out) {
  __shared__ double sh[RED_THREADS];
  double m = INFINITY;
  for (int t = threadIdx.x; t < n; t += blockDim.x) m = fmin(m, partial[t]);
  double b = block_min(m, sh);
  if (threadIdx.x == 0) out[0] = b;
}

// per-variable interior sums: block partials [block][5] then a fixed-order final sum
__global__ __launch_bounds__(RED_THREADS) void k_sum_partial(const double* __restrict__ st, Grid G,
                                                             double* __restrict__ partial) {
  __shared__ double sh[5][RED_THREADS];
  int64_t n = (int64_t)G.nz * G.ny * G.nx;
  double acc[5] = {0, 0, 0, 0, 0};
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    int i = t % G.nx;
    int64_t rr = t / G.nx;
    int j = rr % G.ny;
    int k = rr / G.ny;
    int64_t c = (int64_t)(G.hs + k) * G.sz + (int64_t)(G.hs + j) * G.sy + G.hs + i;
#pragma unroll
    for (int l = 0; l < 5; l++) acc[l] += st[l * G.sl + c];
  }
#pragma unroll
  for (int l = 0; l < 5; l++) sh[l][threadIdx.x] = acc[l];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int l = 0; l < 5; l++) sh[l][threadIdx.x] += sh[l][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 5) partial[blockIdx.x * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_sum_final(const double* __restrict__ partial, int nb, double* __restrict__ out) {
  if (threadIdx.x < 5) {
    double s = 0;
    for (int b = 0; b < nb; b++) s += partial[b * 5 + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// initial data: Profiles.h:8-73 (awfl_profiles.hpp) and the ord^3 GLL
// quadrature of S:428-485
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__host__ __device__ inline double uniform_pm1(uint64_t seed, uint64_t idx) {
  uint64_t h = splitmix64(splitmix64(seed) ^ idx);
  return (double)(h >> 11) * 0x1.0p-52 - 1.0;
}

struct InitDev {
  int nbub;
  int shape[4];
  double x0[4], y0[4], z0[4], xr[4], yr[4], zr[4], amp[4], pwr[4];
  double noise_amp;
  uint64_t noise_seed;
  int nmodes;
  const double* modes;
  double u0, v0;
  double y2d, perturb_amp;
  uint64_t perturb_seed;
  int igw;
  double xlen, ylen, zlen;
};

template <int ORD>
__global__ void k_init(double* __restrict__ st, Grid G, InitDev S) {
  constexpr int hs = (ORD - 1) / 2;
  int64_t n = (int64_t)G.nz * G.ny * G.nx;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    int i = t % G.nx;
    int64_t rr = t / G.nx;
    int j = rr % G.ny;
    int k = rr / G.ny;
    int64_t c = (int64_t)(hs + k) * G.sz + (int64_t)(hs + j) * G.sy + hs + i;
    double sT = 0, sU = 0, sV = 0;
    double cell_tp = 0;
    if (S.noise_amp != 0) cell_tp += S.noise_amp * uniform_pm1(S.noise_seed, ((uint64_t)k * G.ny + j) * G.nx + i);
    if (S.nmodes > 0) {
      double xc = (i + 0.5) * G.dx, yc = (j + 0.5) * G.dy, zc = (k + 0.5) * G.dz;
      double acc = 0;
      for (int m = 0; m < S.nmodes; m++) {
        const double* md = S.modes + 5 * m;
        acc += md[3] * cos(2 * M_PI * (md[0] * xc / S.xlen + md[1] * yc / S.ylen + md[2] * zc / S.zlen) + md[4]);
      }
      cell_tp += acc;
    }
    for (int kk = 0; kk < ORD; kk++) {
      double zloc = (k + 0.5) * G.dz + Gll<ORD>::pts[kk] * G.dz;
      double th, rh;
      if (!S.igw) {
        th = 300;
        rh = const_theta_density(th, zloc);
      } else {
        rh = bvf_density(300., 0.01, zloc);
        th = bvf_theta(300., 0.01, zloc);
      }
      for (int jj = 0; jj < ORD; jj++)
        for (int ii = 0; ii < ORD; ii++) {
          double yloc = G.sim2d ? (S.y2d != 0 ? S.y2d : S.ylen / 2) : (j + 0.5) * G.dy + Gll<ORD>::pts[jj] * G.dy;
          double xloc = (i + 0.5) * G.dx + Gll<ORD>::pts[ii] * G.dx;
          double wt = Gll<ORD>::wts[kk] * Gll<ORD>::wts[jj] * Gll<ORD>::wts[ii];
          double tp = 0;
          for (int b = 0; b < S.nbub; b++) {
            double xn = (xloc - S.x0[b]) / S.xr[b], yn = (yloc - S.y0[b]) / S.yr[b], zn = (zloc - S.z0[b]) / S.zr[b];
            double dist = sqrt(xn * xn + yn * yn + zn * zn);
            if (S.shape[b] == 0) tp += S.amp[b] * fmax(1. - dist, 0.);
            else if (S.shape[b] == 2) tp += S.amp[b] * sin(M_PI * zloc / S.zlen) / sqrt(1 + xn * xn + yn * yn);
            else tp += dist <= 1. ? S.amp[b] * pow((cos(M_PI * dist) + 1) / 2, S.pwr[b]) : 0.;
          }
          tp += cell_tp;
          double tt = th + tp;
          sT += (rh * tt - rh * th) * wt;
          if (S.u0 != 0) sU += S.u0 * rh * wt;
          if (S.v0 != 0 && !G.sim2d) sV += S.v0 * rh * wt;
        }
    }
    if (S.perturb_amp != 0) {     // S:488-499 with a seeded generator
      uint64_t h = splitmix64(splitmix64(S.perturb_seed) ^ (((uint64_t)k * G.ny + j) * G.nx + i));
      sT *= 1. + (double)(h >> 11) * 0x1.0p-53 * S.perturb_amp;
    }
    st[0 * G.sl + c] = 0;
    st[1 * G.sl + c] = sU;
    st[2 * G.sl + c] = sV;
    st[3 * G.sl + c] = 0;
    st[4 * G.sl + c] = sT;
  }
}

// ---------------------------------------------------------------------------
// compiled (ord, ngll) variants
// ---------------------------------------------------------------------------
// the in-place x kernel of awfl_time_step_noalloc: the per-cell x kernel, or
// for (9, 5) the x line kernel
template <int O, int N>
constexpr void (*inplace_x())(const double*, double*, Grid, SweepK) {
  if constexpr (O == 9) return awfl_line::k_sweep_line<O, N, 0, false, true>;
  else return k_sweep_x<O, N, false, 0, true>;
}

// the lane-split line sweeps (awfl_sweep_line.hpp), used for (9, 5), whose
// per-cell kernels need more registers than a lane has
using LineFn = void (*)(const double*, double*, Grid, SweepK);
template <int O, int N, int D, bool T>
constexpr LineFn line_fn() {
  if constexpr (O == 9) return awfl_line::k_sweep_line<O, N, D, T>;
  else return nullptr;
}

struct Variant {
  int ord, ngll;
  void (*sweep_x[2])(const double*, double*, Grid, SweepK);
  void (*sweep_x_inplace)(const double*, double*, Grid, SweepK);
  void (*sweep_y[2])(const double*, double*, Grid, SweepK, int);
  void (*sweep_z[2])(const double*, double*, Grid, SweepK, int);
  void (*init)(double*, Grid, InitDev);
  LineFn sweep_line[3][2];             // [dir][tend]; null: the per-cell kernels above
};

// per-cell sweep kernels, for the orders that do not use the line kernels
template <int O, int N, bool T>
constexpr void (*cell_x())(const double*, double*, Grid, SweepK) {
  if constexpr (O == 9) return nullptr;
  else return k_sweep_x<O, N, T>;
}
template <int O, int N, int D, bool T>
constexpr void (*cell_col())(const double*, double*, Grid, SweepK, int) {
  if constexpr (O == 9) return nullptr;
  else return k_sweep_col<O, N, D, T>;
}

#define AWFL_VARIANT(O, N)                                                                        \
  Variant{O, N,                                                                                   \
          {cell_x<O, N, false>(), cell_x<O, N, true>()},                                          \
          inplace_x<O, N>(),                                                                      \
          {cell_col<O, N, 1, false>(), cell_col<O, N, 1, true>()},                                \
          {cell_col<O, N, 2, false>(), cell_col<O, N, 2, true>()},                                \
          k_init<O>,                                                                              \
          {{line_fn<O, N, 0, false>(), line_fn<O, N, 0, true>()},                                 \
           {line_fn<O, N, 1, false>(), line_fn<O, N, 1, true>()},                                 \
           {line_fn<O, N, 2, false>(), line_fn<O, N, 2, true>()}}}

#ifdef AWFL_LEAN
// parity builds: the FP64 (5, 3) sweeps only, no FP32 form in the library
const Variant kVariants[] = {AWFL_VARIANT(5, 3)};
}  // namespace
namespace awfl_detail {
bool has_f32_variant(int, int) { return false; }
}  // namespace awfl_detail
namespace {
#else
const Variant kVariants[] = {AWFL_VARIANT(3, 2), AWFL_VARIANT(5, 3), AWFL_VARIANT(9, 5)};
#endif

const Variant* find_variant(int ord, int ngll) {
  for (const Variant& v : kVariants)
    if (v.ord == ord && v.ngll == ngll) return &v;
  return nullptr;
}

}  // namespace

namespace awfl_detail {
thread_local std::string g_create_error;
}  // namespace awfl_detail
using awfl_detail::g_create_error;


// S:370-425: hydrostatic background, FP64 on the host
static void build_background(awfl_ctx* c) {
  int hs = c->hs, ord = c->cfg.ord, ngll = c->cfg.ngll, nz = c->cfg.nz, nzh = nz + 2 * hs;
  bool igw = c->cfg.data_spec == 4;
  const double* pts_o = nullptr; const double* wts_o = nullptr; const double* pts_g = nullptr;
  switch (ord) {
    case 3: pts_o = awfl_gll_pts_3; wts_o = awfl_gll_wts_3; break;
    case 5: pts_o = awfl_gll_pts_5; wts_o = awfl_gll_wts_5; break;
    case 7: pts_o = awfl_gll_pts_7; wts_o = awfl_gll_wts_7; break;
    case 9: pts_o = awfl_gll_pts_9; wts_o = awfl_gll_wts_9; break;
  }
  switch (ngll) {
    case 2: pts_g = awfl_gll_pts_2; break; case 3: pts_g = awfl_gll_pts_3; break;
    case 4: pts_g = awfl_gll_pts_4; break; case 5: pts_g = awfl_gll_pts_5; break;
    case 6: pts_g = awfl_gll_pts_6; break; case 7: pts_g = awfl_gll_pts_7; break;
    case 8: pts_g = awfl_gll_pts_8; break; case 9: pts_g = awfl_gll_pts_9; break;
  }
  for (int w = 0; w < 4; w++) c->hy[w].assign(nzh, 0.0);
  for (int w = 4; w < 8; w++) c->hy[w].assign((size_t)nz * ngll, 0.0);
  double dz = c->dz;
  for (int k = 0; k < nzh; k++) {
    for (int kk = 0; kk < ord; kk++) {
      double zloc = (k - hs + 0.5) * dz + pts_o[kk] * dz;
      double th, rh, ph;
      if (!igw) { th = 300; rh = const_theta_density(th, zloc); ph = const_theta_pressure(th, zloc); }
      else { th = bvf_theta(300., 0.01, zloc); rh = bvf_density(300., 0.01, zloc); ph = bvf_pressure(300., 0.01, zloc); }
      double wt = wts_o[kk];
      c->hy[0][k] += rh * wt;
      c->hy[2][k] += th * wt;
      c->hy[3][k] += rh * th * wt;
      c->hy[1][k] += ph * wt;
    }
  }
  for (int k = 0; k < nz; k++)
    for (int kk = 0; kk < ngll; kk++) {
      double zloc = (k + 0.5) * dz + pts_g[kk] * dz;
      double th, rh, ph;
      if (!igw) { th = 300; rh = const_theta_density(th, zloc); ph = const_theta_pressure(th, zloc); }
      else { th = bvf_theta(300., 0.01, zloc); rh = bvf_density(300., 0.01, zloc); ph = bvf_pressure(300., 0.01, zloc); }
      c->hy[4][(size_t)k * ngll + kk] = rh;
      c->hy[6][(size_t)k * ngll + kk] = th;
      c->hy[7][(size_t)k * ngll + kk] = rh * th;
      c->hy[5][(size_t)k * ngll + kk] = ph;
    }
}

// B^GAMMA as a double-double (long double powl, 64-bit mantissa) and 1/B
static PowRef power_ref(double B) {
  long double v = powl((long double)B, (long double)GAMMA);
  PowRef r;
  r.h = (double)v;
  r.l = (double)(v - (long double)r.h);
  r.rb = 1.0 / B;
  r.b = B;
  return r;
}

// reference points of pow_gamma_near, laid out prT[nz] | prRT[nz] | prG[nz*ngll] | prF[nz+1]
static std::vector<PowRef> power_references(const awfl_ctx* c) {
  int nz = c->cfg.nz, ngll = c->cfg.ngll, hs = c->hs;
  std::vector<PowRef> out;
  for (int k = 0; k < nz; k++) out.push_back(power_ref(c->hy[3][hs + k]));
  for (int k = 0; k < nz; k++) out.push_back(power_ref(c->hy[0][hs + k] * c->hy[3][hs + k]));
  for (int k = 0; k < nz; k++)
    for (int kk = 0; kk < ngll; kk++) out.push_back(power_ref(c->hy[7][(size_t)k * ngll + kk]));
  for (int f = 0; f <= nz; f++)
    out.push_back(power_ref(f < nz ? c->hy[7][(size_t)f * ngll] : c->hy[7][(size_t)(nz - 1) * ngll + ngll - 1]));
  return out;
}

// Scratch of awfl_time_step_noalloc: the one state-sized intermediate of a
// step, in a file-scope device array so that the entry point allocates
// nothing.  Each device holds its own copy from the library's first use on it
// until the process ends.  Host code never uses the array's name as an
// address: it reaches each device's copy through the address
// hipGetSymbolAddress returns with that device current, obtained once per
// device by awfl_create and kept in g_noalloc_ptr.  Nothing in it needs to
// start at zero: every cell a sweep reads from it was written by the sweep
// before (halo cells are never read).
#ifdef AWFL_LEAN
// (parity test builds carry no scratch: awfl_time_step_noalloc refuses)
static __device__ __align__(256) double scratch[32];
#else
static __device__ __align__(256) double scratch[AWFL_NOALLOC_MAX_BYTES / sizeof(double)];
#endif

namespace {
std::mutex g_noalloc_mu;
std::map<int, double*> g_noalloc_ptr;

// this device's scratch address (device current); error text on failure
double* noalloc_scratch(int device, std::string& err) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  auto it = g_noalloc_ptr.find(device);
  if (it != g_noalloc_ptr.end()) return it->second;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return nullptr;
  }
  g_noalloc_ptr[device] = static_cast<double*>(p);
  return static_cast<double*>(p);
}
}  // namespace

extern "C" {

const char* awfl_last_error(const awfl_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int awfl_create(const awfl_config* cfg, awfl_ctx** out) {
  if (!cfg || !out) return fail(nullptr, AWFL_E_ARG, "null argument");
  *out = nullptr;
  if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return fail(nullptr, AWFL_E_ARG, "nx, ny, nz must be >= 1");
  if (!(cfg->xlen > 0 && cfg->ylen > 0 && cfg->zlen > 0)) return fail(nullptr, AWFL_E_ARG, "lengths must be > 0");
  for (int b : {cfg->bc_x, cfg->bc_y, cfg->bc_z})
    if (b != AWFL_BC_PERIODIC && b != AWFL_BC_WALL) return fail(nullptr, AWFL_E_ARG, "bc must be periodic or wall");
  if (cfg->data_spec < 1 || cfg->data_spec > 4) return fail(nullptr, AWFL_E_ARG, "data_spec must be 1..4");
  if (cfg->precision != AWFL_FP64 && cfg->precision != AWFL_FP32_PERT)
    return fail(nullptr, AWFL_E_ARG, "precision must be AWFL_FP64 or AWFL_FP32_PERT");
  const Variant* v = find_variant(cfg->ord, cfg->ngll);
  if (!v) return fail(nullptr, AWFL_E_UNSUPPORTED, "(ord, ngll) not compiled in");
  if (cfg->precision == AWFL_FP32_PERT && !awfl_detail::has_f32_variant(cfg->ord, cfg->ngll))
    return fail(nullptr, AWFL_E_UNSUPPORTED, "(ord, ngll) not compiled in for AWFL_FP32_PERT");
  int hs = (cfg->ord - 1) / 2;
  if (cfg->bc_x == AWFL_BC_PERIODIC && cfg->nx < hs) return fail(nullptr, AWFL_E_ARG, "nx < hs");
  int64_t lds = (int64_t)8 * (5 * (cfg->nx + 2 * hs) + 10 * cfg->nx);
  if (lds > 160 * 1024) return fail(nullptr, AWFL_E_UNSUPPORTED, "nx too large for the x-line kernel's LDS");
  awfl_ctx* c = new awfl_ctx();
  c->cfg = *cfg;
  c->var = v;
  c->hs = hs;
  c->sim2d = cfg->ny == 1;
  c->dx = cfg->xlen / cfg->nx;
  c->dy = cfg->ylen / cfg->ny;
  c->dz = cfg->zlen / cfg->nz;
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) {
    delete c;
    return fail(nullptr, AWFL_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  }
  c->noalloc = noalloc_scratch(c->device, c->noalloc_err);
  build_background(c);
  int nzh = cfg->nz + 2 * hs, ng = cfg->nz * cfg->ngll;
  std::vector<double> packed;
  packed.insert(packed.end(), c->hy[0].begin(), c->hy[0].end());
  packed.insert(packed.end(), c->hy[3].begin(), c->hy[3].end());
  packed.insert(packed.end(), c->hy[4].begin(), c->hy[4].end());
  packed.insert(packed.end(), c->hy[7].begin(), c->hy[7].end());
  packed.insert(packed.end(), c->hy[5].begin(), c->hy[5].end());
  packed.insert(packed.end(), c->hy[1].begin(), c->hy[1].end());
  (void)nzh; (void)ng;
  e = hipMalloc(&c->d_hy, packed.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(c->d_hy, packed.data(), packed.size() * sizeof(double), hipMemcpyHostToDevice);
  std::vector<PowRef> pref = power_references(c);
  if (e == hipSuccess) e = hipMalloc(&c->d_pref, pref.size() * sizeof(PowRef));
  if (e == hipSuccess) e = hipMemcpy(c->d_pref, pref.data(), pref.size() * sizeof(PowRef), hipMemcpyHostToDevice);
  c->n_red = 1024;
  if (e == hipSuccess) e = hipMalloc(&c->d_red, (size_t)(c->n_red * 5 + 8) * sizeof(double));
  if (e != hipSuccess) {
    std::string m = std::string("device allocation: ") + hipGetErrorString(e);
    awfl_destroy(c);
    return fail(nullptr, AWFL_E_HIP, m);
  }
  *out = c;
  return AWFL_OK;
}

static void timing_free(awfl_ctx* c);

void awfl_destroy(awfl_ctx* c) {
  if (!c) return;
  timing_free(c);
  if (c->d_hy) (void)hipFree(c->d_hy);
  if (c->d_pref) (void)hipFree(c->d_pref);
  if (c->d_red) (void)hipFree(c->d_red);
  for (double* b : c->d_buf)
    if (b) (void)hipFree(b);
  if (c->d_bgf) (void)hipFree(c->d_bgf);
  for (float* b : c->f_buf)
    if (b) (void)hipFree(b);
  for (double* b : {c->a_scr, c->a_tmp, c->a_states, c->a_states2, c->r_tend[0], c->r_tend[1], c->t_buf[0], c->t_buf[1]})
    if (b) (void)hipFree(b);
  for (hipEvent_t e : c->a_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->a_s2) (void)hipStreamDestroy(c->a_s2);
  delete c;
}

int awfl_layout(const awfl_ctx* c, int* hs, int64_t dims_state[4], int64_t dims_tend[4]) {
  if (!c) return AWFL_E_ARG;
  if (hs) *hs = c->hs;
  if (dims_state) {
    dims_state[0] = 5; dims_state[1] = c->cfg.nz + 2 * c->hs;
    dims_state[2] = c->cfg.ny + 2 * c->hs; dims_state[3] = c->cfg.nx + 2 * c->hs;
  }
  if (dims_tend) {
    dims_tend[0] = 5; dims_tend[1] = c->cfg.nz; dims_tend[2] = c->cfg.ny; dims_tend[3] = c->cfg.nx;
  }
  return AWFL_OK;
}

int awfl_num_split(const awfl_ctx* c) { return c ? 3 : AWFL_E_ARG; }

int awfl_set_stream(awfl_ctx* c, void* s) {
  if (!c) return AWFL_E_ARG;
  // split inputs kept by awfl_forward_ckpt were written on the old stream:
  // not reused from the new one
  if ((hipStream_t)s != c->stream) c->a_rec_ckpt = nullptr;
  c->stream = (hipStream_t)s;
  return AWFL_OK;
}

int awfl_get_dim_switch(const awfl_ctx* c) { return c ? c->dimSwitch : AWFL_E_ARG; }
int awfl_set_dim_switch(awfl_ctx* c, int v) {
  if (!c) return AWFL_E_ARG;
  c->dimSwitch = v ? 1 : 0;
  return AWFL_OK;
}

int awfl_get_background(const awfl_ctx* c, int which, double* h_out, int64_t cap) {
  if (!c || !h_out || which < 0 || which > 7) return AWFL_E_ARG;
  const std::vector<double>& v = c->hy[which];
  if ((int64_t)v.size() > cap) return AWFL_E_ARG;
  std::memcpy(h_out, v.data(), v.size() * sizeof(double));
  return (int)v.size();
}

int awfl_init_state(awfl_ctx* c, double* d_state, const awfl_init_spec* s) {
  if (!c) return AWFL_E_ARG;
  if (!d_state || !s) return fail(c, AWFL_E_ARG, "null state or spec");
  if (s->nbub < 0 || s->nbub > 4) return fail(c, AWFL_E_ARG, "nbub must be 0..4");
  if (s->nmodes < 0 || (s->nmodes > 0 && !s->modes)) return fail(c, AWFL_E_ARG, "bad modes");
  InitDev D{};
  D.nbub = s->nbub;
  for (int b = 0; b < 4; b++) {
    D.shape[b] = s->shape[b]; D.x0[b] = s->x0[b]; D.y0[b] = s->y0[b]; D.z0[b] = s->z0[b];
    D.xr[b] = s->xr[b]; D.yr[b] = s->yr[b]; D.zr[b] = s->zr[b]; D.amp[b] = s->amp[b]; D.pwr[b] = s->pwr[b];
  }
  D.noise_amp = s->noise_amp; D.noise_seed = s->noise_seed;
  D.u0 = s->u0; D.v0 = s->v0; D.igw = c->cfg.data_spec == 4;
  D.y2d = s->y2d; D.perturb_amp = s->perturb_amp; D.perturb_seed = s->perturb_seed;
  D.xlen = c->cfg.xlen; D.ylen = c->cfg.ylen; D.zlen = c->cfg.zlen;
  double* d_modes = nullptr;
  if (s->nmodes > 0) {
    HIP_TRY(c, hipMalloc(&d_modes, sizeof(double) * 5 * s->nmodes));
    HIP_TRY(c, hipMemcpy(d_modes, s->modes, sizeof(double) * 5 * s->nmodes, hipMemcpyHostToDevice));
  }
  D.nmodes = s->nmodes; D.modes = d_modes;
  Grid G = c->grid();
  int64_t n = (int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz;
  int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  static_cast<const Variant*>(c->var)->init<<<blocks, 256, 0, c->stream>>>(d_state, G, D);
  hipError_t e = hipGetLastError();
  if (d_modes) {
    hipError_t e2 = hipStreamSynchronize(c->stream);
    (void)hipFree(d_modes);
    if (e == hipSuccess) e = e2;
  }
  if (e != hipSuccess) return fail(c, AWFL_E_HIP, std::string("init kernel: ") + hipGetErrorString(e));
  return AWFL_OK;
}

int awfl_set_time_step(awfl_ctx* c, double dt) {
  if (!c) return AWFL_E_ARG;
  c->dtInit = dt;
  return AWFL_OK;
}

int awfl_compute_time_step(awfl_ctx* c, const double* d_state, double cfl, double* dt) {
  if (!c) return AWFL_E_ARG;
  if (!dt) return fail(c, AWFL_E_ARG, "null dt");
  if (c->dtInit <= 0) {
    if (!d_state) return fail(c, AWFL_E_ARG, "null state");
    Grid G = c->grid();
    int64_t n = (int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz;
    int blocks = (int)std::min<int64_t>((n + RED_THREADS - 1) / RED_THREADS, c->n_red);
    k_dt_partial<<<blocks, RED_THREADS, 0, c->stream>>>(d_state, G, cfl, c->d_red);
    LAUNCH_CHECK(c);
    k_min_final<<<1, RED_THREADS, 0, c->stream>>>(c->d_red, blocks, c->d_red + 5 * c->n_red);
    LAUNCH_CHECK(c);
    double h = 0;
    HIP_TRY(c, hipMemcpyAsync(&h, c->d_red + 5 * c->n_red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->dtInit = h;
  }
  *dt = c->dtInit;
  return AWFL_OK;
}

int awfl_eos_pressure(awfl_ctx* c, const double* d_rt, double* d_p, int64_t n) {
  if (!c) return AWFL_E_ARG;
  if (!d_rt || !d_p || n < 0) return fail(c, AWFL_E_ARG, "bad eos arguments");
  if (n == 0) return AWFL_OK;
  int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  k_eos<<<blocks, 256, 0, c->stream>>>(d_rt, d_p, n);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int awfl_diagnostics(awfl_ctx* c, const double* d_state, double* d_out) {
  if (!c) return AWFL_E_ARG;
  if (!d_state || !d_out) return fail(c, AWFL_E_ARG, "null argument");
  int64_t n = (int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz;
  int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  k_diag<<<blocks, 256, 0, c->stream>>>(d_state, c->grid(), d_out);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int awfl_state_sums(awfl_ctx* c, const double* d_state, double* h_out) {
  if (!c) return AWFL_E_ARG;
  if (!d_state || !h_out) return fail(c, AWFL_E_ARG, "null argument");
  Grid G = c->grid();
  int64_t n = (int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz;
  int blocks = (int)std::min<int64_t>((n + RED_THREADS - 1) / RED_THREADS, c->n_red);
  k_sum_partial<<<blocks, RED_THREADS, 0, c->stream>>>(d_state, G, c->d_red);
  LAUNCH_CHECK(c);
  k_sum_final<<<1, 64, 0, c->stream>>>(c->d_red, blocks, c->d_red + 5 * c->n_red);
  LAUNCH_CHECK(c);
  HIP_TRY(c, hipMemcpyAsync(h_out, c->d_red + 5 * c->n_red, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return AWFL_OK;
}

}  // extern "C"


using SweepXFn = void (*)(const double*, double*, Grid, SweepK);
using SweepColFn = void (*)(const double*, double*, Grid, SweepK, int);

// launch one sweep with the given kernels (x: kx; y, z: kcol)
// (x_flux_lds: the in-place x kernel keeps its face fluxes apart from the line)
static int launch_sweep_with(awfl_ctx* c, const double* in, double* out, const SweepK& K, int dir, SweepXFn kx,
                             SweepColFn kcol, bool x_flux_lds = false) {
  Grid G = c->grid();
  if (dir == 0) {
    int nx = c->cfg.nx;
    int threads = line_threads(nx);
    size_t lds = sizeof(double) * (size_t)(5 * (nx + 2 * c->hs) + 10 * nx + (x_flux_lds ? 5 * (nx + 1) : 0));
    if (lds > 64 * 1024)
      HIP_TRY(c, hipFuncSetAttribute((const void*)kx, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    kx<<<c->cfg.ny * c->cfg.nz, threads, lds, c->stream>>>(in, out, G, K);
    LAUNCH_CHECK(c);
    return AWFL_OK;
  }
  int n = dir == 1 ? c->cfg.ny : c->cfg.nz;
  int other = dir == 1 ? c->cfg.nz : c->cfg.ny;
  int bx = (c->cfg.nx + 63) / 64;
  // segments of at most 32 cells, and shorter (down to 16 cells) where that
  // is needed for ~4 waves per SIMD (256 CUs x 4 SIMDs).  Longer segments on
  // large grids (256 cells at 512x512x256) measured 4-7% slower.
  int64_t layer = (int64_t)bx * other;
  int nseg = (int)std::max<int64_t>(1, std::min<int64_t>(std::max<int64_t>((4096 + layer - 1) / layer, (n + 31) / 32),
                                                           std::max(1, n / 16)));
  int seglen = (n + nseg - 1) / nseg;
  nseg = (n + seglen - 1) / seglen;
  dim3 grid(bx, other, nseg);
  kcol<<<grid, 64, 0, c->stream>>>(in, out, G, K, seglen);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

// one lane-split line sweep: a work-group of awfl_line::LS_THREADS per line
static int launch_line(awfl_ctx* c, const double* in, double* out, const SweepK& K, int dir, LineFn k,
                       bool inplace = false) {
  const int n = dir == 0 ? c->cfg.nx : dir == 1 ? c->cfg.ny : c->cfg.nz;
  const unsigned lines = (unsigned)((int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz / n);
  const size_t lds = sizeof(double) * awfl_line::line_lds_doubles(n, c->hs, c->cfg.ngll, inplace);
  if (lds > 160 * 1024) return fail(c, AWFL_E_UNSUPPORTED, "line too long for the lane-split sweep's LDS");
  if (lds > 64 * 1024)
    HIP_TRY(c, hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  k<<<lines, awfl_line::LS_THREADS, lds, c->stream>>>(in, out, c->grid(), K);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

// launch one sweep; tend selects the tendency output (layout [5][nz][ny][nx])
static int launch_sweep(awfl_ctx* c, const double* in, double* out, double dt, int dir, bool tend) {
  const Variant* v = static_cast<const Variant*>(c->var);
  if (v->sweep_line[dir][0])
    return launch_line(c, in, out, sweep_constants(c, dt, dir), dir, v->sweep_line[dir][tend ? 1 : 0]);
  return launch_sweep_with(c, in, out, sweep_constants(c, dt, dir), dir, v->sweep_x[tend ? 1 : 0],
                           dir == 1 ? v->sweep_y[tend ? 1 : 0] : v->sweep_z[tend ? 1 : 0]);
}

static int fill(awfl_ctx* c, double* p, int64_t n, double v) {
  int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  k_fill<<<blocks, 256, 0, c->stream>>>(p, n, v);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int launch_sweep_fp64(awfl_ctx* c, const double* in, double* out, double dt, int dir, bool tend) {
  return launch_sweep(c, in, out, dt, dir, tend);
}

extern "C" {

int awfl_sweep(awfl_ctx* c, const double* d_in, double* d_out, double dt, int dir) {
  if (!c) return AWFL_E_ARG;
  if (!d_in || !d_out) return fail(c, AWFL_E_ARG, "null buffer");
  if (d_in == d_out) return fail(c, AWFL_E_ARG, "sweeps run out of place: d_in and d_out alias");
  if (dir < 0 || dir > 2) return fail(c, AWFL_E_ARG, "dir must be 0, 1 or 2");
  if (dir == 1 && c->sim2d) {
    HIP_TRY(c, hipMemcpyAsync(d_out, d_in, sizeof(double) * c->state_elems(), hipMemcpyDeviceToDevice, c->stream));
    return AWFL_OK;
  }
  return launch_sweep(c, d_in, d_out, dt, dir, false);
}

int awfl_compute_tendencies(awfl_ctx* c, const double* d_state, double* d_tend, double dt, int split) {
  if (!c) return AWFL_E_ARG;
  if (!d_state || !d_tend) return fail(c, AWFL_E_ARG, "null buffer");
  if (split < 0 || split > 2) return fail(c, AWFL_E_ARG, "split must be 0, 1 or 2");
  static const int fwd[3] = {0, 1, 2}, bwd[3] = {2, 1, 0};
  int dir = c->dimSwitch ? fwd[split] : bwd[split];
  int rc;
  if (dir == 1 && c->sim2d) rc = fill(c, d_tend, c->tend_elems(), 0.0);   // S:572-575
  else rc = launch_sweep(c, d_state, d_tend, dt, dir, true);
  if (rc != AWFL_OK) return rc;
  if (split == 2) c->dimSwitch = !c->dimSwitch;     // S:593
  return AWFL_OK;
}

int awfl_apply_tendencies(awfl_ctx* c, double* d_state, const double* d_tend, double dt, int split) {
  if (!c) return AWFL_E_ARG;
  if (!d_state || !d_tend) return fail(c, AWFL_E_ARG, "null buffer");
  if (split < 0 || split > 2) return fail(c, AWFL_E_ARG, "split must be 0, 1 or 2");
  Grid G = c->grid();
  int64_t n = c->tend_elems();
  int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  k_apply<<<blocks, 256, 0, c->stream>>>(d_state, d_tend, G, dt);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int awfl_time_step_branches(awfl_ctx* c, double* d_state, double dt, unsigned char* d_branch, int mode) {
  if (!c) return AWFL_E_ARG;
  if (!d_state || !d_branch) return fail(c, AWFL_E_ARG, "null state or branch buffer");
  if (mode != 1 && mode != 2) return fail(c, AWFL_E_ARG, "mode must be 1 (record) or 2 (replay)");
#ifdef AWFL_LEAN
  (void)dt;
  return fail(c, AWFL_E_UNSUPPORTED, "branch record / replay is not built into parity builds");
#else
  if (c->cfg.ord != 5 || c->cfg.ngll != 3)
    return fail(c, AWFL_E_UNSUPPORTED, "branch record / replay is compiled for (ord, ngll) = (5, 3) only");
  for (int b = 0; b < 2; b++)
    if (!c->d_buf[b]) HIP_TRY(c, hipMalloc(&c->d_buf[b], sizeof(double) * c->state_elems()));
  static const int fwd[3] = {0, 1, 2}, bwd[3] = {2, 1, 0};
  const int64_t per_split = (int64_t)(c->cfg.nx + 1) * (c->cfg.ny + 1) * (c->cfg.nz + 1);
  int dirs[3], slot[3], nd = 0;
  for (int s = 0; s < 3; s++) {
    int d = c->dimSwitch ? fwd[s] : bwd[s];
    if (d == 1 && c->sim2d) continue;
    slot[nd] = s;
    dirs[nd++] = d;
  }
  const double* cur = d_state;
  for (int s = 0; s < nd; s++) {
    double* dst = (s == nd - 1) ? d_state : (cur == c->d_buf[0] ? c->d_buf[1] : c->d_buf[0]);
    SweepK K = sweep_constants(c, dt, dirs[s]);
    K.br = d_branch + per_split * slot[s];
    int rc = mode == 1 ? launch_sweep_with(c, cur, dst, K, dirs[s], k_sweep_x<5, 3, false, 1>,
                                           dirs[s] == 1 ? k_sweep_col<5, 3, 1, false, 1> : k_sweep_col<5, 3, 2, false, 1>)
                       : launch_sweep_with(c, cur, dst, K, dirs[s], k_sweep_x<5, 3, false, 2>,
                                           dirs[s] == 1 ? k_sweep_col<5, 3, 1, false, 2> : k_sweep_col<5, 3, 2, false, 2>);
    if (rc != AWFL_OK) return rc;
    cur = dst;
  }
  c->dimSwitch = !c->dimSwitch;
  return AWFL_OK;
#endif
}

int awfl_time_step_noalloc(awfl_ctx* c, double* d_state, double dt) {
  if (!c) return AWFL_E_ARG;
  if (!d_state) return fail(c, AWFL_E_ARG, "null state");
  const int64_t bytes = (int64_t)sizeof(double) * c->state_elems();
#ifdef AWFL_LEAN
  return fail(c, AWFL_E_UNSUPPORTED, "awfl_time_step_noalloc is not built into parity builds");
#endif
  if (bytes > (int64_t)AWFL_NOALLOC_MAX_BYTES)
    return fail(c, AWFL_E_ARG, "awfl_time_step_noalloc: a state of " + std::to_string(bytes) +
                                   " bytes exceeds its scratch of " + std::to_string((long long)AWFL_NOALLOC_MAX_BYTES) +
                                   " bytes; use awfl_time_step");
  const int64_t lds = (int64_t)sizeof(double) * (5 * (c->cfg.nx + 2 * c->hs) + 15 * c->cfg.nx + 5);
  if (lds > 160 * 1024) return fail(c, AWFL_E_UNSUPPORTED, "awfl_time_step_noalloc: nx too large for the in-place x line");
  const Variant* v = static_cast<const Variant*>(c->var);
  if (!c->noalloc) return fail(c, AWFL_E_HIP, "awfl_time_step_noalloc: no scratch address: " + c->noalloc_err);
  int dev = -1;
  HIP_TRY(c, hipGetDevice(&dev));
  if (dev != c->device) return fail(c, AWFL_E_ARG, "awfl_time_step_noalloc: the context's device is not current");
  int dirs[3], nd = 0;
  static const int fwd[3] = {0, 1, 2}, bwd[3] = {2, 1, 0};
  for (int s = 0; s < 3; s++) {
    int d = c->dimSwitch ? fwd[s] : bwd[s];
    if (d == 1 && c->sim2d) continue;                // tendency is zero (S:572-575)
    dirs[nd++] = d;
  }
  // One scratch state S.  Two sweeps: state -> S -> state.  Three: the x
  // sweep runs in place (its work-group owns the whole line), so x, y, z is
  // state -> state -> S -> state and z, y, x is state -> S -> state -> state.
  double* S = c->noalloc;
  for (int s = 0; s < nd; s++) {
    const int d = dirs[s];
    const double* src;
    double* dst;
    if (nd == 2) {
      src = s == 0 ? d_state : S;
      dst = s == 0 ? S : d_state;
    } else if (d == 0) {
      src = dst = d_state;
    } else {
      const bool first_after_state = (dirs[0] == 0) ? s == 1 : s == 0;
      src = first_after_state ? d_state : S;
      dst = first_after_state ? S : d_state;
    }
    SweepK K = sweep_constants(c, dt, d);
    int rc;
    if (v->sweep_line[d][0])          // (9, 5): line kernels
      rc = (d == 0 && src == dst) ? launch_line(c, src, dst, K, 0, v->sweep_x_inplace, true)
                                  : launch_line(c, src, dst, K, d, v->sweep_line[d][0]);
    else
      rc = (d == 0 && src == dst)
               ? launch_sweep_with(c, src, dst, K, 0, v->sweep_x_inplace, nullptr, true)
               : launch_sweep_with(c, src, dst, K, d, v->sweep_x[0], d == 1 ? v->sweep_y[0] : v->sweep_z[0]);
    if (rc != AWFL_OK) return rc;
  }
  c->dimSwitch = !c->dimSwitch;
  return AWFL_OK;
}

int awfl_time_step(awfl_ctx* c, double* d_state, double dt) {
  if (!c) return AWFL_E_ARG;
  if (!d_state) return fail(c, AWFL_E_ARG, "null state");
  for (int b = 0; b < 2; b++)
    if (!c->d_buf[b]) HIP_TRY(c, hipMalloc(&c->d_buf[b], sizeof(double) * c->state_elems()));
  int dirs[3], nd = 0;
  static const int fwd[3] = {0, 1, 2}, bwd[3] = {2, 1, 0};
  for (int s = 0; s < 3; s++) {
    int d = c->dimSwitch ? fwd[s] : bwd[s];
    if (d == 1 && c->sim2d) continue;                // tendency is zero (S:572-575)
    dirs[nd++] = d;
  }
  const double* cur = d_state;
  for (int s = 0; s < nd; s++) {
    double* dst = (s == nd - 1) ? d_state : (cur == c->d_buf[0] ? c->d_buf[1] : c->d_buf[0]);
    const bool timed = 2 * (size_t)c->t_n + 1 < c->t_ev.size();
    if (timed) HIP_TRY(c, hipEventRecord(c->t_ev[2 * c->t_n], c->stream));
    int rc = launch_sweep(c, cur, dst, dt, dirs[s], false);
    if (rc != AWFL_OK) return rc;
    if (timed) {
      HIP_TRY(c, hipEventRecord(c->t_ev[2 * c->t_n + 1], c->stream));
      c->t_dir[c->t_n++] = dirs[s];
    }
    cur = dst;
  }
  c->dimSwitch = !c->dimSwitch;
  return AWFL_OK;
}

static void timing_free(awfl_ctx* c) {
  for (hipEvent_t e : c->t_ev) (void)hipEventDestroy(e);
  c->t_ev.clear();
  c->t_dir.clear();
  c->t_n = 0;
}

int awfl_sweep_timing(awfl_ctx* c, int max_sweeps) {
  if (!c) return AWFL_E_ARG;
  if (max_sweeps < 0) return fail(c, AWFL_E_ARG, "max_sweeps must be >= 0");
  timing_free(c);
  if (max_sweeps == 0) return AWFL_OK;
  c->t_ev.resize(2 * (size_t)max_sweeps);
  c->t_dir.resize(max_sweeps);
  for (size_t i = 0; i < c->t_ev.size(); i++) {
    hipError_t e = hipEventCreate(&c->t_ev[i]);
    if (e != hipSuccess) {
      c->t_ev.resize(i);
      timing_free(c);
      return fail(c, AWFL_E_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
    }
  }
  return AWFL_OK;
}

int awfl_sweep_times(awfl_ctx* c, double ms[3], int64_t count[3]) {
  if (!c || !ms || !count) return AWFL_E_ARG;
  for (int d = 0; d < 3; d++) {
    ms[d] = 0;
    count[d] = 0;
  }
  if (c->t_n > 0) HIP_TRY(c, hipEventSynchronize(c->t_ev[2 * c->t_n - 1]));
  for (int t = 0; t < c->t_n; t++) {
    float m = 0;
    HIP_TRY(c, hipEventElapsedTime(&m, c->t_ev[2 * t], c->t_ev[2 * t + 1]));
    ms[c->t_dir[t]] += m;
    count[c->t_dir[t]]++;
  }
  for (int d = 0; d < 3; d++)
    if (count[d]) ms[d] /= (double)count[d];
  c->t_n = 0;
  return AWFL_OK;
}

}  // extern "C"
