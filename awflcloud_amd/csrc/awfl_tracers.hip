// awfl_tracers.hip -- tracer transport with the split sweeps (C ABI
// awfl_set_tracers / awfl_sweep_tracers / awfl_compute_tendencies_tracers /
// awfl_apply_tracer_tendencies / awfl_time_step_tracers in
// include/awfl_dycore.h).
//
// The reference carries tracers as masses rho*q next to the state and
// compiles every tracer block out (#define NO_TRACERS, S:10, with S =
// Spatial_euler3d_cons_expl_cart_fv_Agrid.h).  The blocks it would run, per
// split direction:
//   divide the tracer masses by the full density (S:622-627) and fill the
//   halos of the mixing ratio (periodic copy, or the end cell at a wall);
//   per cell, after the state's differential transform: reconstruct the
//   mixing ratio to the GLL points, multiply by the density there, clip at
//   zero for a positive tracer, form the flux rho*u*q, run the tracer
//   transform diffTransformTracer (S:2095-2126) on the density and normal
//   momentum levels of the state, time-average the edges and clip again
//   (S:760-809);
//   per face: the state's mass flux times rho*q/rho of the upwind edge,
//   upwinded with the Riemann flux's normal velocity (S:889-901);
//   a positivity flux limiter for positive tracers (S:904-935);
//   the flux-difference tendency, and the tracer masses multiplied back
//   (S:948-955).
// Here one work-group runs one line along the sweep direction with the line
// in LDS, state and tracers together; the state part runs the production
// sweep's statements (cell_edges / riemann), so the state it returns is
// awfl_sweep's up to the compiler's contraction choices in this kernel.
// Departures from the source text, where it cannot run as written:
//   T1 the limiter forms every cell's reduction factor from the unlimited
//      fluxes and then applies it to the faces it is upwind of; the
//      reference scales tracerFlux in place while neighbouring faces read it
//      (an order-dependent result);
//   T2 the y limiter's v < 0 branch reads the y faces of the upwind cell (the
//      source indexes tracerFlux(tr,k,j,ind_j), S:1283-1284);
//   T3 the z limiter wraps its upwind index on a periodic z axis (the source
//      would read k = -1 and nz+1, S:1657-1666);
//   T4 the tracer masses are not rounded through the mixing ratio and back:
//      the update adds dt * tendency to the masses as given.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "awfl_cell.hpp"
#include "awfl_internal.hpp"
#include "awfl_profiles.hpp"

using namespace awfl;

namespace {

constexpr double TRACER_EPS = 1.e-10;   // S:906

__device__ __forceinline__ int axis_n(const Grid& G, int dir) { return dir == 0 ? G.nx : dir == 1 ? G.ny : G.nz; }

// stored cell of logical index a along an axis of n cells (periodic wrap or
// wall clamp); zn marks a clamped slot
__device__ __forceinline__ int wrap_axis(int a, int n, bool per, bool& zn) {
  zn = false;
  if (per) {
    while (a < 0) a += n;
    while (a >= n) a -= n;
    return a;
  }
  if (a < 0) {
    zn = true;
    return 0;
  }
  if (a >= n) {
    zn = true;
    return n - 1;
  }
  return a;
}

// One line of a DIR sweep with ntr tracers.  LDS (doubles), n cells, W =
// n + 2hs, NF = n + 1 faces:
//   state line [5][W] (then the state fluxes [5][NF]), mixing ratios
//   [ntr][W] (then the tracer fluxes [ntr][NF]), state edges [10][n], tracer
//   edges [2 ntr][n] (then the limiter's cell factors), cell densities [n],
//   and for z the sources [n].  The limiter forms the direction of a face
//   again from the state edges, which stay in place to the end.
// Divisions by the density and by the edge densities are multiplications by
// one reciprocal (rcp, ~1 ulp) shared by all tracers.
//
// With one cell per thread (every line up to 256 cells) a thread issues all
// loads of its cell and of one halo cell before its first LDS store (one
// memory round trip to stage the line), and the state and tracer values of
// its cell that the limiter and the update read again are loaded when the
// faces start, in flight behind them.  Register arrays per tracer have
// AWFL_MAX_TRACERS entries, each used under `t < ntr`.  (Compiling the
// kernel for exactly three tracers was measured slower: the unrolled
// tracer loops took another 16 VGPRs, one wave per SIMD fewer.)
template <int ORD, int NGLL, int DIR, bool TEND>
__device__ __forceinline__ void sweep_tr(const double* __restrict__ in, const double* __restrict__ tr_in,
                                         double* __restrict__ out, double* __restrict__ tr_out, const Grid& G,
                                         const SweepK& K, int ntr, unsigned pos) {
  constexpr int MT = AWFL_MAX_TRACERS;
  constexpr int hs = (ORD - 1) / 2;
  constexpr int NM = 1 + DIR;
  constexpr int NA = NGLL;
  extern __shared__ double sm[];
  const int n = axis_n(G, DIR);
  const int W = n + 2 * hs;
  const int NF = n + 1;
  // line -> first interior cell, stride along DIR, level of an x or y line
  const unsigned line = tile_index();
  int64_t base, stride;
  int kl = 0, jl = 0, il = 0;
  if constexpr (DIR == 0) {
    jl = line % G.ny;
    kl = line / G.ny;
    base = (int64_t)(hs + kl) * G.sz + (int64_t)(hs + jl) * G.sy + hs;
    stride = 1;
  } else if constexpr (DIR == 1) {
    il = line % G.nx;
    kl = line / G.nx;
    base = (int64_t)(hs + kl) * G.sz + (int64_t)hs * G.sy + hs + il;
    stride = G.sy;
  } else {
    il = line % G.nx;
    jl = line / G.nx;
    base = (int64_t)hs * G.sz + (int64_t)(hs + jl) * G.sy + hs + il;
    stride = G.sz;
  }
  double* s_st = sm;                   // [5][W]
  double* s_F = sm;                    // [5][NF], over the state line once the cells are done
  double* s_tq = s_st + 5 * W;         // [ntr][W]
  double* s_TF = s_tq;                 // [ntr][NF], over the mixing ratios once the cells are done
  double* s_e = s_tq + ntr * W;        // [10][n]
  double* s_te = s_e + 10 * n;         // [2 ntr][n]
  double* s_dens = s_te + 2 * ntr * n; // [n]
  double* s_src = s_dens + n;          // [n], z only
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  auto level = [&](int c) { return DIR == 2 ? c : kl; };

  const int tid = threadIdx.x;
  const bool one = n <= (int)blockDim.x;

  // 1. the state line (normal momentum zero in wall ghosts), the full
  // density of every cell (S:624-626) and the tracers' mixing ratios (ghosts
  // copy the end cell at a wall)
  if (one) {
    // ts: the thread index behind an empty asm, so that the compiler shares
    // no address arithmetic between the stage and the cells (it would keep it
    // in registers through them)
    int ts = tid;
    asm volatile("" : "+v"(ts));
    const int hc = ts < hs ? ts : n + ts;  // padded slot of halo cell ts (ts < 2hs)
    double v[5], q[MT], hv[5], hq[MT], bgc = 0.0, bgh = 0.0;
    bool hzn = false;
    if (ts < n) {
      sfor<5>([&](auto l) { v[l] = in[l * G.sl + base + ts * stride]; });
      sfor<MT>([&](auto t) {
        if (t < ntr) q[t] = tr_in[t * G.sl + base + ts * stride];
      });
      bgc = G.hyDensCells[hs + level(ts)];
    }
    if (ts < 2 * hs) {
      const int src = wrap_axis(hc - hs, n, per, hzn);
      sfor<5>([&](auto l) { hv[l] = in[l * G.sl + base + src * stride]; });
      sfor<MT>([&](auto t) {
        if (t < ntr) hq[t] = tr_in[t * G.sl + base + src * stride];
      });
      bgh = G.hyDensCells[hs + level(src)];
    }
    if (ts < n) {
      const double dens = v[idR] + bgc;
      s_dens[ts] = dens;
      sfor<5>([&](auto l) { s_st[l * W + hs + ts] = v[l]; });
      const double rr = rcp(dens);
      sfor<MT>([&](auto t) {
        if (t < ntr) s_tq[t * W + hs + ts] = q[t] * rr;
      });
    }
    if (ts < 2 * hs) {
      sfor<5>([&](auto l) { s_st[l * W + hc] = (hzn && l == NM) ? 0.0 : hv[l]; });
      const double rr = rcp(hv[idR] + bgh);
      sfor<MT>([&](auto t) {
        if (t < ntr) s_tq[t * W + hc] = hq[t] * rr;
      });
    }
  } else {
    for (int c = tid; c < W; c += blockDim.x) {
      bool zn;
      const int src = wrap_axis(c - hs, n, per, zn);
      sfor<5>([&](auto l) { s_st[l * W + c] = (zn && l == NM) ? 0.0 : in[l * G.sl + base + src * stride]; });
    }
    __syncthreads();
    for (int c = tid; c < n; c += blockDim.x) s_dens[c] = s_st[idR * W + c + hs] + G.hyDensCells[hs + level(c)];
    __syncthreads();
    for (int c = tid; c < W; c += blockDim.x) {
      bool zn;
      const int src = wrap_axis(c - hs, n, per, zn);
      const double rr = rcp(s_dens[src]);
      for (int t = 0; t < ntr; t++) s_tq[t * W + c] = tr_in[t * G.sl + base + src * stride] * rr;
    }
  }
  __syncthreads();

  // 2. cells: state edges (and z source), then per tracer its edges from the
  // density and normal-momentum levels of the state's transform
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    CellBg<NGLL> bg{};
    if constexpr (DIR == 2) {
      sfor<NGLL>([&](auto ii) {
        bg.rG[ii] = G.hyDensGLL[c * NGLL + ii];
        bg.rtG[ii] = G.hyDensThetaGLL[c * NGLL + ii];
        bg.pG[ii] = G.hyPressureGLL[c * NGLL + ii];
        bg.ptG[ii] = G.prG[c * NGLL + ii];
      });
      bg.wall_bot = !per && c == 0;
      bg.wall_top = !per && c == n - 1;
    } else {
      bg.r = G.hyDensCells[hs + kl];
      bg.rt = G.hyDensThetaCells[hs + kl];
      bg.pt = G.prT[kl];
      bg.wall_bot = bg.wall_top = false;
    }
    double st[5][ORD], eL[5], eR[5], sa;
    double rl[NA][NGLL], nl[NA][NGLL];
    sfor<5>([&](auto l) { sfor<ORD>([&](auto s) { st[l][s] = s_st[l * W + c + s]; }); });
    cell_edges<ORD, NGLL, DIR>(st, bg, G.weno_s, G.weno_w, K, eL, eR, sa, [&](const auto& q) {
      sfor<NA>([&](auto kt) {
        sfor<NGLL>([&](auto ii) {
          rl[kt][ii] = q[idR][kt][ii];
          nl[kt][ii] = q[NM][kt][ii];
        });
      });
    });
    sfor<5>([&](auto l) {
      s_e[l * n + c] = eL[l];
      s_e[(5 + l) * n + c] = eR[l];
    });
    if constexpr (DIR == 2) s_src[c] = sa;
    double rr0[NGLL];
    sfor<NGLL>([&](auto ii) { rr0[ii] = rcp(rl[0][ii]); });
    for (int t = 0; t < ntr; t++) {
      const bool tp = (pos >> t) & 1u;
      double u[ORD], g[NGLL];
      sfor<ORD>([&](auto s) { u[s] = s_tq[t * W + c + s]; });
      reconstruct<ORD, NGLL>(u, G.weno_s, g);
      double rt[NA][NGLL], rnt[NA][NGLL];
      sfor<NGLL>([&](auto ii) {
        rt[0][ii] = g[ii] * rl[0][ii];
        if (tp) rt[0][ii] = fmax(0.0, rt[0][ii]);
        rnt[0][ii] = rt[0][ii] * nl[0][ii] * rr0[ii];
      });
      // diffTransformTracer (S:2095-2126)
      sfor<NA - 1>([&](auto kt) {
        sfor<NGLL>([&](auto ii) {
          double df = 0.0;
          sfor<NGLL>([&](auto s) { df += Deriv<NGLL>::m[s][ii] * rnt[kt][s]; });
          rt[kt + 1][ii] = df * K.ndt[kt];
        });
        if constexpr (kt + 2 < NA) {     // the flux at the last level is not used
          sfor<NGLL>([&](auto ii) {
            double tot = 0.0;
            sfor<kt + 2>([&](auto ir) {
              if constexpr (ir == 0) tot += nl[ir][ii] * rt[kt + 1 - ir][ii];
              else tot += nl[ir][ii] * rt[kt + 1 - ir][ii] - rl[ir][ii] * rnt[kt + 1 - ir][ii];
            });
            rnt[kt + 1][ii] = tot * rr0[ii];
          });
        }
      });
      double a0 = rt[0][0], a1 = rt[0][NGLL - 1];
      sfor<NA - 1>([&](auto km) {
        constexpr int kt = km + 1;
        a0 += rt[kt][0] * K.tav[kt];
        a1 += rt[kt][NGLL - 1] * K.tav[kt];
      });
      if (tp) {
        a0 = fmax(0.0, a0);
        a1 = fmax(0.0, a1);
      }
      s_te[(2 * t) * n + c] = a0;
      s_te[(2 * t + 1) * n + c] = a1;
    }
  }
  __syncthreads();

  // 3. faces: state flux, tracer fluxes upwinded with the state's normal
  // velocity, and the limiter's direction (the normal momentum average,
  // S:907).  A periodic x or y line's face n is its face 0; a periodic z
  // line keeps both, whose fluxes differ by the background pressure each
  // subtracts (S:1628-1629).
  const bool wrapF = per && DIR != 2;
  const int nfaces = wrapF ? n : n + 1;
  // this thread's cell as it came in, for the limiter and the update; the
  // empty asm keeps these loads after the barrier (issued earlier, their
  // registers would be live through the cells)
  double kin[5], ktr[MT];
  if (one && tid < n) {
    int64_t o = base + tid * stride;
    asm volatile("" : "+v"(o));
    sfor<5>([&](auto l) { kin[l] = in[l * G.sl + o]; });
    sfor<MT>([&](auto t) {
      if (t < ntr) ktr[t] = tr_in[t * G.sl + o];
    });
  }
  for (int f = threadIdx.x; f < nfaces; f += blockDim.x) {
    const bool wl = !per && f == 0, wr = !per && f == n;
    const int a = wl ? 0 : (f == 0 ? n - 1 : f - 1);
    const int b = wr ? n - 1 : (f == n ? 0 : f);
    double L[5], R[5], F[5], un;
    sfor<5>([&](auto l) {
      L[l] = wl ? s_e[l * n + 0] : s_e[(5 + l) * n + a];
      R[l] = wr ? s_e[(5 + l) * n + n - 1] : s_e[l * n + b];
    });
    FaceRef fr;
    if constexpr (DIR == 0) {
      fr.rt = fr.q5 = G.prT[kl];
    } else if constexpr (DIR == 1) {
      fr.rt = G.prRT[kl];
      fr.q5 = G.prT[kl];
    } else {
      fr.rt = fr.q5 = G.prF[f];
    }
    riemann<DIR>(L, R, fr, F, un);
    if constexpr (DIR == 2) {          // S:1628-1629
      if (f < G.nz) F[idW] -= G.hyPressureGLL[f * NGLL + 0];
      else F[idW] -= G.hyPressureGLL[(G.nz - 1) * NGLL + NGLL - 1];
    }
    sfor<5>([&](auto l) { s_F[l * NF + f] = F[l]; });
    const double up = un > 0 ? F[0] * rcp(L[0]) : F[0] * rcp(R[0]);
    for (int t = 0; t < ntr; t++) {
      // rho*q of the upwind side: the low cell's high edge, or the high
      // cell's low edge (a wall face takes both from its one cell)
      const double tL = wl ? s_te[(2 * t) * n + 0] : s_te[(2 * t + 1) * n + a];
      const double tR = wr ? s_te[(2 * t + 1) * n + n - 1] : s_te[(2 * t) * n + b];
      s_TF[t * NF + f] = (un > 0 ? tL : tR) * up;
    }
  }
  __syncthreads();

  // 4. positivity limiter (T1): cell factors from the unlimited fluxes over
  // the tracer-edge region, then each face takes its upwind cell's factor
  if (pos) {
    double* s_fac = s_te;              // [ntr][n]
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      const int cp = (wrapF && c + 1 == n) ? 0 : c + 1;
      sfor<MT>([&](auto t) {
        if (t < ntr && ((pos >> t) & 1u)) {
          const double f1 = fmin(s_TF[t * NF + c], 0.0);
          const double f2 = fmax(s_TF[t * NF + cp], 0.0);
          const double fluxOut = K.dt * (f2 - f1) * K.rd;
          // the cell's mixing ratio times its density, as staged in step 1
          const double m = one ? ktr[t] : tr_in[t * G.sl + base + c * stride];
          const double q = m * rcp(s_dens[c]);
          s_fac[t * n + c] = fmin(1.0, q * s_dens[c] * rcp(fluxOut + TRACER_EPS));
        }
      });
    }
    __syncthreads();
    for (int f = threadIdx.x; f < nfaces; f += blockDim.x) {
      // the normal momentum average of the face's edge states (S:907)
      const bool wl = !per && f == 0, wr = !per && f == n;
      const int fa = wl ? 0 : (f == 0 ? n - 1 : f - 1);
      const int fb = wr ? n - 1 : (f == n ? 0 : f);
      const double uL = wl ? s_e[NM * n + 0] : s_e[(5 + NM) * n + fa];
      const double uR = wr ? s_e[(5 + NM) * n + n - 1] : s_e[NM * n + fb];
      const double u = 0.5 * (uL + uR);
      int a = -1;
      if (u > 0) a = f == 0 ? n - 1 : f - 1;
      else if (u < 0) a = f == n ? 0 : f;
      if (a < 0) continue;
      for (int t = 0; t < ntr; t++)
        if ((pos >> t) & 1u) s_TF[t * NF + f] *= s_fac[t * n + a];
    }
    __syncthreads();
  }

  // 5. tendencies: the state as the production sweep forms them, the
  // tracers as -(flux difference) / d (S:950) by the same reciprocal
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    const int cp = (wrapF && c + 1 == n) ? 0 : c + 1;
    int k = kl, j = jl, i = il;
    if constexpr (DIR == 0) i = c;
    else if constexpr (DIR == 1) j = c;
    else k = c;
    const int64_t o = base + c * stride;
    const int64_t ot = ((int64_t)k * G.ny + j) * G.nx + i;
    const int64_t nc = (int64_t)G.nx * G.ny * G.nz;
    sfor<5>([&](auto l) {
      double td = -(s_F[l * NF + cp] - s_F[l * NF + c]) * K.rd;
      if (DIR == 2 && l == idW) td = s_src[c] + td;
      if (G.sim2d && l == idV) td = 0.0;
      if constexpr (TEND) out[l * nc + ot] = td;
      else out[l * G.sl + o] = (one ? kin[l] : in[l * G.sl + o]) + K.dt * td;
    });
    sfor<MT>([&](auto t) {
      if (t < ntr) {
        const double td = -(s_TF[t * NF + cp] - s_TF[t * NF + c]) * K.rd;
        if constexpr (TEND) tr_out[t * nc + ot] = td;
        else tr_out[t * G.sl + o] = (one ? ktr[t] : tr_in[t * G.sl + o]) + K.dt * td;
      }
    });
  }
}

// Entry points.  The z kernel carries a bound of three waves per SIMD: its
// registers otherwise come to 171, two waves.  The x and y kernels stay
// under 168 without it and were measured faster without it.
template <int ORD, int NGLL, int DIR, bool TEND>
__global__ __launch_bounds__(256) void k_sweep_tr(const double* __restrict__ in, const double* __restrict__ tr_in,
                                                  double* __restrict__ out, double* __restrict__ tr_out, Grid G,
                                                  SweepK K, int ntr, unsigned pos) {
  sweep_tr<ORD, NGLL, DIR, TEND>(in, tr_in, out, tr_out, G, K, ntr, pos);
}
template <int ORD, int NGLL, bool TEND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_sweep_tr_z(
    const double* __restrict__ in, const double* __restrict__ tr_in, double* __restrict__ out,
    double* __restrict__ tr_out, Grid G, SweepK K, int ntr, unsigned pos) {
  sweep_tr<ORD, NGLL, 2, TEND>(in, tr_in, out, tr_out, G, K, ntr, pos);
}

// tracers += dt * tracer_tend over the interior (Temporal_ader.h:67-69)
__global__ void k_apply_tr(double* __restrict__ tr, const double* __restrict__ ttend, Grid G, int ntr, double dt) {
  const int64_t nc = (int64_t)G.nx * G.ny * G.nz;
  const int64_t total = nc * ntr;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = e / nc, r = e - t * nc;
    const int i = (int)(r % G.nx);
    const int64_t q = r / G.nx;
    const int j = (int)(q % G.ny), k = (int)(q / G.ny);
    const int64_t o = t * G.sl + (int64_t)(G.hs + k) * G.sz + (int64_t)(G.hs + j) * G.sy + G.hs + i;
    tr[o] = tr[o] + dt * ttend[e];
  }
}

// initTracers (S:505-560): tracer 0 uniform, 1 a replica of the thermal's
// theta bubble, 2 a block, each times the constant-theta (300 K) density and
// integrated with the ord^3 GLL quadrature; later tracers zero
template <int ORD>
__global__ void k_init_tr(double* __restrict__ tr, Grid G, int ntr, double xlen, double ylen) {
  constexpr int hs = (ORD - 1) / 2;
  const int64_t n = (int64_t)G.nz * G.ny * G.nx;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int i = t % G.nx;
    const int64_t r = t / G.nx;
    const int j = r % G.ny, k = r / G.ny;
    const int64_t o = (int64_t)(hs + k) * G.sz + (int64_t)(hs + j) * G.sy + hs + i;
    bool inside = k >= 1 * G.nz / 10 && k < 3 * G.nz / 10 && i >= 4 * G.nx / 10 && i < 6 * G.nx / 10;
    if (!G.sim2d) inside = inside && j >= 4 * G.ny / 10 && j < 6 * G.ny / 10;
    for (int l = 0; l < ntr; l++) {
      double acc = 0;
      if (l < 3) {
        for (int kk = 0; kk < ORD; kk++)
          for (int jj = 0; jj < ORD; jj++)
            for (int ii = 0; ii < ORD; ii++) {
              const double zloc = (k + 0.5) * G.dz + Gll<ORD>::pts[kk] * G.dz;
              const double yloc = G.sim2d ? ylen / 2 : (j + 0.5) * G.dy + Gll<ORD>::pts[jj] * G.dy;
              const double xloc = (i + 0.5) * G.dx + Gll<ORD>::pts[ii] * G.dx;
              const double wt = Gll<ORD>::wts[kk] * Gll<ORD>::wts[jj] * Gll<ORD>::wts[ii];
              const double rh = const_theta_density(300., zloc);
              double tval = 1;
              if (l == 1) tval = ellipsoid_linear(xloc, yloc, zloc, xlen / 2, ylen / 2, 2000, 2000, 2000, 2000, 2);
              else if (l == 2) tval = inside ? 1 : 0;
              acc += rh * tval * wt;
            }
      }
      tr[l * G.sl + o] = acc;
    }
  }
}

__global__ void k_fill_tr(double* __restrict__ p, int64_t n, double v) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) p[e] = v;
}

using KernelTr = void (*)(const double*, const double*, double*, double*, Grid, SweepK, int, unsigned);

struct VariantTr {
  int ord, ngll;
  KernelTr upd[3], tend[3];
  void (*init)(double*, Grid, int, double, double);
};

#define AWFL_VARIANT_TR(O, N)                                                                  \
  VariantTr {                                                                                  \
    O, N, {k_sweep_tr<O, N, 0, false>, k_sweep_tr<O, N, 1, false>, k_sweep_tr_z<O, N, false>}, \
        {k_sweep_tr<O, N, 0, true>, k_sweep_tr<O, N, 1, true>, k_sweep_tr_z<O, N, true>},       \
        k_init_tr<O>,                                                                          \
  }

const VariantTr kVariantsTr[] = {AWFL_VARIANT_TR(3, 2), AWFL_VARIANT_TR(5, 3), AWFL_VARIANT_TR(9, 5)};

const VariantTr* find_variant_tr(int ord, int ngll) {
  for (const VariantTr& v : kVariantsTr)
    if (v.ord == ord && v.ngll == ngll) return &v;
  return nullptr;
}

int64_t tracer_elems(const awfl_ctx* c) { return (int64_t)c->ntr * (c->state_elems() / 5); }
int64_t tracer_tend_elems(const awfl_ctx* c) { return (int64_t)c->ntr * c->cfg.nx * c->cfg.ny * c->cfg.nz; }

// one sweep of state and tracers along dir: update (out = in + dt tend) or
// the tendencies themselves
int launch_sweep_tr(awfl_ctx* c, const double* in, const double* tr_in, double* out, double* tr_out, double dt,
                    int dir, bool tend) {
  const VariantTr* v = find_variant_tr(c->cfg.ord, c->cfg.ngll);
  if (!v) return fail(c, AWFL_E_UNSUPPORTED, "tracers: (ord, ngll) not compiled in");
  const int n = dir == 0 ? c->cfg.nx : dir == 1 ? c->cfg.ny : c->cfg.nz;
  const int W = n + 2 * c->hs, ntr = c->ntr;
  const size_t lds = sizeof(double) * (size_t)(5 * W + ntr * W + 10 * n + 2 * ntr * n + (dir == 2 ? 2 : 1) * n);
  if (lds > 160 * 1024) return fail(c, AWFL_E_UNSUPPORTED, "tracers: line too long for the tracer sweep's LDS");
  KernelTr k = tend ? v->tend[dir] : v->upd[dir];
  if (lds > 64 * 1024)
    HIP_TRY(c, hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int64_t nc = (int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz;
  const unsigned lines = (unsigned)(nc / n);
  const int threads = line_threads(n);
  k<<<lines, threads, lds, c->stream>>>(in, tr_in, out, tr_out, c->grid(), sweep_constants(c, dt, dir), ntr,
                                        c->tpos);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int fill_tr(awfl_ctx* c, double* p, int64_t n) {
  if (n <= 0) return AWFL_OK;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  k_fill_tr<<<blocks, 256, 0, c->stream>>>(p, n, 0.0);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int need_tracers(awfl_ctx* c) {
  if (c->ntr < 1) return fail(c, AWFL_E_ARG, "no tracers set (awfl_set_tracers)");
  if (c->cfg.precision != AWFL_FP64) return fail(c, AWFL_E_UNSUPPORTED, "tracers run with the FP64 state");
  return AWFL_OK;
}

}  // namespace

extern "C" {

int awfl_set_tracers(awfl_ctx* c, int num_tracers, const int* h_positive) {
  if (!c) return AWFL_E_ARG;
  if (num_tracers < 0 || num_tracers > AWFL_MAX_TRACERS)
    return fail(c, AWFL_E_ARG, "num_tracers must be in [0, AWFL_MAX_TRACERS]");
  if (num_tracers > 0 && !h_positive) return fail(c, AWFL_E_ARG, "null positivity flags");
  if (!find_variant_tr(c->cfg.ord, c->cfg.ngll)) return fail(c, AWFL_E_UNSUPPORTED, "tracers: (ord, ngll) not compiled in");
  unsigned mask = 0;
  for (int t = 0; t < num_tracers; t++)
    if (h_positive[t]) mask |= 1u << t;
  if (num_tracers != c->ntr) {
    for (double*& b : c->t_buf) {
      if (b) (void)hipFree(b);
      b = nullptr;
    }
  }
  c->ntr = num_tracers;
  c->tpos = mask;
  return AWFL_OK;
}

int awfl_num_tracers(const awfl_ctx* c) { return c ? c->ntr : AWFL_E_ARG; }

int awfl_init_tracers(awfl_ctx* c, double* d_tracers) {
  if (!c) return AWFL_E_ARG;
  int rc = need_tracers(c);
  if (rc != AWFL_OK) return rc;
  if (!d_tracers) return fail(c, AWFL_E_ARG, "null buffer");
  const VariantTr* v = find_variant_tr(c->cfg.ord, c->cfg.ngll);
  if (!v) return fail(c, AWFL_E_UNSUPPORTED, "tracers: (ord, ngll) not compiled in");
  rc = fill_tr(c, d_tracers, tracer_elems(c));   // halos zero
  if (rc != AWFL_OK) return rc;
  const int64_t n = (int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz;
  const int blocks = (int)std::min<int64_t>((n + 127) / 128, 65536);
  v->init<<<blocks, 128, 0, c->stream>>>(d_tracers, c->grid(), c->ntr, c->cfg.xlen, c->cfg.ylen);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int awfl_sweep_tracers(awfl_ctx* c, const double* d_in, const double* d_tr_in, double* d_out, double* d_tr_out,
                       double dt, int dir) {
  if (!c) return AWFL_E_ARG;
  int rc = need_tracers(c);
  if (rc != AWFL_OK) return rc;
  if (!d_in || !d_tr_in || !d_out || !d_tr_out) return fail(c, AWFL_E_ARG, "null buffer");
  if (d_in == d_out || d_tr_in == d_tr_out) return fail(c, AWFL_E_ARG, "sweeps run out of place");
  if (dir < 0 || dir > 2) return fail(c, AWFL_E_ARG, "dir must be 0, 1 or 2");
  if (dir == 1 && c->sim2d) {
    HIP_TRY(c, hipMemcpyAsync(d_out, d_in, sizeof(double) * c->state_elems(), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(d_tr_out, d_tr_in, sizeof(double) * tracer_elems(c), hipMemcpyDeviceToDevice,
                              c->stream));
    return AWFL_OK;
  }
  return launch_sweep_tr(c, d_in, d_tr_in, d_out, d_tr_out, dt, dir, false);
}

int awfl_compute_tendencies_tracers(awfl_ctx* c, const double* d_state, const double* d_tracers, double* d_tend,
                                    double* d_tracer_tend, double dt, int split) {
  if (!c) return AWFL_E_ARG;
  int rc = need_tracers(c);
  if (rc != AWFL_OK) return rc;
  if (!d_state || !d_tracers || !d_tend || !d_tracer_tend) return fail(c, AWFL_E_ARG, "null buffer");
  if (split < 0 || split > 2) return fail(c, AWFL_E_ARG, "split must be 0, 1 or 2");
  static const int fwd[3] = {0, 1, 2}, bwd[3] = {2, 1, 0};
  const int dir = c->dimSwitch ? fwd[split] : bwd[split];
  if (dir == 1 && c->sim2d) {          // S:572-575
    rc = fill_tr(c, d_tend, c->tend_elems());
    if (rc == AWFL_OK) rc = fill_tr(c, d_tracer_tend, tracer_tend_elems(c));
  } else {
    rc = launch_sweep_tr(c, d_state, d_tracers, d_tend, d_tracer_tend, dt, dir, true);
  }
  if (rc != AWFL_OK) return rc;
  if (split == 2) c->dimSwitch = !c->dimSwitch;   // S:593
  return AWFL_OK;
}

int awfl_apply_tracer_tendencies(awfl_ctx* c, double* d_tracers, const double* d_tracer_tend, double dt) {
  if (!c) return AWFL_E_ARG;
  int rc = need_tracers(c);
  if (rc != AWFL_OK) return rc;
  if (!d_tracers || !d_tracer_tend) return fail(c, AWFL_E_ARG, "null buffer");
  const int64_t n = tracer_tend_elems(c);
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  k_apply_tr<<<blocks, 256, 0, c->stream>>>(d_tracers, d_tracer_tend, c->grid(), c->ntr, dt);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int awfl_time_step_tracers(awfl_ctx* c, double* d_state, double* d_tracers, double dt) {
  if (!c) return AWFL_E_ARG;
  int rc = need_tracers(c);
  if (rc != AWFL_OK) return rc;
  if (!d_state || !d_tracers) return fail(c, AWFL_E_ARG, "null buffer");
  for (int b = 0; b < 2; b++) {
    if (!c->d_buf[b]) HIP_TRY(c, hipMalloc(&c->d_buf[b], sizeof(double) * c->state_elems()));
    if (!c->t_buf[b]) HIP_TRY(c, hipMalloc(&c->t_buf[b], sizeof(double) * tracer_elems(c)));
  }
  int dirs[3], nd = 0;
  static const int fwd[3] = {0, 1, 2}, bwd[3] = {2, 1, 0};
  for (int s = 0; s < 3; s++) {
    const int d = c->dimSwitch ? fwd[s] : bwd[s];
    if (d == 1 && c->sim2d) continue;   // tendency is zero (S:572-575)
    dirs[nd++] = d;
  }
  const double* cur = d_state;
  const double* curt = d_tracers;
  int pair = 0;
  for (int s = 0; s < nd; s++) {
    double *dst, *dstt;
    if (s == nd - 1) {
      dst = d_state;
      dstt = d_tracers;
    } else {
      dst = c->d_buf[pair];
      dstt = c->t_buf[pair];
      pair ^= 1;
    }
    rc = launch_sweep_tr(c, cur, curt, dst, dstt, dt, dirs[s], false);
    if (rc != AWFL_OK) return rc;
    cur = dst;
    curt = dstt;
  }
  c->dimSwitch = !c->dimSwitch;
  return AWFL_OK;
}

}  // extern "C"
