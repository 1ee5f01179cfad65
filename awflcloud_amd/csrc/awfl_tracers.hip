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
// Here the x sweep runs one line per work-group with the line in LDS, state
// and tracers together (k_sweep_tr), and the y and z sweeps run the walker
// (k_walk_tr: four x-adjacent lines per work-group walked in 64-cell steps);
// both run the production sweep's statements (cell_edges / riemann) for the
// state, so the state they return is awfl_sweep's up to the compiler's
// contraction choices in these kernels.
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

// ---------------------------------------------------------------------------
// y and z sweeps: the tracer walker.
//
// A y or z line is strided in memory (each cell on its own cache line), so
// the line kernel above, one line per work-group, issues one cache line per
// lane for every load and store.  Here a work-group of four waves takes
// WK_WX = 4 lines that are neighbours in x, wave w on line w, and walks them
// in steps of WK_SEG = 64 cells, lane c on cell X = c0 + c of the step.
// - Loads are issued with the lanes across x (the four lines' values of one
//   cell are one 32-byte run) into a ring of each line's cells in LDS: the
//   state, the mixing ratios, the tracer masses and (z) the cell sources.
//   The ring holds the step's stencils, cells c0 - HB .. c0 + 63 + hs.
// - A lane forms cell X's edges (the production cell_edges and the tracer
//   transform of the line kernel), then face X, the left face of its cell,
//   from the right edge of cell X - 1 (the next lower lane by a shuffle, or
//   the last cell of the previous step kept in LDS).
// - The positivity limiter needs a cell's factor from its two faces' fluxes
//   and a face's factor from its upwind cell, so a lane forms the factor of
//   cell X - 1, limits face X - 1 and updates cell X - 2: the update trails
//   the faces by two cells.  What the next step's first lanes need of the
//   last cells of this one is kept in a small carry area.  Per tracer the
//   edges, fluxes, factors and limited fluxes pass between lanes by
//   shuffles, in one loop.
// - Outputs go to an LDS buffer and are stored with the lanes across x at
//   the next step boundary, while that step's ring loads are in flight.
// - Cells 0 and 1 need face 0, which a periodic line forms from the last
//   cell's edge (and the limiter of either boundary may take the factor of
//   the cell at the other end, as the line kernel's indexing does), so the
//   first step keeps what they need in a head area and a tail pass after the
//   last step forms faces 0 and n and updates cells n-2, n-1, 0 and 1.
// Every per-line exchange stays inside the line's wave; the work-group meets
// at barriers only around the ring's loads.  The arithmetic of every cell,
// face, factor and update is the line kernel's, statement for statement.
// ---------------------------------------------------------------------------
constexpr int WK_SEG = 64;   // cells of a line per step, one per lane
constexpr int WK_WX = 4;     // lines per work-group, one per wave

template <int ORD>
struct WalkGeom {
  static constexpr int hs = (ORD - 1) / 2;
  static constexpr int HB = hs > 2 ? hs : 2;       // ring cells behind a step's first cell
  static constexpr int R = WK_SEG + HB + hs;       // ring length
};

// per-line LDS layout, in doubles
struct WalkLds {
  int ring;   // rows of R: state [5], mixing ratio [ntr], mass [ntr], (z) source [1]
  int carry;  // of the step's last cell X: right edge [5], F(X) [5], F(X-1) [5], limiter direction [1],
              // then per tracer its high edge, flux of face X, factor of cell X-1, limited flux of face X-1
  int outb;   // [5 + ntr][SEG] the step's outputs (cells c0-2 .. c0+61), stored with the lanes across x
  int head;   // cells 0 and 1 (first step), for the tail
  int tail;   // faces 0 and n, factors of cells 0 and n-1, limited faces n-1, 0, n, 1, n-2; then
              // the constants the tail reads (loaded by the first part of the walk)
  int size;
};
// head offsets
constexpr int WH_EL0 = 0, WH_F1 = 5, WH_F2 = 10, WH_DIR1 = 15, WH_TR = 16;   // then tlo0, TF1, fac1, TFl2 [ntr] each, then two cells
__host__ __device__ inline WalkLds walk_lds(int ntr, int R, bool z) {
  WalkLds L{};
  int o = 0;
  L.ring = o;  o += (5 + 2 * ntr + (z ? 1 : 0)) * R;
  L.carry = o; o += 16 + 4 * ntr;
  L.outb = o;  o += (5 + ntr) * WK_SEG;
  L.head = o;  o += WH_TR + 4 * ntr + 2 * (6 + ntr);
  L.tail = o;  o += 12 + 9 * ntr + 20;
  L.size = o;
  return L;
}

// orders one wave's LDS accesses across its lanes (a lane reads what another
// lane of the same wave wrote): no instruction, a fence for the compiler
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The walker's arguments, one struct at the start of the kernel argument
// segment.  Each part of the walk reads them again through a pointer the
// compiler cannot follow (walk_args): loaded once before the walk's loop,
// the sweep constants and grid fields stay live through every step's cells
// (registers and scalar registers spilled), where each part now loads what
// it uses, as the line kernel does.
struct WalkArgs {
  const double* in;
  const double* tr_in;
  double* out;
  double* tr_out;
  Grid G;
  SweepK K;
  int ntr;
  unsigned pos;
};
__device__ __forceinline__ WalkArgs walk_args() {
#if __HIP_DEVICE_COMPILE__
  uint64_t v = (uint64_t)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(v));
  const __attribute__((address_space(4))) WalkArgs* p = (const __attribute__((address_space(4))) WalkArgs*)v;
  WalkArgs r;
  r.in = p->in;
  r.tr_in = p->tr_in;
  r.out = p->out;
  r.tr_out = p->tr_out;
  r.G = p->G;
  r.K = p->K;
  r.ntr = p->ntr;
  r.pos = p->pos;
  return r;
#else
  return WalkArgs{};   // host pass: never called
#endif
}

// one part of the walk: 0 the first step's ring and the carries, 1 step s,
// 2 the outputs of step s-1 and the ring loads of step s, 3 the tail, 4 the
// outputs of step s
template <int ORD, int NGLL, int DIR, bool TEND>
__device__ __forceinline__ void walk(const WalkArgs& A, int mode, int s) {
  static_assert(DIR == 1 || DIR == 2, "the walker runs the y and z sweeps");
  const double* __restrict__ in = A.in;
  const double* __restrict__ tr_in = A.tr_in;
  double* __restrict__ out = A.out;
  double* __restrict__ tr_out = A.tr_out;
  const Grid& G = A.G;
  const SweepK& K = A.K;
  const int ntr = A.ntr;
  const unsigned pos = A.pos;
  using WGm = WalkGeom<ORD>;
  constexpr int hs = WGm::hs, HB = WGm::HB, R = WGm::R, SEG = WK_SEG, WX = WK_WX;
  constexpr int MT = AWFL_MAX_TRACERS;
  constexpr int NM = 1 + DIR;
  constexpr int NA = NGLL;
  extern __shared__ double sm[];
  const int n = DIR == 1 ? G.ny : G.nz;
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  const int ntx = (G.nx + WX - 1) / WX;
  const unsigned tile = tile_index();
  const int x0 = (int)(tile % (unsigned)ntx) * WX;
  const int other = (int)(tile / (unsigned)ntx);   // y: the level; z: the row
  const int kl = DIR == 1 ? other : 0;
  const int jl = DIR == 2 ? other : 0;
  const int64_t stride = DIR == 1 ? G.sy : G.sz;
  const int64_t base0 = DIR == 1 ? (int64_t)(hs + kl) * G.sz + (int64_t)hs * G.sy + hs + x0
                                 : (int64_t)hs * G.sz + (int64_t)(hs + jl) * G.sy + hs + x0;
  const int64_t nc = (int64_t)G.nx * G.ny * G.nz;
  const WalkLds Lo = walk_lds(ntr, R, DIR == 2);
  const int tid = threadIdx.x;
  const int lw = __builtin_amdgcn_readfirstlane(tid >> 6);
  int c = tid & 63;
  asm volatile("" : "+v"(c));   // lane addresses formed in each part, not held across the walk
  const bool lvalid = x0 + lw < G.nx;
  double* const ln = sm + lw * Lo.size;
  double* const rg = ln + Lo.ring;                 // state rows, row l at rg + l * R
  double* const mr = rg + 5 * R;                   // mixing ratios
  double* const ms = mr + ntr * R;                 // masses
  double* const srcr = ms + ntr * R;               // z sources
  double* const ob = ln + Lo.outb;
  double* const cy = ln + Lo.carry;
  double* const ct_hi = cy + 16;                   // per tracer: high edge of the last cell
  double* const ct_tf = ct_hi + ntr;               // flux of its face (unlimited)
  double* const ct_fac = ct_tf + ntr;              // factor of the cell before it
  double* const ct_tfl = ct_fac + ntr;             // limited flux of the face before it
  double* const hd = ln + Lo.head;
  double* const tl = ln + Lo.tail;
  double* const h_tlo0 = hd + WH_TR;
  double* const h_tf1 = h_tlo0 + ntr;
  double* const h_fac1 = h_tf1 + ntr;
  double* const h_tfl2 = h_fac1 + ntr;
  double* const h_cell = h_tfl2 + ntr;             // [2][6 + ntr]: state, masses, source
  // tail: F of faces 0 and n [2][5], their limiter directions [2], tracer fluxes [2][ntr],
  // factors of cells 0 and n-1 [2][ntr], limited faces n-1, 0, n, 1, n-2 [5][ntr]
  double* const t_F = tl;
  double* const t_dir = tl + 10;
  double* const t_tf = tl + 12;
  double* const t_fac = t_tf + 2 * ntr;
  double* const t_lim = t_fac + 2 * ntr;
  // tail constants: power references of faces 0 and n [2][8], the background
  // pressure each subtracts (z) [2], background densities of cells 0 and n-1 [2]
  double* const t_con = t_lim + 5 * ntr;
  auto slot = [](int a) { return (a + HB) % R; };   // a >= -HB
  auto bg_dens = [&](int cell) {                    // background density of a cell clamped into the line
    const int cc = cell < 0 ? 0 : (cell >= n ? n - 1 : cell);
    return G.hyDensCells[hs + (DIR == 2 ? cc : kl)];
  };
  auto upwind_dir = [](double u) { return u > 0 ? -1 : (u < 0 ? 1 : 0); };
  auto face_ref = [&](int f) {
    FaceRef fr;
    if constexpr (DIR == 1) {
      fr.rt = G.prRT[kl];
      fr.q5 = G.prT[kl];
    } else {
      fr.rt = fr.q5 = G.prF[f];
    }
    return fr;
  };
  auto face_bg = [&](int f) {                      // background pressure a z face subtracts (S:1628-1629)
    return f < G.nz ? G.hyPressureGLL[f * NGLL + 0] : G.hyPressureGLL[(G.nz - 1) * NGLL + NGLL - 1];
  };

  // ring loads: cells a0 .. a0 + cnt - 1 of every line of the tile, lanes across x
  // mid: between the loads and the LDS stores (the previous step's outputs
  // are stored there, while the loads are in flight)
  auto stage = [&](int a0, int cnt, auto&& mid) {
    const int lt = tid % WX;
    if (x0 + lt >= G.nx) {
      mid();
      return;
    }
    double* const rt = sm + lt * Lo.size + Lo.ring;
    for (int k = tid / WX; k < cnt; k += SEG) {
      const int a = a0 + k;
      bool zn;
      const int src = wrap_axis(a, n, per, zn);
      const int64_t o = base0 + lt + (int64_t)src * stride;
      double v[5], q[MT];
      sfor<5>([&](auto l) { v[l] = in[l * G.sl + o]; });
      sfor<MT>([&](auto t) {
        if (t < ntr) q[t] = tr_in[t * G.sl + o];
      });
      const double bg = G.hyDensCells[hs + (DIR == 2 ? src : kl)];
      if (k == tid / WX) mid();
      const int sl = slot(a);
      sfor<5>([&](auto l) { rt[l * R + sl] = (zn && l == NM) ? 0.0 : v[l]; });
      const double rr = rcp(v[idR] + bg);
      sfor<MT>([&](auto t) {
        if (t < ntr) {
          rt[(5 + t) * R + sl] = q[t] * rr;
          rt[(5 + ntr + t) * R + sl] = q[t];
        }
      });
    }
  };
  // state and tracer output of cell `cell` (the tail's cells)
  auto put = [&](int cell, const double (&Fl)[5], const double (&Fr)[5], const double* tfl, const double* tfr, int ts,
                 const double* st0, const double* ms0, int mstride, double sa) {
    int k = kl, j = jl;
    if constexpr (DIR == 1) j = cell;
    else k = cell;
    const int64_t o = base0 + lw + (int64_t)cell * stride;
    const int64_t ot = ((int64_t)k * G.ny + j) * G.nx + x0 + lw;
    sfor<5>([&](auto l) {
      double td = -(Fr[l] - Fl[l]) * K.rd;
      if (DIR == 2 && l == idW) td = sa + td;
      if (G.sim2d && l == idV) td = 0.0;
      if constexpr (TEND) out[l * nc + ot] = td;
      else out[l * G.sl + o] = st0[l * mstride] + K.dt * td;
    });
    for (int t = 0; t < ntr; t++) {
      const double td = -(tfr[t * ts] - tfl[t * ts]) * K.rd;
      if constexpr (TEND) tr_out[t * nc + ot] = td;
      else tr_out[t * G.sl + o] = ms0[t * mstride] + K.dt * td;
    }
  };

  const int S = (n + SEG - 1) / SEG;
  if (mode == 0) {
    if (lvalid && c == 0) {   // zeros for what step 0's lane 0 reads of a previous step; the tail's constants
      for (int e = 0; e < 16 + 4 * ntr; e++) cy[e] = 0.0;
      for (int f = 0; f < 2; f++) {
        const FaceRef fr = face_ref(f == 0 ? 0 : n);
        const PowRef* pr[2] = {&fr.rt, &fr.q5};
        for (int r = 0; r < 2; r++) {
          t_con[f * 8 + r * 4 + 0] = pr[r]->h;
          t_con[f * 8 + r * 4 + 1] = pr[r]->l;
          t_con[f * 8 + r * 4 + 2] = pr[r]->rb;
          t_con[f * 8 + r * 4 + 3] = pr[r]->b;
        }
        t_con[16 + f] = DIR == 2 ? face_bg(f == 0 ? 0 : n) : 0.0;
      }
      t_con[18] = bg_dens(0);
      t_con[19] = bg_dens(n - 1);
    }
    stage(-HB, R, [] {});
    return;
  }
  // the outputs of step sf, lanes across x: cell X-2 for X = sf*SEG + k
  auto flush = [&](int sf) {
    const int lt = tid % WX, k = tid / WX;
    const int X = sf * SEG + k;
    if (x0 + lt >= G.nx || X < 4 || X >= n) return;
    const double* const obt = sm + lt * Lo.size + Lo.outb;
    const int cell = X - 2;
    int kk = kl, j = jl;
    if constexpr (DIR == 1) j = cell;
    else kk = cell;
    const int64_t o = base0 + lt + (int64_t)cell * stride;
    const int64_t ot = ((int64_t)kk * G.ny + j) * G.nx + x0 + lt;
    sfor<5>([&](auto l) {
      if constexpr (TEND) out[l * nc + ot] = obt[l * SEG + k];
      else out[l * G.sl + o] = obt[l * SEG + k];
    });
    for (int t = 0; t < ntr; t++) {
      if constexpr (TEND) tr_out[t * nc + ot] = obt[(5 + t) * SEG + k];
      else tr_out[t * G.sl + o] = obt[(5 + t) * SEG + k];
    }
  };
  if (mode == 2) {
    stage(s * SEG + hs, SEG, [&] { flush(s - 1); });
    return;
  }
  if (mode == 4) {
    flush(s);
    return;
  }
  if (mode == 1) {
    const int c0 = s * SEG;
    const int X = c0 + c;
    const int cw = s == S - 1 ? n - 1 - c0 : SEG - 1;   // the lane whose cell the next step (or the tail) continues from
    if (!lvalid) return;
    // (a) cell X: state edges, the source, and the density and normal
    // momentum levels the tracer transform reads
    double eL[5], eR[5], sa;
    double rl[NA][NGLL], nl[NA][NGLL];
    {
      const int Xc = X < n ? X : n - 1;
      CellBg<NGLL> bg{};
      if constexpr (DIR == 2) {
        sfor<NGLL>([&](auto ii) {
          bg.rG[ii] = G.hyDensGLL[Xc * NGLL + ii];
          bg.rtG[ii] = G.hyDensThetaGLL[Xc * NGLL + ii];
          bg.pG[ii] = G.hyPressureGLL[Xc * NGLL + ii];
          bg.ptG[ii] = G.prG[Xc * NGLL + ii];
        });
        bg.wall_bot = !per && X == 0;
        bg.wall_top = !per && X == n - 1;
      } else {
        bg.r = G.hyDensCells[hs + kl];
        bg.rt = G.hyDensThetaCells[hs + kl];
        bg.pt = G.prT[kl];
        bg.wall_bot = bg.wall_top = false;
      }
      double st[5][ORD];
      const int s0 = slot(X - hs);
      sfor<ORD>([&](auto k) {
        const int sk = s0 + k < R ? s0 + k : s0 + k - R;
        sfor<5>([&](auto l) { st[l][k] = rg[l * R + sk]; });
      });
      cell_edges<ORD, NGLL, DIR>(st, bg, G.weno_s, G.weno_w, K, eL, eR, sa, [&](const auto& q) {
        sfor<NA>([&](auto kt) {
          sfor<NGLL>([&](auto ii) {
            rl[kt][ii] = q[idR][kt][ii];
            nl[kt][ii] = q[NM][kt][ii];
          });
        });
      });
    }
    if constexpr (DIR == 2) srcr[slot(X)] = sa;
    // lane X-2 reads this source in the update below; the fence also ends the
    // cell's schedule here (measured: y 0.803 -> 0.786 ms, its spill halved)
    wave_sync();
    if (s == 0 && c == 0) sfor<5>([&](auto l) { hd[WH_EL0 + l] = eL[l]; });
    // (b) face X from the right edge of cell X-1 (the next lower lane, or
    // the carried cell c0-1; at the first step lane 0 takes its own left
    // edge: face 0 is formed again by the tail)
    double F[5], un, up;
    int dX;
    {
      double L[5], Rr[5];
      sfor<5>([&](auto l) {
        const double sh = __shfl_up(eR[l], 1);
        L[l] = c == 0 ? (s == 0 ? eL[l] : cy[l]) : sh;
        Rr[l] = eL[l];
      });
      if (c == cw) sfor<5>([&](auto l) { cy[l] = eR[l]; });
      riemann<DIR>(L, Rr, face_ref(X < n ? X : n), F, un);
      if constexpr (DIR == 2) F[idW] -= face_bg(X < n ? X : n);
      up = un > 0 ? F[0] * rcp(L[0]) : F[0] * rcp(Rr[0]);
      dX = upwind_dir(0.5 * (L[NM] + Rr[NM]));   // the normal momentum average (S:907)
    }
    if (s == 0 && c == 1) hd[WH_DIR1] = (double)dX;
    const int dm1i = __shfl_up(dX, 1);
    const int dm1 = c == 0 ? (int)cy[15] : dm1i;   // direction of face X-1
    if (c == cw) cy[15] = (double)dX;
    // (c) the state of cell X-2 from faces X-2 and X-1 (the update trails
    // the faces by two cells: the tracers' limiter needs the factor of the
    // cell on either side of a face)
    const bool upd = X >= 4 && X < n;
    const int sc = slot(X - 2);
    {
      double F1[5], F2[5];
      sfor<5>([&](auto l) {
        const double a = __shfl_up(F[l], 1);
        F1[l] = c == 0 ? cy[5 + l] : a;
      });
      if (c == cw) sfor<5>([&](auto l) { cy[5 + l] = F[l]; });
      sfor<5>([&](auto l) {
        const double a = __shfl_up(F1[l], 1);
        F2[l] = c == 0 ? cy[10 + l] : a;
      });
      if (c == cw) sfor<5>([&](auto l) { cy[10 + l] = F1[l]; });
      if (upd) {
        const double sa2 = DIR == 2 ? srcr[sc] : 0.0;
        sfor<5>([&](auto l) {
          double td = -(F1[l] - F2[l]) * K.rd;
          if (DIR == 2 && l == idW) td = sa2 + td;
          if (G.sim2d && l == idV) td = 0.0;
          ob[l * SEG + c] = TEND ? td : rg[l * R + sc] + K.dt * td;
        });
      } else if (s == 0 && (c == 2 || c == 3)) {
        // cells 0 and 1 as they came in, and faces 1 and 2, for the tail
        double* const hc = h_cell + (c - 2) * (6 + ntr);
        sfor<5>([&](auto l) { hc[l] = rg[l * R + sc]; });
        for (int t = 0; t < ntr; t++) hc[5 + t] = ms[t * R + sc];
        hc[5 + ntr] = DIR == 2 ? srcr[sc] : 0.0;
        if (c == 3) sfor<5>([&](auto l) {
            hd[WH_F1 + l] = F2[l];
            hd[WH_F2 + l] = F1[l];
          });
      }
    }
    // (d) per tracer: cell X's edges (the line kernel's transform), face X's
    // flux upwinded with the state's normal velocity, the limiter factor of
    // cell X-1 (T1), face X-1 limited by its upwind cell, and cell X-2's
    // update.  Neighbouring lanes' values come by shuffles; lane 0 takes the
    // previous step's last cell from the carry area, which lane cw then
    // overwrites for the next step.
    double rr0[NGLL];
    sfor<NGLL>([&](auto ii) { rr0[ii] = rcp(rl[0][ii]); });
    const int sm1 = slot(X - 1);
    const double dns = pos ? rg[idR * R + sm1] + bg_dens(X - 1) : 0.0;
    for (int t = 0; t < ntr; t++) {
      const bool tp = (pos >> t) & 1u;
      double u[ORD], g[NGLL];
      const int s1 = slot(X - hs);
      sfor<ORD>([&](auto kk) {
        const int sk = s1 + kk < R ? s1 + kk : s1 + kk - R;
        u[kk] = mr[t * R + sk];
      });
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
          sfor<NGLL>([&](auto s2) { df += Deriv<NGLL>::m[s2][ii] * rnt[kt][s2]; });
          rt[kt + 1][ii] = df * K.ndt[kt];
        });
        if constexpr (kt + 2 < NA) {
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
      if (s == 0 && c == 0) h_tlo0[t] = a0;
      // face X: rho*q of the upwind side, the high edge of cell X-1 or the low edge of cell X
      const double a1s = __shfl_up(a1, 1);
      const double a1m = c == 0 ? ct_hi[t] : a1s;
      const double tfX = (un > 0 ? a1m : a0) * up;
      if (s == 0 && c == 1) h_tf1[t] = tfX;
      const double tfs = __shfl_up(tfX, 1);
      const double tfm = c == 0 ? ct_tf[t] : tfs;   // face X-1, unlimited
      double lim = tfm, fv = 0.0;
      if (tp) {
        const double f1 = fmin(tfm, 0.0);
        const double f2 = fmax(tfX, 0.0);
        const double fluxOut = K.dt * (f2 - f1) * K.rd;
        const double q = mr[t * R + sm1];
        fv = fmin(1.0, q * dns * rcp(fluxOut + TRACER_EPS));   // factor of cell X-1
        if (s == 0 && c == 2) h_fac1[t] = fv;
        const double fs = __shfl_up(fv, 1);
        const double fm = c == 0 ? ct_fac[t] : fs;   // factor of cell X-2
        if (dm1 != 0) lim *= dm1 < 0 ? fm : fv;
      }
      if (s == 0 && c == 3) h_tfl2[t] = lim;
      const double ls = __shfl_up(lim, 1);
      const double lm = c == 0 ? ct_tfl[t] : ls;     // face X-2, limited
      if (upd) {
        const double td = -(lim - lm) * K.rd;
        ob[(5 + t) * SEG + c] = TEND ? td : ms[t * R + sc] + K.dt * td;
      }
      if (c == cw) {
        ct_hi[t] = a1;
        ct_tf[t] = tfX;
        ct_fac[t] = fv;
        ct_tfl[t] = lim;
      }
    }
    return;
  }
  if (!lvalid) return;
  wave_sync();

  // tail: faces 0 (lane 0) and n (lane 1).  A periodic y line's face n is its
  // face 0 (both lanes form the same flux); a periodic z line keeps both,
  // which differ by the background pressure each subtracts.
  if (c < 2) {
    const bool f0 = c == 0;
    const double* const eRn = cy;                  // right edge of cell n-1
    const double* const eL0 = hd + WH_EL0;         // left edge of cell 0
    double L[5], Rr[5], F[5], un;
    sfor<5>([&](auto l) {
      L[l] = (f0 && !per) ? eL0[l] : eRn[l];
      Rr[l] = (!f0 && !per) ? eRn[l] : eL0[l];
    });
    FaceRef fr;
    const double* const tc = t_con + c * 8;
    fr.rt = PowRef{tc[0], tc[1], tc[2], tc[3]};
    fr.q5 = PowRef{tc[4], tc[5], tc[6], tc[7]};
    riemann<DIR>(L, Rr, fr, F, un);
    if constexpr (DIR == 2) F[idW] -= t_con[16 + c];
    sfor<5>([&](auto l) { t_F[c * 5 + l] = F[l]; });
    t_dir[c] = (double)upwind_dir(0.5 * (L[NM] + Rr[NM]));
    const double up = un > 0 ? F[0] * rcp(L[0]) : F[0] * rcp(Rr[0]);
    for (int t = 0; t < ntr; t++) {
      const double thn = ct_hi[t];                 // high edge of cell n-1
      const double tl0 = h_tlo0[t];                // low edge of cell 0
      const double tL = (f0 && !per) ? tl0 : thn;
      const double tR = (!f0 && !per) ? thn : tl0;
      t_tf[c * ntr + t] = (un > 0 ? tL : tR) * up;
    }
  }
  wave_sync();
  // factors of cell 0 (lane 0: faces 0 and 1) and cell n-1 (lane 1: faces n-1 and n)
  if (pos && c < 2) {
    const bool f0 = c == 0;
    const int sn = slot(n - 1);
    const double* const hc0 = h_cell;
    const double dns = f0 ? hc0[idR] + t_con[18] : rg[idR * R + sn] + t_con[19];
    const int o_l = f0 ? Lo.tail + 12 : Lo.carry + 16 + ntr;
    const int o_r = f0 ? Lo.head + WH_TR + ntr : Lo.tail + 12 + ntr;
    for (int t = 0; t < ntr; t++) {
      if ((pos >> t) & 1u) {
        const double fl = ln[o_l + t];
        const double fr = ln[o_r + t];
        const double f1 = fmin(fl, 0.0);
        const double f2 = fmax(fr, 0.0);
        const double fluxOut = K.dt * (f2 - f1) * K.rd;
        const double q = f0 ? ln[Lo.head + WH_TR + 4 * ntr + 5 + t] * rcp(dns) : mr[t * R + sn];
        t_fac[c * ntr + t] = fmin(1.0, q * dns * rcp(fluxOut + TRACER_EPS));
      }
    }
  }
  wave_sync();
  // limited faces n-1, 0, n, 1 (lanes 0..3): per lane, the LDS offsets of
  // the face's unlimited flux, its direction and the factors of the cells
  // on its two sides, chosen as numbers (a chain of branches over lanes
  // loading different addresses was compiled with one lane's address
  // taken for another's)
  if (c < 4) {
    const int LT = Lo.tail;
    const int o_tf = c == 0 ? Lo.carry + 16 + ntr : c == 1 ? LT + 12 : c == 2 ? LT + 12 + ntr : Lo.head + WH_TR + ntr;
    const int o_d = c == 0 ? Lo.carry + 15 : c == 1 ? LT + 10 : c == 2 ? LT + 11 : Lo.head + WH_DIR1;
    const int o_fl = c == 0 ? Lo.carry + 16 + 2 * ntr : c == 3 ? LT + 12 + 2 * ntr : LT + 12 + 3 * ntr;
    const int o_fr = c == 0 ? LT + 12 + 3 * ntr : c == 3 ? Lo.head + WH_TR + 2 * ntr : LT + 12 + 2 * ntr;
    const double d = ln[o_d];
    for (int t = 0; t < ntr; t++) {
      double v = ln[o_tf + t];
      if (((pos >> t) & 1u) && d != 0.0) v *= d < 0 ? ln[o_fl + t] : ln[o_fr + t];
      t_lim[c * ntr + t] = v;
      if (c == 0) t_lim[4 * ntr + t] = ct_tfl[t];    // face n-2, limited by the last step
    }
  }
  wave_sync();
  // cells n-2, n-1, 0, 1 (lanes 0..3), their inputs chosen by LDS offset
  if (c < 4) {
    const int sc = slot(n - 2 + (c & 1));
    const int o_hc = Lo.head + WH_TR + 4 * ntr + (c - 2) * (6 + ntr);   // c >= 2: the cell kept by the first step
    // state fluxes of the left and right faces: F(n-2) | F(n-1) | F(n) | F(0) | F(1) | F(2)
    const int o_Fl = c == 0 ? Lo.carry + 10 : c == 1 ? Lo.carry + 5 : c == 2 ? Lo.tail : Lo.head + WH_F1;
    const int o_Fr = c == 0 ? Lo.carry + 5 : c == 1 ? Lo.tail + 5 : c == 2 ? Lo.head + WH_F1 : Lo.head + WH_F2;
    double Fl[5], Fr[5];
    sfor<5>([&](auto l) {
      Fl[l] = ln[o_Fl + l];
      Fr[l] = ln[o_Fr + l];
    });
    // the cell as it came in: the ring (cells n-2, n-1) or the head (cells 0, 1)
    const int o_st = c < 2 ? Lo.ring + sc : o_hc;
    const int mst = c < 2 ? R : 1;
    const int o_ms = c < 2 ? Lo.ring + (5 + ntr) * R + sc : o_hc + 5;
    const double sa = DIR == 2 ? ln[c < 2 ? Lo.ring + (5 + 2 * ntr) * R + sc : o_hc + 5 + ntr] : 0.0;
    // limited fluxes of the cell's faces: n-2|n-1, n-1|n, 0|1, 1|2
    const int o_tl = Lo.tail + 12 + 4 * ntr;
    const int o_l = c == 0 ? o_tl + 4 * ntr : c == 1 ? o_tl : c == 2 ? o_tl + ntr : o_tl + 3 * ntr;
    const int o_r = c == 0 ? o_tl : c == 1 ? o_tl + 2 * ntr : c == 2 ? o_tl + 3 * ntr : Lo.head + WH_TR + 3 * ntr;
    put(c < 2 ? n - 2 + c : c - 2, Fl, Fr, ln + o_l, ln + o_r, 1, ln + o_st, ln + o_ms, mst, sa);
  }
}

template <int ORD, int NGLL, int DIR, bool TEND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_walk_tr(WalkArgs) {
  const int n = DIR == 1 ? walk_args().G.ny : walk_args().G.nz;
  const int S = (n + WK_SEG - 1) / WK_SEG;
  walk<ORD, NGLL, DIR, TEND>(walk_args(), 0, 0);
  __syncthreads();
  for (int s = 0; s < S; s++) {
    walk<ORD, NGLL, DIR, TEND>(walk_args(), 1, s);
    if (s + 1 < S) {
      __syncthreads();
      walk<ORD, NGLL, DIR, TEND>(walk_args(), 2, s + 1);
      __syncthreads();
    }
  }
  walk<ORD, NGLL, DIR, TEND>(walk_args(), 3, S - 1);   // the tail: this wave's line only
  __syncthreads();
  walk<ORD, NGLL, DIR, TEND>(walk_args(), 4, S - 1);
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
using KernelWk = void (*)(WalkArgs);

struct VariantTr {
  int ord, ngll;
  KernelTr upd[3], tend[3];
  KernelWk wupd[2], wtend[2];   // y and z walkers (null: the line kernels run y and z)
  void (*init)(double*, Grid, int, double, double);
};

#define AWFL_VARIANT_TR(O, N, WU, WT)                                                          \
  VariantTr {                                                                                  \
    O, N, {k_sweep_tr<O, N, 0, false>, k_sweep_tr<O, N, 1, false>, k_sweep_tr_z<O, N, false>}, \
        {k_sweep_tr<O, N, 0, true>, k_sweep_tr<O, N, 1, true>, k_sweep_tr_z<O, N, true>}, WU, WT, \
        k_init_tr<O>,                                                                          \
  }
#define AWFL_WALK_TR(O, N, T) {k_walk_tr<O, N, 1, T>, k_walk_tr<O, N, 2, T>}

// (9, 5) keeps the line kernels in y and z: its cell is too large for a
// work-group of four lines
const VariantTr kVariantsTr[] = {
    AWFL_VARIANT_TR(3, 2, AWFL_WALK_TR(3, 2, false), AWFL_WALK_TR(3, 2, true)),
    AWFL_VARIANT_TR(5, 3, AWFL_WALK_TR(5, 3, false), AWFL_WALK_TR(5, 3, true)),
    AWFL_VARIANT_TR(9, 5, {}, {})};

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
  const int64_t nc = (int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz;
  if (dir > 0 && n >= 8 && v->wupd[dir - 1]) {
    // the walker: tiles of WK_WX x-adjacent lines
    const int hs = c->hs, HB = hs > 2 ? hs : 2;
    const WalkLds L = walk_lds(ntr, WK_SEG + HB + hs, dir == 2);
    const size_t lds = sizeof(double) * (size_t)L.size * WK_WX;
    KernelWk k = tend ? v->wtend[dir - 1] : v->wupd[dir - 1];
    if (lds > 64 * 1024)
      HIP_TRY(c, hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const unsigned tiles = (unsigned)((c->cfg.nx + WK_WX - 1) / WK_WX) * (unsigned)(dir == 1 ? c->cfg.nz : c->cfg.ny);
    const WalkArgs a{in, tr_in, out, tr_out, c->grid(), sweep_constants(c, dt, dir), ntr, c->tpos};
    k<<<tiles, 64 * WK_WX, lds, c->stream>>>(a);
    LAUNCH_CHECK(c);
    return AWFL_OK;
  }
  const size_t lds = sizeof(double) * (size_t)(5 * W + ntr * W + 10 * n + 2 * ntr * n + (dir == 2 ? 2 : 1) * n);
  if (lds > 160 * 1024) return fail(c, AWFL_E_UNSUPPORTED, "tracers: line too long for the tracer sweep's LDS");
  KernelTr k = tend ? v->tend[dir] : v->upd[dir];
  if (lds > 64 * 1024)
    HIP_TRY(c, hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
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
