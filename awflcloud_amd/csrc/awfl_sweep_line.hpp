// awfl_sweep_line.hpp -- lane-split split sweep for the high orders
// (WENO9/ADER5), included by awfl_dycore.hip.
//
// One work-group of 256 lanes per line along the sweep direction; the line
// and every intermediate stay in LDS.  The per-cell body of the production
// kernels (awfl_cell.hpp: cell_edges) keeps the whole transform of a cell in
// one lane's registers, 5 variables x NA levels x NGLL points of state plus
// the fluxes, which at (9, 5) is ~275 doubles: one wave per SIMD and spills.
// Here the work of a cell is spread over lanes:
//   1. stage the line with the boundary rule (periodic wrap or wall clamp,
//      S:631-662, 994-1025, 1351-1382) into LDS [5][n+2hs];
//   2. per chunk of 48 cells:
//      a. WENO, one lane per (cell, variable): the ORD-cell stencil -> NGLL
//         GLL values (WenoLimiter.h:93-172, S:1800-1833) into LDS;
//      b. the differential transform, one lane per (cell, GLL point): each
//         lane carries its point's levels of the 5 variables and the 4
//         fluxes, the derivative across the cell's points (S:1863-1882 and
//         analogues) reads the NGLL lanes' flux values from LDS, the
//         products (S:1885-1915) and the z source (S:2049-2055) are
//         pointwise; the edge lanes form the time-averaged edge states
//         (S:2142-2177), the z source is averaged at every point and summed
//         with the GLL weights (S:1489-1497);
//      the NGLL lanes of a cell sit in one wave (12 cells per wave, 4 lanes
//      idle), so their exchanges need no work-group barrier;
//   3. one lane per face: the Riemann flux (S:842-887, 1202-1245, 1583-1629)
//      from the edge states, with the edge rule at walls (S:816-837 and
//      analogues);
//   4. one lane per cell: the flux difference and the update (S:940-956,
//      1298-1310, 1680-1696; T:63-65), or the tendency alone.
// The arithmetic is the production kernels' (reconstruct, riemann and the
// transform's statements in their order), so results agree with them to
// rounding.
#pragma once
#include "awfl_cell.hpp"
#include "awfl_internal.hpp"

// Waves per SIMD the line kernels are compiled for.  Bounded to three, the
// (9, 5) kernels take 150 (x, y) and 158 (z) VGPRs and no scratch; bounded
// to four they spill 208-240 bytes per lane and run 1.6x slower.
#ifndef AWFL_LS_WAVES
#define AWFL_LS_WAVES 3
#endif

namespace awfl_line {

using namespace awfl;

constexpr int LS_THREADS = 256;
constexpr int LS_WAVES = LS_THREADS / 64;
constexpr int LS_CELLS_PER_WAVE = 12;                  // 12 cells x NGLL (<= 5) points <= 64 lanes
constexpr int LS_CHUNK = LS_WAVES * LS_CELLS_PER_WAVE; // cells per chunk

// LDS exchange among the lanes of one wave: the fence keeps the compiler from
// moving LDS accesses across it (LDS operations of a wave complete in order)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// LDS of one line of n cells: staged line (reused for the face fluxes), GLL
// values / flux exchange of a chunk, the source value of each chunk cell's
// last GLL point, edge states, z source averages (and, in place, the face
// fluxes apart from the line)
inline size_t line_lds_doubles(int n, int hs, int ngll, bool inplace = false) {
  return (size_t)5 * (n + 2 * hs) + (size_t)5 * ngll * LS_CHUNK + LS_CHUNK + (size_t)10 * n + n +
         (inplace ? (size_t)5 * (n + 1) : 0);
}

// INPLACE (in == out allowed): every global read of the line happens while
// staging it, before the first barrier; the face fluxes then get LDS of their
// own and the update takes each cell's old value from the staged line.
template <int ORD, int NGLL, int DIR, bool TEND, bool INPLACE = false>
__global__ __launch_bounds__(LS_THREADS) __attribute__((amdgpu_waves_per_eu(AWFL_LS_WAVES))) void k_sweep_line(const double* __restrict__ in, double* __restrict__ out,
                                                           Grid G, SweepK K) {
  constexpr int hs = (ORD - 1) / 2;
  constexpr int NA = NGLL;
  constexpr int NM = 1 + DIR;
  static_assert(NGLL * LS_CELLS_PER_WAVE <= 64, "a cell's GLL lanes must share a wave");
  extern __shared__ double sm[];
  const int n = DIR == 0 ? G.nx : DIR == 1 ? G.ny : G.nz;
  const int W = n + 2 * hs;
  // line -> first interior cell, stride along DIR, level k of x and y lines
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
  double* s_st = sm;                                   // [5][W]; then face fluxes [5][n+1]
  double* s_x = s_st + 5 * W;                          // [5 NGLL][CHUNK]: GLL values, then flux exchange
  double* s_sv = s_x + 5 * NGLL * LS_CHUNK;            // [CHUNK]
  double* s_e = s_sv + LS_CHUNK;                       // [10][n]: eL | eR
  double* s_src = s_e + 10 * n;                        // [n]: z source averages
  double* s_F = INPLACE ? s_src + n : sm;               // [5][n+1]
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;

  // 1. stage the line
  for (int c = threadIdx.x; c < W; c += LS_THREADS) {
    int a = c - hs, src;
    bool zn = false;
    if (per) {
      src = a % n;
      if (src < 0) src += n;
    } else if (a < 0) {
      src = 0; zn = true;
    } else if (a >= n) {
      src = n - 1; zn = true;
    } else {
      src = a;
    }
    sfor<5>([&](auto l) {
      const double v = in[l * G.sl + base + (int64_t)src * stride];
      s_st[l * W + c] = (zn && l == NM) ? 0.0 : v;
    });
  }
  __syncthreads();

  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  // 2b lane: cell of the chunk and GLL point
  const int cl_t = wave * LS_CELLS_PER_WAVE + lane / NGLL;
  const int ii = lane % NGLL;
  const bool lane_ok = lane < NGLL * LS_CELLS_PER_WAVE;
  // derivative-matrix column of this lane's point: d(ii) = sum_s D[s][ii] F(s)
  double Dc[NGLL];
  sfor<NGLL>([&](auto s) { Dc[s] = Deriv<NGLL>::m[s][ii]; });

  for (int c0 = 0; c0 < n; c0 += LS_CHUNK) {
    const int cnt = min(LS_CHUNK, n - c0);
    // 2a. WENO: lane -> (cell c0 + cl, variable v)
    if (threadIdx.x < 5 * LS_CHUNK) {
      const int v = threadIdx.x / LS_CHUNK, cl = threadIdx.x % LS_CHUNK;
      if (cl < cnt) {
        double u[ORD], gv[NGLL];
        sfor<ORD>([&](auto s) { u[s] = s_st[v * W + c0 + cl + s]; });
        const bool dw = (v == idR || v == idT) ? G.weno_s : G.weno_w;
        reconstruct<ORD, NGLL>(u, dw, gv);
        sfor<NGLL>([&](auto p) { s_x[(v * NGLL + p) * LS_CHUNK + cl] = gv[p]; });
      }
    }
    __syncthreads();
    // 2b. transform: lane -> (cell c0 + cl_t, point ii)
    const int cl = cl_t;
    const int c = c0 + cl;
    const bool act = lane_ok && cl < cnt;
    // LDS slot of this lane's cell: an idle lane reads the last cell's
    // values and stores nothing
    const int clr = min(cl, LS_CHUNK - 1);
    const int kc = DIR == 2 ? min(c, n - 1) : kl;       // level of the cell's background
    double q[5][NA], f[4][NA], g[NA];
    double src0 = 0.0;                                  // z: the source at level 0 (Q2 broadcast)
    // z gravity source at level kt: the level-0 value, then -rho_kt g (S:2055),
    // formed where it is used rather than kept
    auto src = [&](auto kt) -> double {
      if constexpr (kt == 0) return src0;
      else return -q[idR][kt] * GRAV;
    };
    double rr = 1.0, rrt = 1.0;
    {
      double q0[5];
      sfor<5>([&](auto l) { q0[l] = s_x[(l * NGLL + ii) * LS_CHUNK + clr]; });
      double brho, brt;
      PowRef pr;
      if constexpr (DIR == 2) {
        brho = G.hyDensGLL[kc * NGLL + ii];
        brt = G.hyDensThetaGLL[kc * NGLL + ii];
        pr = G.prG[kc * NGLL + ii];
      } else {
        brho = G.hyDensCells[hs + kl];
        brt = G.hyDensThetaCells[hs + kl];
        pr = G.prT[kl];
      }
      const double et = q0[idT] * pr.rb;
      q[idR][0] = q0[idR] + brho;
      q[idT][0] = q0[idT] + brt;
      sfor<3>([&](auto m) { q[m + 1][0] = q0[m + 1]; });
      if constexpr (DIR == 2) {                         // S:1413-1416
        if (!per && c == n - 1 && ii == NGLL - 1) q[idW][0] = 0.0;
        if (!per && c == 0 && ii == 0) q[idW][0] = 0.0;
      }
      // act == false: a harmless positive state, never stored
      if (!act) {
        q[idR][0] = 1.0;
        q[idT][0] = 300.0;
      }
      const double r = q[idR][0];
      rr = rcp(r);
      double vel[5];
      sfor<4>([&](auto m) { vel[m + 1] = q[m + 1][0] * rr; });
      sfor<4>([&](auto m) { f[m][0] = r * vel[NM] * vel[m + 1]; });
      g[0] = act ? pow_gamma_near(q[idT][0], et, pr.h, pr.l) : 1.0;
      rrt = rcp(q[idT][0]);
      if constexpr (DIR == 2) {
        // S:1439 assigns a scalar to the whole SArray: every entry of the
        // source holds the value of the last GLL point
        const double sv = -(q[idR][0] - brho) * GRAV;
        if (act && ii == NGLL - 1) s_sv[cl] = sv;
        wave_lds_sync();
        const double svl = s_sv[clr];
        src0 = svl;
      }
    }
    sfor<NA - 1>([&](auto kt) {
      constexpr double rkt1 = 1.0 / (kt + 1.0);
      // this point's five flux values at level kt into the exchange rows
      double Fl[5];
      Fl[0] = q[NM][kt];
      sfor<4>([&](auto m) {
        constexpr int l = m + 1;
        if constexpr (l == NM) {
          if constexpr (kt == 0) {
            if constexpr (DIR == 2) Fl[l] = f[m][kt] + C0 * g[kt] - G.hyPressureGLL[kc * NGLL + ii];
            else Fl[l] = f[m][kt] + C0 * g[kt];
          } else {
            Fl[l] = f[m][kt] + (0.5 * C0) * g[kt];
          }
        } else {
          Fl[l] = f[m][kt];
        }
      });
      wave_lds_sync();                                  // the cell's lanes are done with the previous rows
      if (act) sfor<5>([&](auto l) { s_x[(l * NGLL + ii) * LS_CHUNK + cl] = Fl[l]; });
      wave_lds_sync();
      double d[5];
      sfor<5>([&](auto l) {
        double t = -0.0;
        sfor<NGLL>([&](auto s) { t += Dc[s] * s_x[(l * NGLL + s) * LS_CHUNK + clr]; });
        d[l] = t;
      });
      const double ndt = K.ndt[kt];
      sfor<5>([&](auto l) {
        if constexpr (DIR == 2 && l == idW) q[l][kt + 1] = d[l] * ndt + src(kt) * rkt1;
        else q[l][kt + 1] = d[l] * ndt;
      });
      if constexpr (DIR == 2) {                         // S:2051-2055
        if (!per && c == n - 1 && ii == NGLL - 1) q[idW][kt + 1] = 0.0;
        if (!per && c == 0 && ii == 0) q[idW][kt + 1] = 0.0;
      }
      if constexpr (kt + 2 < NA) {                      // products at level kt+1
        sfor<4>([&](auto m) {
          constexpr int l = m + 1;
          double tot = 0.0;
          sfor<kt + 2>([&](auto ir) {
            if constexpr (ir == 0) tot += q[NM][ir] * q[l][kt + 1 - ir];
            else tot += q[NM][ir] * q[l][kt + 1 - ir] - q[idR][ir] * f[m][kt + 1 - ir];
          });
          f[m][kt + 1] = tot * rr;
        });
        double tg = 0.0;
        sfor<kt + 1>([&](auto ir) {
          constexpr double wgt = kt + 1.0 - ir;
          if constexpr (ir == 0)
            tg += wgt * (GAMMA * g[ir] * q[idT][kt + 1 - ir]);
          else
            tg += wgt * (GAMMA * g[ir] * q[idT][kt + 1 - ir] - q[idT][ir] * g[kt + 1 - ir]);
        });
        g[kt + 1] = (GAMMA * g[0] * q[idT][kt + 1] + tg * rkt1) * rrt;
      }
    });
    // time-averaged edge states from the edge points (S:2142-2177)
    if (act && (ii == 0 || ii == NGLL - 1)) {
      const int row = ii == 0 ? 0 : 5;
      sfor<5>([&](auto l) {
        double a = q[l][0];
        sfor<NA - 1>([&](auto km) { a += q[l][km + 1] * K.tav[km + 1]; });
        s_e[(row + l) * n + c] = a;
      });
    }
    if constexpr (DIR == 2) {                           // S:1462, 1489-1492
      double s0 = src(std::integral_constant<int, 0>{});
      sfor<NA - 1>([&](auto km) { s0 += src(std::integral_constant<int, km + 1>{}) * K.tav[km + 1]; });
      wave_lds_sync();
      if (act) s_x[ii * LS_CHUNK + cl] = s0 * Gll<NGLL>::wts[ii];
      wave_lds_sync();
      if (act && ii == 0) {
        double sa = 0.0;
        sfor<NGLL>([&](auto p) { sa += s_x[p * LS_CHUNK + cl]; });
        s_src[c] = sa;
      }
    }
    __syncthreads();                                    // the next chunk rewrites s_x
  }

  // 3. faces (s_F over the staged line, which no lane reads any more)
  const int nfaces = (per && DIR != 2) ? n : n + 1;
  for (int f = threadIdx.x; f < nfaces; f += LS_THREADS) {
    int a, b;                                           // qL: right edge of a, qR: left edge of b
    bool aL = false, bR = false;
    if (per) {
      a = f == 0 ? n - 1 : f - 1;
      b = f == n ? 0 : f;
    } else if (f == 0) {
      a = 0; b = 0; aL = true;
    } else if (f == n) {
      a = n - 1; b = n - 1; bR = true;
    } else {
      a = f - 1; b = f;
    }
    double qL[5], qR[5], F[5];
    sfor<5>([&](auto l) {
      qL[l] = aL ? s_e[l * n + a] : s_e[(5 + l) * n + a];
      qR[l] = bR ? s_e[(5 + l) * n + b] : s_e[l * n + b];
    });
    FaceRef fr;
    if constexpr (DIR == 0) {
      fr.rt = G.prT[kl];
      fr.q5 = fr.rt;
    } else if constexpr (DIR == 1) {
      fr.rt = G.prRT[kl];
      fr.q5 = G.prT[kl];
    } else {
      fr.rt = G.prF[f];
      fr.q5 = fr.rt;
    }
    riemann<DIR>(qL, qR, fr, F);
    if constexpr (DIR == 2) {                           // S:1628-1629
      if (f < n) F[idW] -= G.hyPressureGLL[f * NGLL + 0];
      else F[idW] -= G.hyPressureGLL[(n - 1) * NGLL + NGLL - 1];
    }
    sfor<5>([&](auto l) { s_F[l * (n + 1) + f] = F[l]; });
  }
  __syncthreads();

  // 4. flux difference and update
  for (int c = threadIdx.x; c < n; c += LS_THREADS) {
    const int cp = (per && DIR != 2 && c + 1 == n) ? 0 : c + 1;
    sfor<5>([&](auto l) {
      double td = -(s_F[l * (n + 1) + cp] - s_F[l * (n + 1) + c]) * K.rd;
      if constexpr (DIR == 2 && l == idW) td = s_src[c] + td;
      if (G.sim2d && l == idV) td = 0.0;
      if constexpr (TEND) {
        const int64_t ti = DIR == 0   ? (((int64_t)l * G.nz + kl) * G.ny + jl) * G.nx + c
                           : DIR == 1 ? (((int64_t)l * G.nz + kl) * G.ny + c) * G.nx + il
                                      : (((int64_t)l * G.nz + c) * G.ny + jl) * G.nx + il;
        out[ti] = td;
      } else {
        const int64_t o = l * G.sl + base + (int64_t)c * stride;
        if constexpr (INPLACE) out[o] = s_st[l * W + hs + c] + K.dt * td;
        else out[o] = in[o] + K.dt * td;
      }
    });
  }
}

}  // namespace awfl_line
