// awfl_cell_ref.hpp -- the per-cell and per-face numerics of one split sweep
// in the reference's own operation order, for the parity builds of the FP64
// sweeps (AWFL_PARITY with AWFL_PAR_ORDER, compiled with -ffp-contract=off).
//
// Every division, square root and power stands where the reference has one,
// every sum runs in the reference's order, nothing is folded or hoisted:
//   WENO ..................... WenoLimiter.h:12-25, 93-172
//   reconstruction ........... S:1800-1833
//   level-0 products ......... S:704-715, 1067-1078, 1428-1440
//   differential transform ... S:1837-1916 (x), 1920-2000 (y), 2004-2091 (z)
//   time averages ............ S:2130-2177
//   Riemann fluxes ........... S:842-887 (x), 1202-1245 (y), 1583-1629 (z)
// With IEEE division and square root (the hardware sequences the compiler
// emits for `/` and sqrt are correctly rounded) and the same power function
// on both sides, the result is the CPU oracle's to the last bit, except where
// the two compilers order a sum differently.  The production kernels
// (awfl_cell.hpp) differ from this by the rewrites that make them fast.
#pragma once
#include <cmath>

#include "awfl_cell.hpp"

namespace awfl {
namespace ref {

// (rho theta)^GAMMA: the general pow_gamma (the same double-double method on
// the host and the device) or the device library's pow
__host__ __device__ __forceinline__ double eos_pow(double x) {
#if defined(AWFL_PARITY) && (AWFL_PARITY & AWFL_PAR_LIBPOW)
  return ::pow(x, GAMMA);
#else
  return pow_gamma(x);
#endif
}

// generated truncated-TV quadratic form, terms summed left to right with the
// square terms multiplied as c * (a * a) (TransformMatrices.h:188-193)
template <int N>
__host__ __device__ __forceinline__ double tv_terms(const double (&a)[N]) {
  double r = 0.0;
  sfor<Tv<N>::n>([&](auto t) {
    constexpr int j = Tv<N>::j[t], k = Tv<N>::k[t];
    constexpr double c = Tv<N>::c[t];
    const double term = (j == k) ? c * (a[j] * a[k]) : c * a[j] * a[k];
    r = (t == 0) ? term : r + term;
  });
  return r;
}

// weno::compute_weno_coefs with convexify and map_weights (WenoLimiter.h:12-25, 93-172)
template <int ORD>
__host__ __device__ __forceinline__ void weno_coefs(const double (&u)[ORD], double (&aw)[ORD]) {
  using W = Weno<ORD>;
  constexpr int hs = W::hs;
  constexpr int NP = hs + 2;
  const double eps = 1.0e-20;
  double a[NP][ORD];
  for (int j = 0; j < NP; j++)
    for (int i = 0; i < ORD; i++) a[j][i] = 0.0;
  sfor<hs + 1>([&](auto i) {
    sfor<hs + 1>([&](auto ii) {
      sfor<hs + 1>([&](auto s) { a[i][ii] += W::recon[i][s][ii] * u[i + s]; });
    });
  });
  sfor<ORD>([&](auto ii) { sfor<ORD>([&](auto s) { a[hs + 1][ii] += W::recon[hs + 1][s][ii] * u[s]; }); });
  sfor<hs + 1>([&](auto i) { sfor<hs + 1>([&](auto ii) { a[hs + 1][ii] -= W::idl[i] * a[i][ii]; }); });
  sfor<ORD>([&](auto ii) { a[hs + 1][ii] /= W::idl[hs + 1]; });
  double tv[NP];
  sfor<hs + 1>([&](auto i) {
    double lo[hs + 1];
    sfor<hs + 1>([&](auto ii) { lo[ii] = a[i][ii]; });
    tv[i] = tv_terms<hs + 1>(lo);
  });
  {
    double hi[ORD];
    sfor<ORD>([&](auto ii) { hi[ii] = a[hs + 1][ii]; });
    tv[hs + 1] = tv_terms<ORD>(hi);
  }
  double lo_avg = 0.0;
  sfor<hs + 1>([&](auto i) { lo_avg += tv[i]; });
  lo_avg /= hs + 1;
  tv[hs + 1] = lo_avg + (tv[hs + 1] - lo_avg) * W::sigma;
  double wts[NP];
  sfor<NP>([&](auto i) { wts[i] = W::idl[i] / (tv[i] * tv[i] + eps); });
  auto convexify = [&]() {
    double sum = 0.0;
    sfor<NP>([&](auto i) { sum += wts[i]; });
    sfor<NP>([&](auto i) { wts[i] /= (sum + eps); });
  };
  convexify();
  sfor<NP>([&](auto i) {
    constexpr double d = W::idl[i];
    wts[i] = wts[i] * (d + d * d - 3.0 * d * wts[i] + wts[i] * wts[i]) / (d * d + wts[i] * (1.0 - 2.0 * d));
  });
  convexify();
  for (int i = 0; i < ORD; i++) aw[i] = 0.0;
  sfor<NP>([&](auto i) { sfor<ORD>([&](auto ii) { aw[ii] += wts[i] * a[i][ii]; }); });
}

// reconstruct_gll_values (S:1800-1833)
template <int ORD, int NGLL>
__host__ __device__ __forceinline__ void reconstruct(const double (&u)[ORD], bool doweno, double (&g)[NGLL]) {
  using P = Proj<ORD, NGLL>;
  if (doweno) {
    double wc[ORD];
    weno_coefs<ORD>(u, wc);
    sfor<NGLL>([&](auto ii) {
      double t = 0.0;
      sfor<ORD>([&](auto s) { t += P::c2g[s][ii] * wc[s]; });
      g[ii] = t;
    });
  } else {
    sfor<NGLL>([&](auto ii) {
      double t = 0.0;
      sfor<ORD>([&](auto s) { t += P::s2g[s][ii] * u[s]; });
      g[ii] = t;
    });
  }
}

// One cell of a DIR sweep: reconstruction, level-0 products, differential
// transform, time averages (the production cell_edges, in reference order).
// K.dsw is the grid spacing along the sweep.
template <int ORD, int NGLL, int DIR>
__host__ __device__ __forceinline__ void cell_edges(const double (&st)[5][ORD], const CellBg<NGLL>& bg, bool weno_s,
                                                    bool weno_w, const SweepK& K, double (&eL)[5], double (&eR)[5],
                                                    double& src_avg) {
  constexpr int NA = NGLL;
  constexpr int NM = 1 + DIR;
  double q[5][NA][NGLL];   // r, ru, rv, rw, rt
  double f[4][NA][NGLL];   // q_nm * (u, v, w, t) in variable order
  double g[NA][NGLL];      // rt_gamma
  double src[NA][NGLL];
  sfor<5>([&](auto l) {
    double gv[NGLL];
    reconstruct<ORD, NGLL>(st[l], (l == idR || l == idT) ? weno_s : weno_w, gv);
    sfor<NGLL>([&](auto ii) { q[l][0][ii] = gv[ii]; });
    if constexpr (l == idR) {
      sfor<NGLL>([&](auto ii) { q[idR][0][ii] += DIR == 2 ? bg.rG[ii] : bg.r; });
    }
    if constexpr (DIR == 2 && l == idW) {        // S:1413-1416
      if (bg.wall_top) q[idW][0][NGLL - 1] = 0.0;
      if (bg.wall_bot) q[idW][0][0] = 0.0;
    }
    if constexpr (l == idT) {
      sfor<NGLL>([&](auto ii) { q[idT][0][ii] += DIR == 2 ? bg.rtG[ii] : bg.rt; });
    }
  });
  sfor<NGLL>([&](auto ii) {
    const double rr = q[idR][0][ii];
    double vel[5];
    sfor<4>([&](auto m) { vel[m + 1] = q[m + 1][0][ii] / rr; });
    sfor<4>([&](auto m) { f[m][0][ii] = rr * vel[NM] * vel[m + 1]; });
    g[0][ii] = eos_pow(rr * vel[idT]);
    if constexpr (DIR == 2) {                      // S:1439: every entry takes the last point's value
      const double sv = -(q[idR][0][ii] - bg.rG[ii]) * GRAV;
      sfor<NA>([&](auto a) { sfor<NGLL>([&](auto b) { src[a][b] = sv; }); });
    }
  });
  sfor<NA - 1>([&](auto km) {
    constexpr int kt = km + 1;
    sfor<NGLL>([&](auto ii) {
      sfor<4>([&](auto m) { f[m][kt][ii] = 0.0; });
      g[kt][ii] = 0.0;
    });
  });
  sfor<NA - 1>([&](auto kt) {
    sfor<NGLL>([&](auto ii) {
      double d[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      sfor<NGLL>([&](auto s) {
        constexpr double D = Deriv<NGLL>::m[s][ii];
        d[0] += D * q[NM][kt][s];
        sfor<4>([&](auto m) {
          constexpr int l = m + 1;
          if constexpr (l == NM) {
            if constexpr (kt == 0) {
              if constexpr (DIR == 2) d[l] += D * (f[m][kt][s] + C0 * g[kt][s] - bg.pG[s]);
              else d[l] += D * (f[m][kt][s] + C0 * g[kt][s]);
            } else {
              d[l] += D * (f[m][kt][s] + C0 * g[kt][s] / 2);
            }
          } else {
            d[l] += D * f[m][kt][s];
          }
        });
      });
      // S:1878-1882 (x: r, ru, rv, rw, rt from d1..d5 with d1 the mass flux)
      sfor<5>([&](auto l) {
        if constexpr (DIR == 2 && l == idW) q[l][kt + 1][ii] = -d[l] / K.dsw / (kt + 1) + src[kt][ii] / (kt + 1);
        else q[l][kt + 1][ii] = -d[l] / K.dsw / (kt + 1.0);
      });
      if constexpr (DIR == 2) {                    // S:2051-2055
        if (bg.wall_top) q[idW][kt + 1][NGLL - 1] = 0.0;
        if (bg.wall_bot) q[idW][kt + 1][0] = 0.0;
        src[kt + 1][ii] = -q[idR][kt + 1][ii] * GRAV;
      }
    });
    sfor<NGLL>([&](auto ii) {
      double t[4] = {0.0, 0.0, 0.0, 0.0};
      sfor<kt + 2>([&](auto l) {
        sfor<4>([&](auto m) {
          t[m] += q[NM][l][ii] * q[m + 1][kt + 1 - l][ii] - q[idR][l][ii] * f[m][kt + 1 - l][ii];
        });
      });
      sfor<4>([&](auto m) { f[m][kt + 1][ii] = t[m] / q[idR][0][ii]; });
      double tg = 0.0;
      sfor<kt + 1>([&](auto l) {
        tg += (kt + 1.0 - l) * (GAMMA * g[l][ii] * q[idT][kt + 1 - l][ii] - q[idT][l][ii] * g[kt + 1 - l][ii]);
      });
      g[kt + 1][ii] = (GAMMA * g[0][ii] * q[idT][kt + 1][ii] + tg / (kt + 1.0)) / q[idT][0][ii];
    });
  });
  // compute_timeAvg_edges (S:2142-2177): dts(kt) * dt^kt / (kt+1), dt^kt by
  // repeated multiplication
  sfor<5>([&](auto l) {
    double a0 = q[l][0][0], a1 = q[l][0][NGLL - 1];
    double dtmult = K.dt;
    sfor<NA - 1>([&](auto km) {
      constexpr int kt = km + 1;
      a0 += q[l][kt][0] * dtmult / (kt + 1.0);
      a1 += q[l][kt][NGLL - 1] * dtmult / (kt + 1.0);
      dtmult *= K.dt;
    });
    eL[l] = a0;
    eR[l] = a1;
  });
  if constexpr (DIR == 2) {                        // S:2130-2138, 1489-1492
    double dtmult = K.dt;
    sfor<NA - 1>([&](auto km) {
      constexpr int kt = km + 1;
      sfor<NGLL>([&](auto ii) { src[0][ii] += src[kt][ii] * dtmult / (kt + 1.0); });
      dtmult *= K.dt;
    });
    double sa = 0.0;
    sfor<NGLL>([&](auto ii) { sa += src[0][ii] * Gll<NGLL>::wts[ii]; });
    src_avg = sa;
  } else {
    src_avg = 0.0;
  }
}

// Upwind characteristic flux at one face (S:842-887, 1202-1245, 1583-1629);
// the z background pressure is subtracted by the caller.
template <int DIR>
__host__ __device__ __forceinline__ void riemann(const double (&L)[5], const double (&R)[5], double (&F)[5]) {
  constexpr int NM = 1 + DIR;
  // the momenta w2 and w3 take, in the reference's order
  constexpr int T1 = DIR == 0 ? 2 : 1;
  constexpr int T2 = DIR == 2 ? 2 : 3;
  const double r = 0.5 * (L[0] + R[0]);
  double vel[5];
  if constexpr (DIR == 1) {                        // conservative averages (S:1210-1214)
    sfor<4>([&](auto m) { vel[m + 1] = 0.5 * (L[m + 1] + R[m + 1]); });
  } else {
    sfor<4>([&](auto m) { vel[m + 1] = 0.5 * (L[m + 1] / L[0] + R[m + 1] / R[0]); });
  }
  const double un = vel[NM], t = vel[4];
  const double p = C0 * eos_pow(r * t);
  const double cs2 = GAMMA * p / r;
  const double cs = ::sqrt(cs2);
  double w1, w2, w3;
  if (un > 0) {
    w1 = L[0] - L[4] / t;
    w2 = L[T1] - vel[T1] * L[4] / t;
    w3 = L[T2] - vel[T2] * L[4] / t;
  } else {
    w1 = R[0] - R[4] / t;
    w2 = R[T1] - vel[T1] * R[4] / t;
    w3 = R[T2] - vel[T2] * R[4] / t;
  }
  const double w5 = un * R[0] / (2 * cs) - R[NM] / (2 * cs) + R[4] / (2 * t);
  const double w6 = -un * L[0] / (2 * cs) + L[NM] / (2 * cs) + L[4] / (2 * t);
  double q[5];
  q[0] = w1 + w5 + w6;
  q[NM] = un * w1 + (un - cs) * w5 + (un + cs) * w6;
  q[T1] = w2 + vel[T1] * w5 + vel[T1] * w6;
  q[T2] = w3 + vel[T2] * w5 + vel[T2] * w6;
  q[4] = t * w5 + t * w6;
  F[0] = q[NM];
  sfor<3>([&](auto m) {
    constexpr int l = m == 0 ? T1 : (m == 1 ? T2 : 4);
    F[l] = q[NM] * q[l] / q[0];
  });
  F[NM] = q[NM] * q[NM] / q[0] + C0 * eos_pow(q[4]);
}

}  // namespace ref
}  // namespace awfl
