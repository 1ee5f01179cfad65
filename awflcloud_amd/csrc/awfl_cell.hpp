// awfl_cell.hpp -- per-cell and per-face numerics of one split sweep, in
// registers: WENO reconstruction to GLL points, ADER differential transform
// of the Euler fluxes, time averaging of the cell-edge states, and the
// upwind characteristic Riemann flux.
//
// Every routine restates the reference's formula in the reference's operation
// order (the FP64 CPU oracle follows the same order), with three arithmetic
// changes that keep the result within a few ulp per operation: divisions by
// constants or by a value used several times become multiplications by a
// reciprocal (compile-time, launch-time, or rcp() refined to ~1 ulp);
// (rho theta)^GAMMA uses pow_gamma(); products and sums contract to FMA.
//   WENO ..................... WenoLimiter.h:12-25, 93-172
//   reconstruction ........... Spatial_euler3d_cons_expl_cart_fv_Agrid.h (S:) 1800-1833
//   level-0 products ......... S:704-715, 1067-1078, 1428-1440
//   differential transform ... S:1837-1916 (x), 1920-2000 (y), 2004-2091 (z)
//   time averages ............ S:2130-2177
//   Riemann fluxes ........... S:842-887 (x), 1202-1245 (y), 1583-1629 (z)
#pragma once
#include <type_traits>

#include "awfl_fastmath.hpp"   // HIP runtime, or host stand-ins for __host__/__device__
#include "awfl_tables.hpp"

namespace awfl {

// reciprocal of a positive normal value: refined FP64 (rcp) or the FP32
// hardware estimate (~1 ulp)
__host__ __device__ __forceinline__ double rcp_t(double x) { return rcp(x); }
__host__ __device__ __forceinline__ float rcp_t(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// phys_params.h:4-10
constexpr double GRAV = 9.81;
constexpr double C0 = 27.5629410929725921310572974482;
constexpr double GAMMA = 1.40027894002789400278940027894;
constexpr double WENO_EPS = 1.0e-20;     // WenoLimiter.h:22, 100

enum { idR = 0, idU = 1, idV = 2, idW = 3, idT = 4 };

// ---------------------------------------------------------------------------
// truncated total-variation indicator: the generated quadratic form of
// coefs_to_tv<N>, terms summed left to right
// ---------------------------------------------------------------------------
template <int N, int M, class T>
__host__ __device__ __forceinline__ T tv_eval(const T (&a)[M]) {
  // grouped by the first coefficient: sum_j a_j (sum_{terms (j,k)} c a_k)
  T r = 0;
  bool first = true;
  sfor<N>([&](auto j) {
    T inner = 0;
    bool any = false;
    sfor<Tv<N>::n>([&](auto t) {
      if constexpr (Tv<N>::j[t] == j) {
        constexpr T c = (T)Tv<N>::c[t];
        constexpr int k = Tv<N>::k[t];
        inner = any ? inner + c * a[k] : c * a[k];
        any = true;
      }
    });
    if (any) {
      r = first ? a[j] * inner : r + a[j] * inner;
      first = false;
    }
  });
  return r;
}

// ex[i] = prod_{j != i} b[j] and P = prod_j b[j] (prefix/suffix products)
template <int N, class T>
__host__ __device__ __forceinline__ void excl_products(const T (&b)[N], T (&ex)[N], T& P) {
  T pre[N], suf[N];
  pre[0] = 1.0;
  sfor<N - 1>([&](auto i) {
    if constexpr (i == 0) pre[1] = b[0];
    else pre[i + 1] = pre[i] * b[i];
  });
  suf[N - 1] = 1.0;
  sfor<N - 1>([&](auto m) {
    constexpr int i = N - 2 - m;
    if constexpr (i == N - 2) suf[i] = b[N - 1];
    else suf[i] = suf[i + 1] * b[i + 1];
  });
  sfor<N>([&](auto i) {
    if constexpr (i == 0) ex[i] = suf[0];
    else if constexpr (i == N - 1) ex[i] = pre[N - 1];
    else ex[i] = pre[i] * suf[i];
  });
  P = pre[N - 1] * b[N - 1];
}

// The bridge polynomial's coefficients as one linear map of the stencil,
// [cell][coef]: (recon[hs+1] - sum_i idl_i recon[i] on cells i..i+hs) /
// idl[hs+1], the three statements of WenoLimiter.h:117-131 (full-stencil
// polynomial, subtraction of the weighted candidates, scaling) folded into
// one table at compile time; its values differ from theirs by rounding only.
template <int ORD>
struct BridgeMap {
  double m[ORD][ORD];
  constexpr BridgeMap() : m{} {
    using W = Weno<ORD>;
    constexpr int hs = W::hs;
    for (int s = 0; s < ORD; s++)
      for (int ii = 0; ii < ORD; ii++) {
        double t = W::recon[hs + 1][s][ii];
        if (ii <= hs)
          for (int i = 0; i <= hs; i++)
            if (s - i >= 0 && s - i <= hs) t -= W::idl[i] * W::recon[i][s - i][ii];
        m[s][ii] = t * (1.0 / W::idl[hs + 1]);
      }
  }
};
template <int ORD>
inline constexpr BridgeMap<ORD> kBridge{};

// The bridge map is mirror-symmetric: column c (coefficient c) has
// m[hs+k][c] = (-1)^c m[hs-k][c] (odd columns vanish at the centre), so it is
// applied to the sums u[hs+k] + u[hs-k] (even c) or differences (odd c).
// kind[c]: 1 symmetric, -1 antisymmetric, 0 neither (the plain map is used);
// cm[c] the centre entry, cf[c][k] the coefficient of pair k.
template <int ORD>
struct BridgeMirror {
  int kind[ORD];
  double cm[ORD];
  double cf[ORD][ORD];
  constexpr BridgeMirror() : kind(), cm(), cf() {
    constexpr int hs = (ORD - 1) / 2;
    for (int c = 0; c < ORD; c++) {
      double mx = 0.0;
      for (int s = 0; s < ORD; s++) {
        const double v = kBridge<ORD>.m[s][c];
        mx = v > mx ? v : (-v > mx ? -v : mx);
      }
      const double tol = 1e-12 * mx;
      bool sym = true, anti = true;
      for (int k = 1; k <= hs; k++) {
        const double p = kBridge<ORD>.m[hs + k][c], q = kBridge<ORD>.m[hs - k][c];
        const double ds = p - q, da = p + q;
        if (ds > tol || -ds > tol) sym = false;
        if (da > tol || -da > tol) anti = false;
      }
      const double mid = kBridge<ORD>.m[hs][c];
      if (mid > tol || -mid > tol) anti = false;
      kind[c] = sym ? 1 : (anti ? -1 : 0);
      cm[c] = sym ? mid : 0.0;
      for (int k = 1; k <= hs; k++)
        cf[c][k] = sym ? 0.5 * (kBridge<ORD>.m[hs + k][c] + kBridge<ORD>.m[hs - k][c])
                       : 0.5 * (kBridge<ORD>.m[hs + k][c] - kBridge<ORD>.m[hs - k][c]);
    }
  }
};
template <int ORD>
inline constexpr BridgeMirror<ORD> kBridgeMirror{};

// Every candidate polynomial, the full-stencil polynomial and so the bridge
// and any convex combination of them have the centre cell's average over the
// centre cell [-1/2, 1/2]: a0 + sum_{c even >= 2} a_c / ((c + 1) 2^c) = u_hs.
// So a reconstruction's constant coefficient need not be formed: its GLL
// values are u_hs + sum_{c >= 1} (x^c - [c even] / ((c + 1) 2^c)) a_c.  The
// map m[c][ii] holds those coefficients (row 0 unused).
template <int ORD, int NGLL>
struct AvgProjMap {
  double m[ORD][NGLL];
  constexpr AvgProjMap() : m() {
    using P = Proj<ORD, NGLL>;
    for (int c = 0; c < ORD; c++) {
      double k = 0.0;
      if (c % 2 == 0) {
        double p2 = 1.0;
        for (int e = 0; e < c; e++) p2 *= 2.0;
        k = 1.0 / ((c + 1) * p2);
      }
      for (int ii = 0; ii < NGLL; ii++) m[c][ii] = c == 0 ? 0.0 : P::c2g[c][ii] - P::c2g[0][ii] * k;
    }
  }
};
template <int ORD, int NGLL>
inline constexpr AvgProjMap<ORD, NGLL> kAvgProj{};

// ---------------------------------------------------------------------------
// weno::compute_weno_coefs: ord cell averages u -> ord polynomial coefficients
// ---------------------------------------------------------------------------
// NOAVG: the constant coefficients (candidates, bridge and aw[0]) are not
// formed (aw[0] is left unset); see AvgProjMap.
template <int ORD, class T, bool NOAVG = false>
__host__ __device__ __forceinline__ void weno_coefs(const T (&u)[ORD], T (&aw)[ORD]) {
  using W = Weno<ORD>;
  constexpr int hs = W::hs;
  T a[hs + 2][ORD];
  // hs+1 degree-hs candidates on sub-stencils i..i+hs (WenoLimiter.h:110-116)
  sfor<hs + 1>([&](auto i) {
    sfor<hs + 1>([&](auto ii) {
      if constexpr (NOAVG && ii == 0) return;
      T t = -0.0;
      sfor<hs + 1>([&](auto s) {
        constexpr T cf = (T)W::recon[i][s][ii];
        if constexpr (cf != 0) t += cf * u[i + s];
      });
      a[i][ii] = t;
    });
  });
  // "bridge" polynomial (:117-131): the full-stencil polynomial less the
  // weighted candidates, scaled, as one constant map of the stencil
#if AWFL_PARITY & AWFL_PAR_BRIDGE
  sfor<ORD>([&](auto ii) {
    if constexpr (NOAVG && ii == 0) return;
    T t = -0.0;
    sfor<ORD>([&](auto s) { t += (T)W::recon[hs + 1][s][ii] * u[s]; });
    if constexpr (ii <= hs) sfor<hs + 1>([&](auto i) { t -= (T)W::idl[i] * a[i][ii]; });
    a[hs + 1][ii] = t / (T)W::idl[hs + 1];
  });
#else
  T usum[hs + 1], udif[hs + 1];
  sfor<hs>([&](auto km) {
    constexpr int k = km + 1;
    usum[k] = u[hs + k] + u[hs - k];
    udif[k] = u[hs + k] - u[hs - k];
  });
  sfor<ORD>([&](auto ii) {
    if constexpr (NOAVG && ii == 0) return;
    T t = -0.0;
    if constexpr (kBridgeMirror<ORD>.kind[ii] == 1) {
      if constexpr (kBridgeMirror<ORD>.cm[ii] != 0) t = (T)kBridgeMirror<ORD>.cm[ii] * u[hs];
      sfor<hs>([&](auto km) {
        constexpr T cf = (T)kBridgeMirror<ORD>.cf[ii][km + 1];
        if constexpr (cf != 0) t += cf * usum[km + 1];
      });
    } else if constexpr (kBridgeMirror<ORD>.kind[ii] == -1) {
      sfor<hs>([&](auto km) {
        constexpr T cf = (T)kBridgeMirror<ORD>.cf[ii][km + 1];
        if constexpr (cf != 0) t += cf * udif[km + 1];
      });
    } else {
      sfor<ORD>([&](auto s) {
        constexpr T cf = (T)kBridge<ORD>.m[s][ii];
        if constexpr (cf != 0) t += cf * u[s];
      });
    }
    a[hs + 1][ii] = t;
  });
#endif
  // total variation of every candidate (:133-143), sigma blend (:145-151)
  T tv[hs + 2];
  sfor<hs + 1>([&](auto i) { tv[i] = tv_eval<hs + 1>(a[i]); });
  tv[hs + 1] = tv_eval<ORD>(a[hs + 1]);
  T lo_avg = 0;
  sfor<hs + 1>([&](auto i) { lo_avg += tv[i]; });
  lo_avg *= (T)(1.0 / (hs + 1));
  constexpr T sigma = (T)W::sigma;
  tv[hs + 1] = lo_avg + (tv[hs + 1] - lo_avg) * sigma;
  // weights ~ idl / TV^2, convexify, map, convexify (:153-161).  Each
  // normalisation w_i = (a_i / b_i) / (sum_k a_k / b_k + eps) is formed over
  // the common denominator: a_i prod_{j!=i} b_j / (sum_k a_k prod_{j!=k} b_j
  // + eps prod_j b_j), one reciprocal instead of hs+3 divisions.
  T w[hs + 2];
  if constexpr (std::is_same<T, double>::value) {
    {
      T D[hs + 2], ex[hs + 2], P;
      sfor<hs + 2>([&](auto i) { D[i] = tv[i] * tv[i] + WENO_EPS; });
      excl_products<hs + 2>(D, ex, P);
      T N = 0.0;
      sfor<hs + 2>([&](auto i) {
        constexpr double d = W::idl[i];
        w[i] = d * ex[i];
        N += w[i];
      });
      T rs = rcp(fma(WENO_EPS, P, N));
      sfor<hs + 2>([&](auto i) { w[i] *= rs; });
    }
    {
      T A[hs + 2], E[hs + 2], ex[hs + 2], P;
      sfor<hs + 2>([&](auto i) {
        constexpr double d = W::idl[i];
        A[i] = w[i] * (d + d * d - 3.0 * d * w[i] + w[i] * w[i]);
        E[i] = d * d + w[i] * (1.0 - 2.0 * d);
      });
      excl_products<hs + 2>(E, ex, P);
      T M = 0.0;
      sfor<hs + 2>([&](auto i) {
        w[i] = A[i] * ex[i];
        M += w[i];
      });
      T rs = rcp(fma(WENO_EPS, P, M));
      sfor<hs + 2>([&](auto i) { w[i] *= rs; });
    }
  } else {
    // FP32 (and the tangent-linear scalar): products of the eps-sized TV
    // denominators would underflow, so the first normalisation takes one
    // reciprocal per weight; in FP32 the map's denominators d^2 + w (1 - 2d),
    // bounded below by min(d, 1-d)^2, share one reciprocal
    sfor<hs + 2>([&](auto i) {
      constexpr T d = (T)W::idl[i];
      w[i] = d * rcp_t(tv[i] * tv[i] + (T)WENO_EPS);
    });
    T sum = 0;
    sfor<hs + 2>([&](auto i) { sum += w[i]; });
    T rs = rcp_t(sum + (T)WENO_EPS);
    sfor<hs + 2>([&](auto i) { w[i] *= rs; });
    if constexpr (std::is_same<T, float>::value) {
      T A[hs + 2], E[hs + 2], ex[hs + 2], P;
      sfor<hs + 2>([&](auto i) {
        constexpr T d2 = (T)(W::idl[i] + W::idl[i] * W::idl[i]);
        constexpr T dd = (T)(W::idl[i] * W::idl[i]);
        constexpr T c3 = (T)(3.0 * W::idl[i]);
        constexpr T c12 = (T)(1.0 - 2.0 * W::idl[i]);
        A[i] = w[i] * (d2 - c3 * w[i] + w[i] * w[i]);
        E[i] = dd + w[i] * c12;
      });
      excl_products<hs + 2>(E, ex, P);
      sum = 0;
      sfor<hs + 2>([&](auto i) {
        w[i] = A[i] * ex[i];
        sum += w[i];
      });
      rs = rcp_t(sum + (T)WENO_EPS * P);
      sfor<hs + 2>([&](auto i) { w[i] *= rs; });
    } else {
      // the tangent-linear scalar: the map as written, the statements the
      // hand-written reverse follows
      sfor<hs + 2>([&](auto i) {
        constexpr T d = (T)W::idl[i];
        w[i] = w[i] * (d + d * d - (T)3 * d * w[i] + w[i] * w[i]) * rcp_t(d * d + w[i] * ((T)1 - (T)2 * d));
      });
      sum = 0;
      sfor<hs + 2>([&](auto i) { sum += w[i]; });
      rs = rcp_t(sum + (T)WENO_EPS);
      sfor<hs + 2>([&](auto i) { w[i] *= rs; });
    }
  }
  // weighted sum (:163-171); candidates have no coefficients above hs
  sfor<ORD>([&](auto ii) {
    if constexpr (NOAVG && ii == 0) return;
    T t = -0.0;
    sfor<hs + 2>([&](auto i) {
      if constexpr (i == hs + 1 || ii <= hs) t += w[i] * a[i][ii];
    });
    aw[ii] = t;
  });
}

// reconstruct_gll_values (S:1800-1833): ord stencil -> NGLL point values
template <int ORD, int NGLL, class T>
__host__ __device__ __forceinline__ void reconstruct(const T (&u)[ORD], bool doweno, T (&g)[NGLL]) {
  using P = Proj<ORD, NGLL>;
  if (doweno) {
    // the weighted polynomial's GLL values from the centre cell's average and
    // its non-constant coefficients (AvgProjMap; c2g[0][.] = 1)
    T aw[ORD];
    weno_coefs<ORD, T, true>(u, aw);
    sfor<NGLL>([&](auto ii) {
      T t = u[(ORD - 1) / 2];
      sfor<ORD>([&](auto s) {
        if constexpr (s > 0) {
          constexpr T cf = (T)kAvgProj<ORD, NGLL>.m[s][ii];
          if constexpr (cf != 0) t += cf * aw[s];
        }
      });
      g[ii] = t;
    });
  } else {
    sfor<NGLL>([&](auto ii) {
      T t = -0.0;
      sfor<ORD>([&](auto s) {
        constexpr T cf = (T)P::s2g[s][ii];
        if constexpr (cf != 0) t += cf * u[s];
      });
      g[ii] = t;
    });
  }
}

// Background seen by one cell of a sweep.  X and Y add the cell averages
// hyDensCells/hyDensThetaCells (S:680, 697, 1043, 1060); Z adds the GLL point
// values and subtracts the background pressure (S:1400, 1421, 2042).
// Each background rho*theta value B also comes with B^GAMMA as a double-double
// (h, l) and 1/B, the reference point of pow_gamma_near.
struct PowRef {
  double h, l, rb, b;
};

template <int NGLL>
struct CellBg {
  double r, rt;            // x/y: cell averages of the background
  PowRef pt;               // x/y: power reference of rt
  double rG[NGLL];         // z: hyDensGLL(k, :)
  double rtG[NGLL];        // z: hyDensThetaGLL(k, :)
  double pG[NGLL];         // z: hyPressureGLL(k, :)
  PowRef ptG[NGLL];        // z: power references of rtG
  bool wall_bot, wall_top; // z: solid wall and k == 0 / k == nz-1
};

// power references of a face: for (r t) of the averaged state and for the
// recomposed q5 of the Riemann flux
struct FaceRef {
  PowRef rt, q5;
};

// Launch constants of one sweep: dt, 1/d along the sweep, the differential
// transform factors -1/(d (kt+1)) and the time-average weights dt^kt/(kt+1).
constexpr int KMAX = 9;
struct SweepK {
  double dt;
  double rd;              // 1 / dsw
  double ndt[KMAX];       // -1 / (dsw (kt+1)), kt = 0..NA-2
  double tav[KMAX];       // dt^kt / (kt+1), kt = 1..NA-1 (tav[0] unused)
  double dsw;             // grid spacing along the sweep (reference-order parity builds)
  unsigned char* br;      // upwind branch per face (record / replay kernels only)
};

// ---------------------------------------------------------------------------
// One cell of a DIR sweep (0 x, 1 y, 2 z): stencils of the 5 state variables
// along the sweep axis -> time-averaged left/right edge states (and, for z,
// the time- and cell-averaged gravity source of rho*w).
// ---------------------------------------------------------------------------
// levels(q) receives the transform levels q[var][level][point] (level 0 with
// the background added) once they are complete: the tracer sweeps take the
// density and normal-momentum levels from it (S:757-758).
template <int ORD, int NGLL, int DIR, class Levels>
__host__ __device__ __forceinline__ void cell_edges(const double (&st)[5][ORD], const CellBg<NGLL>& bg, bool weno_s,
                                           bool weno_w, const SweepK& K, double (&eL)[5], double (&eR)[5],
                                           double& src_avg, Levels&& levels) {
  constexpr int NA = NGLL;             // nAder = ngll (Temporal_ader.h:8)
  constexpr int NM = 1 + DIR;          // momentum normal to the sweep
  double q[5][NA][NGLL];               // r, ru, rv, rw, rt at every level
  double f[4][NA][NGLL];               // q_nm*var_l/r for l = 1..4
  double g[NA][NGLL];                  // (r t)^gamma, levels >= 1 carried doubled (S:1904-1915)
  double src[NA][NGLL];                // z gravity source
  double rr[NGLL], rrt[NGLL];          // 1/r and 1/(r t) at level 0

  // level 0: reconstruction + background
  double et[NGLL];                     // (rho theta)'/B at the GLL points
  sfor<5>([&](auto l) {
    double gv[NGLL];
    reconstruct<ORD, NGLL>(st[l], (l == idR || l == idT) ? weno_s : weno_w, gv);
    sfor<NGLL>([&](auto ii) { q[l][0][ii] = gv[ii]; });
  });
  sfor<NGLL>([&](auto ii) {
    if constexpr (DIR == 2) {
      et[ii] = q[idT][0][ii] * bg.ptG[ii].rb;
      q[idR][0][ii] += bg.rG[ii];
      q[idT][0][ii] += bg.rtG[ii];
    } else {
      et[ii] = q[idT][0][ii] * bg.pt.rb;
      q[idR][0][ii] += bg.r;
      q[idT][0][ii] += bg.rt;
    }
  });
  if constexpr (DIR == 2) {            // S:1413-1416
    if (bg.wall_top) q[idW][0][NGLL - 1] = 0.0;
    if (bg.wall_bot) q[idW][0][0] = 0.0;
  }
  double sv = 0.0;
  sfor<NGLL>([&](auto ii) {
    double r = q[idR][0][ii];
    rr[ii] = rcp(r);
    double vel[5];
    sfor<4>([&](auto m) { vel[m + 1] = q[m + 1][0][ii] * rr[ii]; });
    sfor<4>([&](auto m) { f[m][0][ii] = r * vel[NM] * vel[m + 1]; });
    if constexpr (NA > 1) {
      // (r t)^gamma: r t = rho*theta up to an ulp
      const PowRef& pr = DIR == 2 ? bg.ptG[ii] : bg.pt;
      g[0][ii] = pow_gamma_near(q[idT][0][ii], et[ii], pr.h, pr.l);
      rrt[ii] = rcp(q[idT][0][ii]);
    }
    if constexpr (DIR == 2) sv = -(q[idR][0][ii] - bg.rG[ii]) * GRAV;
  });
  if constexpr (DIR == 2) {
    // S:1439 assigns a scalar to the whole SArray: every entry holds the value
    // of the last GLL point
    sfor<NA>([&](auto kt) { sfor<NGLL>([&](auto ii) { src[kt][ii] = sv; }); });
  }

  // differential transform: levels 1..NA-1
  sfor<NA - 1>([&](auto kt) {
    constexpr double rkt1 = 1.0 / (kt + 1.0);
    const double ndt = K.ndt[kt];
    sfor<NGLL>([&](auto ii) {
      double d[5];
      sfor<5>([&](auto l) { d[l] = -0.0; });
      sfor<NGLL>([&](auto s) {
        constexpr double D = Deriv<NGLL>::m[s][ii];
        if constexpr (D == 0.0) return;
        d[0] += D * q[NM][kt][s];
        sfor<4>([&](auto m) {
          constexpr int l = m + 1;
          if constexpr (l == NM) {
            if constexpr (kt == 0) {
              if constexpr (DIR == 2) d[l] += D * (f[m][kt][s] + C0 * g[kt][s] - bg.pG[s]);
              else d[l] += D * (f[m][kt][s] + C0 * g[kt][s]);
            } else {
              d[l] += D * (f[m][kt][s] + (0.5 * C0) * g[kt][s]);
            }
          } else {
            d[l] += D * f[m][kt][s];
          }
        });
      });
      sfor<5>([&](auto l) {
        if constexpr (DIR == 2 && l == idW) q[l][kt + 1][ii] = d[l] * ndt + src[kt][ii] * rkt1;
        else q[l][kt + 1][ii] = d[l] * ndt;
      });
      if constexpr (DIR == 2) {        // S:2051-2055
        if (bg.wall_top) q[idW][kt + 1][NGLL - 1] = 0.0;
        if (bg.wall_bot) q[idW][kt + 1][0] = 0.0;
        src[kt + 1][ii] = -q[idR][kt + 1][ii] * GRAV;
      }
    });
    if constexpr (kt + 2 < NA) {       // products at level kt+1 are used only below the last level
      sfor<NGLL>([&](auto ii) {
        sfor<4>([&](auto m) {
          constexpr int l = m + 1;
          double tot = 0.0;
          sfor<kt + 2>([&](auto ir) {
            if constexpr (ir == 0) tot += q[NM][ir][ii] * q[l][kt + 1 - ir][ii];   // f(kt+1) is still 0
            else tot += q[NM][ir][ii] * q[l][kt + 1 - ir][ii] - q[idR][ir][ii] * f[m][kt + 1 - ir][ii];
          });
          f[m][kt + 1][ii] = tot * rr[ii];
        });
        double tg = 0.0;
        sfor<kt + 1>([&](auto ir) {
          constexpr double wgt = kt + 1.0 - ir;
          if constexpr (ir == 0)          // g(kt+1) is still 0
            tg += wgt * (GAMMA * g[ir][ii] * q[idT][kt + 1 - ir][ii]);
          else
            tg += wgt * (GAMMA * g[ir][ii] * q[idT][kt + 1 - ir][ii] - q[idT][ir][ii] * g[kt + 1 - ir][ii]);
        });
        g[kt + 1][ii] = (GAMMA * g[0][ii] * q[idT][kt + 1][ii] + tg * rkt1) * rrt[ii];
      });
    }
  });

  levels(q);
  // time averages of the edge values (S:2142-2177)
  sfor<5>([&](auto l) {
    double a0 = q[l][0][0], a1 = q[l][0][NGLL - 1];
    sfor<NA - 1>([&](auto km) {
      constexpr int kt = km + 1;
      a0 += q[l][kt][0] * K.tav[kt];
      a1 += q[l][kt][NGLL - 1] * K.tav[kt];
    });
    eL[l] = a0;
    eR[l] = a1;
  });
  if constexpr (DIR == 2) {            // S:1462, 1489-1492
    sfor<NA - 1>([&](auto km) {
      constexpr int kt = km + 1;
      sfor<NGLL>([&](auto ii) { src[0][ii] += src[kt][ii] * K.tav[kt]; });
    });
    double sa = 0.0;
    sfor<NGLL>([&](auto ii) { sa += src[0][ii] * Gll<NGLL>::wts[ii]; });
    src_avg = sa;
  } else {
    src_avg = 0.0;
  }
}

template <int ORD, int NGLL, int DIR>
__host__ __device__ __forceinline__ void cell_edges(const double (&st)[5][ORD], const CellBg<NGLL>& bg, bool weno_s,
                                           bool weno_w, const SweepK& K, double (&eL)[5], double (&eR)[5],
                                           double& src_avg) {
  cell_edges<ORD, NGLL, DIR>(st, bg, weno_s, weno_w, K, eL, eR, src_avg, [](const auto&) {});
}

// ---------------------------------------------------------------------------
// Upwind characteristic flux at one face.  X and Z average primitive
// variables; Y averages the conservative variables (S:1210-1214).  The
// normal velocity test is strict (u > 0).
// ---------------------------------------------------------------------------
// un_out receives the averaged normal velocity the flux upwinds with (the
// tracer fluxes upwind with the same value, S:891-900).  With FORCE the
// upwind side of waves 1-3 is up_forced (true: left) instead of the sign
// test: the flux with a frozen branch, for difference checks of the adjoint.
template <int DIR, bool FORCE = false>
__host__ __device__ __forceinline__ void riemann(const double (&L)[5], const double (&R)[5], const FaceRef& fr,
                                        double (&F)[5], double& un_out, bool up_forced = false) {
  constexpr int NM = 1 + DIR;
  // tangential momentum indices in the order the reference uses them
  constexpr int T1 = DIR == 0 ? 2 : 1;
  constexpr int T2 = DIR == 2 ? 2 : 3;
  double r = 0.5 * (L[0] + R[0]);
  double un, ut1, ut2, t;
  if constexpr (DIR == 1) {
    un = 0.5 * (L[NM] + R[NM]);
    ut1 = 0.5 * (L[T1] + R[T1]);
    ut2 = 0.5 * (L[T2] + R[T2]);
    t = 0.5 * (L[4] + R[4]);
  } else {
    double rL = rcp(L[0]), rR = rcp(R[0]);
    un = 0.5 * (L[NM] * rL + R[NM] * rR);
    ut1 = 0.5 * (L[T1] * rL + R[T1] * rR);
    ut2 = 0.5 * (L[T2] * rL + R[T2] * rR);
    t = 0.5 * (L[4] * rL + R[4] * rR);
  }
  double p = C0 * pow_gamma_near(r * t, fma(r, t, -fr.rt.b) * fr.rt.rb, fr.rt.h, fr.rt.l);
  double cs2 = GAMMA * p * rcp(r);
  double cs, r2c;
  sqrt_halfrsqrt(cs2, cs, r2c);
  double rt = rcp(t);
  double w1, w2, w3;
  if (FORCE ? up_forced : un > 0) {
    double a = L[4] * rt;
    w1 = L[0] - a;
    w2 = L[T1] - ut1 * a;
    w3 = L[T2] - ut2 * a;
  } else {
    double a = R[4] * rt;
    w1 = R[0] - a;
    w2 = R[T1] - ut1 * a;
    w3 = R[T2] - ut2 * a;
  }
  double w5 = (un * R[0] - R[NM]) * r2c + 0.5 * R[4] * rt;
  double w6 = (L[NM] - un * L[0]) * r2c + 0.5 * L[4] * rt;
  double q1 = w1 + w5 + w6;
  double qn = un * w1 + (un - cs) * w5 + (un + cs) * w6;
  double qt1 = w2 + ut1 * w5 + ut1 * w6;
  double qt2 = w3 + ut2 * w5 + ut2 * w6;
  double q5 = t * w5 + t * w6;
  double rq1 = rcp(q1);
  F[0] = qn;
  F[NM] = qn * qn * rq1 + C0 * pow_gamma_near(q5, (q5 - fr.q5.b) * fr.q5.rb, fr.q5.h, fr.q5.l);
  F[T1] = qn * qt1 * rq1;
  F[T2] = qn * qt2 * rq1;
  F[4] = qn * q5 * rq1;
  un_out = un;
}

template <int DIR>
__host__ __device__ __forceinline__ void riemann(const double (&L)[5], const double (&R)[5], const FaceRef& fr,
                                        double (&F)[5]) {
  double un;
  riemann<DIR>(L, R, fr, F, un);
}

}  // namespace awfl
