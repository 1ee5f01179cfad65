// awfl_cell_pert.hpp -- the split-sweep numerics in FP32 perturbation form.
//
// The state is carried as FP32 deviations from the FP64 hydrostatic
// background: rho', rho*u, rho*v, rho*w, (rho theta)' -- the same variables
// the FP64 path stores.  WENO, the ADER differential transform and the
// Riemann flux run in FP32 on deviations wherever a full value would cancel:
//   * pressure only ever appears as p' = pbar ((1 + e)^GAMMA - 1),
//     e = (rho theta)'/(rho theta)bar, from a binomial series -- never as
//     C0 (rho theta)^GAMMA - pbar (catastrophic in FP32 at ~1e5 Pa);
//   * the x/y flux drops the level-constant background pressure (it cancels
//     exactly in the flux difference); the z flux is already written with the
//     background pressure subtracted (S:1628-1629, 2042);
//   * the characteristic decomposition and recomposition of the Riemann flux
//     are linear in the state for fixed eigen-parameters, so they are applied
//     to the deviation vector; the background recomposes to itself.
// Full density and rho*theta (background + deviation, FP32) appear only where
// the reference divides by them or forms products with them.  The FP64
// background enters as FP32 per-level / per-point constants.
#pragma once
#include <type_traits>

#include "awfl_cell.hpp"

namespace awfl {
namespace pert {

constexpr float C0F = (float)C0;
constexpr float GAMMAF = (float)GAMMA;
constexpr float GRAVF = (float)GRAV;

// (1 + e)^GAMMA - 1: binomial series (degree 5: the first omitted term is
// below 5e-9 of the result for |e| <= 1/16, a tenth of an FP32 ulp), else
// exp2/log2 of the hardware.
__device__ __forceinline__ float powm1_gamma(float e) {
  if (fabsf(e) <= 0.0625f) {
    float q = (float)BINOM_GAMMA.c[5];
    q = fmaf(q, e, (float)BINOM_GAMMA.c[4]);
    q = fmaf(q, e, (float)BINOM_GAMMA.c[3]);
    q = fmaf(q, e, (float)BINOM_GAMMA.c[2]);
    q = fmaf(q, e, (float)BINOM_GAMMA.c[1]);
    return q * e;
  }
  return __builtin_amdgcn_exp2f(GAMMAF * __builtin_amdgcn_logf(1.0f + e)) - 1.0f;
}

// background of one GLL point: density, rho*theta, pressure, 1/(rho theta)
struct PointBg {
  float r, rt, p, rrt;
};

template <int NGLL>
struct CellBgP {
  PointBg pt[NGLL];        // x/y: all points equal the level's cell averages (S:680, 697)
  bool wall_bot, wall_top; // z: solid wall and k == 0 / k == nz-1
};

// Launch constants in FP32
// Launch constants in FP32.  dn[kt][s][ii] = D[s][ii] * ndt[kt]: the GLL
// derivative matrix with the transform factor -1/(d (kt+1)) folded in
// (ngll <= 3 here).
constexpr int PK_NA = 3, PK_NG = 3;
struct SweepKF {
  float dt, rd;
  float ndt[KMAX];
  float tav[KMAX];
  float dn[PK_NA - 1][PK_NG][PK_NG];
};

// One cell of a DIR sweep in perturbation form: deviation stencils -> deviation
// edge states (time averaged) and the z gravity source.
template <int ORD, int NGLL, int DIR>
__device__ __forceinline__ void cell_edges_pert(const float (&st)[5][ORD], const CellBgP<NGLL>& bg, bool weno_s,
                                                bool weno_w, const SweepKF& K, float (&eL)[5], float (&eR)[5],
                                                float& src_avg) {
  constexpr int NA = NGLL;
  constexpr int NM = 1 + DIR;
  static_assert(NA <= PK_NA && NGLL <= PK_NG, "FP32 launch constants sized for ngll <= 3");
  float q[5][NA][NGLL];    // level 0: deviations for rho and rho*theta; other levels deviations throughout
  float fr0[NGLL], frt0[NGLL];   // full rho and rho*theta at level 0
  float f[4][NA][NGLL];
  float g[NA][NGLL];       // (rho theta)^gamma: level 0 full, higher levels (doubled) deviations
  float pp[NGLL];          // p' at level 0
  float src[NA][NGLL];
  float rr[NGLL], rrt[NGLL];

  sfor<5>([&](auto l) {
    float gv[NGLL];
    reconstruct<ORD, NGLL>(st[l], (l == idR || l == idT) ? weno_s : weno_w, gv);
    sfor<NGLL>([&](auto ii) { q[l][0][ii] = gv[ii]; });
  });
  if constexpr (DIR == 2) {
    if (bg.wall_top) q[idW][0][NGLL - 1] = 0.0f;
    if (bg.wall_bot) q[idW][0][0] = 0.0f;
  }
  float sv = 0.0f;
  sfor<NGLL>([&](auto ii) {
    const PointBg& b = bg.pt[ii];
    fr0[ii] = b.r + q[idR][0][ii];
    frt0[ii] = b.rt + q[idT][0][ii];
    rr[ii] = __builtin_amdgcn_rcpf(fr0[ii]);
    const float un = q[NM][0][ii] * rr[ii];   // r * (q_nm / r) * (q_l / r) = (q_nm / r) * q_l
    f[0][0][ii] = un * q[1][0][ii];
    f[1][0][ii] = un * q[2][0][ii];
    f[2][0][ii] = un * q[3][0][ii];
    f[3][0][ii] = un * frt0[ii];
    float s = powm1_gamma(q[idT][0][ii] * b.rrt);
    pp[ii] = b.p * s;
    g[0][ii] = (b.p + pp[ii]) * (1.0f / C0F);
    rrt[ii] = __builtin_amdgcn_rcpf(frt0[ii]);
    if constexpr (DIR == 2) sv = -q[idR][0][ii] * GRAVF;
  });
  if constexpr (DIR == 2) {
    sfor<NA>([&](auto kt) { sfor<NGLL>([&](auto ii) { src[kt][ii] = sv; }); });   // S:1439
  }

  sfor<NA - 1>([&](auto kt) {
    constexpr float rkt1 = 1.0f / (kt + 1.0f);
    sfor<NGLL>([&](auto ii) {
      float d[5];
      sfor<5>([&](auto l) { d[l] = -0.0f; });
      sfor<NGLL>([&](auto s) {
        if constexpr (Deriv<NGLL>::m[s][ii] == 0.0) return;
        const float D = K.dn[kt][s][ii];
        d[0] += D * q[NM][kt][s];
        sfor<4>([&](auto m) {
          constexpr int l = m + 1;
          if constexpr (l == NM) {
            // the constant-in-the-sweep background pressure has zero derivative
            // (x/y); z subtracts hyPressureGLL (S:2042): both leave p'
            if constexpr (kt == 0) d[l] += D * (f[m][kt][s] + pp[s]);
            else d[l] += D * (f[m][kt][s] + (0.5f * C0F) * g[kt][s]);
          } else {
            d[l] += D * f[m][kt][s];
          }
        });
      });
      sfor<5>([&](auto l) {
        if constexpr (DIR == 2 && l == idW) q[l][kt + 1][ii] = d[l] + src[kt][ii] * rkt1;
        else q[l][kt + 1][ii] = d[l];
      });
      if constexpr (DIR == 2) {
        if (bg.wall_top) q[idW][kt + 1][NGLL - 1] = 0.0f;
        if (bg.wall_bot) q[idW][kt + 1][0] = 0.0f;
        src[kt + 1][ii] = -q[idR][kt + 1][ii] * GRAVF;
      }
    });
    if constexpr (kt + 2 < NA) {
      sfor<NGLL>([&](auto ii) {
        auto full = [&](int l, int lev) -> float {
          if (lev == 0 && l == idR) return fr0[ii];
          if (lev == 0 && l == idT) return frt0[ii];
          return q[l][lev][ii];
        };
        sfor<4>([&](auto m) {
          constexpr int l = m + 1;
          float tot = 0.0f;
          sfor<kt + 2>([&](auto ir) {
            if constexpr (ir == 0) tot += q[NM][ir][ii] * q[l][kt + 1 - ir][ii];
            else tot += q[NM][ir][ii] * full(l, kt + 1 - ir) - full(idR, ir) * f[m][kt + 1 - ir][ii];
          });
          f[m][kt + 1][ii] = tot * rr[ii];
        });
        float tg = 0.0f;
        sfor<kt + 1>([&](auto ir) {
          constexpr float wgt = kt + 1.0f - ir;
          if constexpr (ir == 0)
            tg += wgt * (GAMMAF * g[ir][ii] * q[idT][kt + 1 - ir][ii]);
          else
            tg += wgt * (GAMMAF * g[ir][ii] * q[idT][kt + 1 - ir][ii] - full(idT, ir) * g[kt + 1 - ir][ii]);
        });
        g[kt + 1][ii] = (GAMMAF * g[0][ii] * q[idT][kt + 1][ii] + tg * rkt1) * rrt[ii];
      });
    }
  });

  sfor<5>([&](auto l) {
    float a0 = q[l][0][0], a1 = q[l][0][NGLL - 1];
    sfor<NA - 1>([&](auto km) {
      constexpr int kt = km + 1;
      a0 += q[l][kt][0] * K.tav[kt];
      a1 += q[l][kt][NGLL - 1] * K.tav[kt];
    });
    eL[l] = a0;
    eR[l] = a1;
  });
  if constexpr (DIR == 2) {
    sfor<NA - 1>([&](auto km) {
      constexpr int kt = km + 1;
      sfor<NGLL>([&](auto ii) { src[0][ii] += src[kt][ii] * K.tav[kt]; });
    });
    float sa = 0.0f;
    sfor<NGLL>([&](auto ii) { sa += src[0][ii] * (float)Gll<NGLL>::wts[ii]; });
    src_avg = sa;
  } else {
    src_avg = 0.0f;
  }
}

// background seen by one face
struct FaceBgP {
  float r, rt, p, rrt;     // density, rho*theta, pressure, 1/(rho theta) at the face
  float pq, rq, bq;        // y flux only: p and 1/B of B = r*rt (its average of conservatives)
};

// Upwind characteristic flux in perturbation form: deviation edge states
// -> flux whose difference gives the tendency (normal-momentum flux without
// the face background pressure).
template <int DIR>
__device__ __forceinline__ void riemann_pert(const float (&L)[5], const float (&R)[5], const FaceBgP& fb,
                                             float (&F)[5]) {
  constexpr int NM = 1 + DIR;
  constexpr int T1 = DIR == 0 ? 2 : 1;
  constexpr int T2 = DIR == 2 ? 2 : 3;
  float rL = fb.r + L[0], rR = fb.r + R[0];
  float tL = fb.rt + L[4], tR = fb.rt + R[4];
  float r = 0.5f * (rL + rR);
  float un, ut1, ut2, t, e;
  if constexpr (DIR == 1) {       // averages of conservative variables (S:1210-1214)
    un = 0.5f * (L[NM] + R[NM]);
    ut1 = 0.5f * (L[T1] + R[T1]);
    ut2 = 0.5f * (L[T2] + R[T2]);
    t = 0.5f * (tL + tR);
    e = fmaf(r, t, -fb.bq) * fb.rq;
  } else {
    float iL = __builtin_amdgcn_rcpf(rL), iR = __builtin_amdgcn_rcpf(rR);
    un = 0.5f * (L[NM] * iL + R[NM] * iR);
    ut1 = 0.5f * (L[T1] * iL + R[T1] * iR);
    ut2 = 0.5f * (L[T2] * iL + R[T2] * iR);
    t = 0.5f * (tL * iL + tR * iR);
    e = fmaf(r, t, -fb.rt) * fb.rrt;
  }
  float pb = DIR == 1 ? fb.pq : fb.p;
  float p = pb + pb * powm1_gamma(e);
  // sound speed and its reciprocal from one hardware reciprocal square root
  const float c2 = GAMMAF * p * __builtin_amdgcn_rcpf(r);
  const float rc = __builtin_amdgcn_rsqf(c2);
  float cs = c2 * rc;
  float it = __builtin_amdgcn_rcpf(t);
  float i2c = 0.5f * rc;
  // characteristic amplitudes of the deviation vector
  float w1, w2, w3;
  if (un > 0) {
    float a = L[4] * it;
    w1 = L[0] - a;
    w2 = L[T1] - ut1 * a;
    w3 = L[T2] - ut2 * a;
  } else {
    float a = R[4] * it;
    w1 = R[0] - a;
    w2 = R[T1] - ut1 * a;
    w3 = R[T2] - ut2 * a;
  }
  float w5 = (un * R[0] - R[NM]) * i2c + 0.5f * R[4] * it;
  float w6 = (L[NM] - un * L[0]) * i2c + 0.5f * L[4] * it;
  float q1 = w1 + w5 + w6;                          // deviations of the recomposed state
  float qn = un * w1 + (un - cs) * w5 + (un + cs) * w6;
  float qt1 = w2 + ut1 * w5 + ut1 * w6;
  float qt2 = w3 + ut2 * w5 + ut2 * w6;
  float q5 = t * w5 + t * w6;
  float iq1 = __builtin_amdgcn_rcpf(fb.r + q1);
  F[0] = qn;
  F[NM] = qn * qn * iq1 + fb.p * powm1_gamma(q5 * fb.rrt);
  F[T1] = qn * qt1 * iq1;
  F[T2] = qn * qt2 * iq1;
  F[4] = qn * (fb.rt + q5) * iq1;
}

}  // namespace pert
}  // namespace awfl
