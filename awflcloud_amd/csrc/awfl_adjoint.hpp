// awfl_adjoint.hpp -- tangent-linear and adjoint (reverse) forms of the
// per-cell and per-face numerics of one FP64 split sweep.
//
// The split sweep out = in + dt * T(in) of the forward kernels is a chain of
//   stencil (5 variables x ORD cells) -> WENO -> GLL values -> ADER
//   differential transform -> time-averaged edge states (and the z gravity
//   source), then per face the upwind characteristic flux, then the flux
//   difference.
// cell_fwd / riemann_fwd restate that chain for a generic scalar T: with
// T = double they are the primal (and fill a Tape of the transform levels);
// with T = Dual (value + one directional derivative) they give the exact
// tangent-linear map.  cell_adj / riemann_adj are the hand-written reverse
// sweeps of the same statements (vector-Jacobian products), used by the
// adjoint kernels.  Everything here is host+device so that a host unit test
// can pin the reverse against the tangent (dot-product test) and the tangent
// against finite differences of the production forward (cell_edges/riemann).
#pragma once
#include "awfl_cell.hpp"

namespace awfl {
namespace adj {

#define AWFL_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------------------
// forward-mode scalar
// ---------------------------------------------------------------------------
struct Dual {
  double v, d;
  AWFL_HD constexpr Dual() : v(0.0), d(0.0) {}
  AWFL_HD constexpr Dual(double x) : v(x), d(0.0) {}
  AWFL_HD constexpr Dual(double x, double dx) : v(x), d(dx) {}
  AWFL_HD Dual& operator+=(const Dual& o) { v += o.v; d += o.d; return *this; }
  AWFL_HD Dual& operator-=(const Dual& o) { v -= o.v; d -= o.d; return *this; }
  AWFL_HD Dual& operator*=(const Dual& o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
};
AWFL_HD constexpr Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
AWFL_HD constexpr Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
AWFL_HD constexpr Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
AWFL_HD constexpr Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
AWFL_HD constexpr bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
AWFL_HD constexpr bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
AWFL_HD constexpr bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }

AWFL_HD double val(double x) { return x; }
AWFL_HD double val(const Dual& x) { return x.v; }

AWFL_HD double rcp_g(double x) { return rcp(x); }
AWFL_HD Dual rcp_g(const Dual& x) {
  double r = rcp(x.v);
  return Dual(r, -x.d * r * r);
}
AWFL_HD double pow_g(double x) { return pow_gamma(x); }
AWFL_HD Dual pow_g(const Dual& x) {
  double p = pow_gamma(x.v);
  return Dual(p, GAMMA * p * rcp(x.v) * x.d);
}
// x^GAMMA near a background value B with known power (PowRef, as the
// production kernels): B^GAMMA (1 + (x-B)/B)^GAMMA
AWFL_HD double pow_at(double x, const PowRef& pr) { return pow_gamma_near(x, (x - pr.b) * pr.rb, pr.h, pr.l); }
AWFL_HD Dual pow_at(const Dual& x, const PowRef& pr) {
  double p = pow_gamma_near(x.v, (x.v - pr.b) * pr.rb, pr.h, pr.l);
  return Dual(p, GAMMA * p * rcp(x.v) * x.d);
}
AWFL_HD double sqrt_g(double x) { return sqrt(x); }
AWFL_HD Dual sqrt_g(const Dual& x) {
  double s = sqrt(x.v);
  return Dual(s, 0.5 * x.d * rcp(s));
}

// the WENO of weno_coefs for T = Dual takes its generic (non-double) branch,
// which calls rcp_t (found here by argument-dependent lookup)
AWFL_HD Dual rcp_t(const Dual& x) { return rcp_g(x); }

// ---------------------------------------------------------------------------
// differential-transform levels of one cell, kept for the reverse sweep
// ---------------------------------------------------------------------------
template <int NGLL, class T>
struct Tape {
  T q[5][NGLL][NGLL];      // [var][level][point]; level 0 full state (background added)
  T f[4][NGLL][NGLL];      // q_nm * var_l / r, l = 1..4
  T g[NGLL][NGLL];         // (r t)^gamma, levels >= 1 doubled
  T rr[NGLL], rrt[NGLL];   // 1/r, 1/(r t) at level 0
};

// WENO stage of one cell: the 5 reconstructions to the GLL points
// (deviations from the background)
template <int ORD, int NGLL, class T>
AWFL_HD void cell_recon(const T (&st)[5][ORD], bool weno_s, bool weno_w, T (&q0)[5][NGLL]) {
  sfor<5>([&](auto l) { reconstruct<ORD, NGLL>(st[l], (l == idR || l == idT) ? weno_s : weno_w, q0[l]); });
}

// Transform stage of one cell, generic scalar: from the reconstructed GLL
// values, the statements of cell_edges (awfl_cell.hpp) with pow/rcp/sqrt
// through pow_g/rcp_g.
template <int NGLL, int DIR, class T>
AWFL_HD void cell_dt(const T (&q0)[5][NGLL], const CellBg<NGLL>& bg, const SweepK& K, Tape<NGLL, T>& tp,
                     T (&eL)[5], T (&eR)[5], T& src_avg) {
  constexpr int NA = NGLL;
  constexpr int NM = 1 + DIR;
  auto& q = tp.q;
  auto& f = tp.f;
  auto& g = tp.g;
  T src[NA][NGLL];
  sfor<5>([&](auto l) { sfor<NGLL>([&](auto ii) { q[l][0][ii] = q0[l][ii]; }); });
  sfor<NGLL>([&](auto ii) {
    if constexpr (DIR == 2) {
      q[idR][0][ii] += T(bg.rG[ii]);
      q[idT][0][ii] += T(bg.rtG[ii]);
    } else {
      q[idR][0][ii] += T(bg.r);
      q[idT][0][ii] += T(bg.rt);
    }
  });
  if constexpr (DIR == 2) {
    if (bg.wall_top) q[idW][0][NGLL - 1] = T(0.0);
    if (bg.wall_bot) q[idW][0][0] = T(0.0);
  }
  T sv = T(0.0);
  sfor<NGLL>([&](auto ii) {
    // 1/rho and 1/(rho theta) from one reciprocal of their product
    const T z = rcp_g(q[idR][0][ii] * q[idT][0][ii]);
    tp.rr[ii] = q[idT][0][ii] * z;
    tp.rrt[ii] = q[idR][0][ii] * z;
    sfor<4>([&](auto m) { f[m][0][ii] = q[NM][0][ii] * q[m + 1][0][ii] * tp.rr[ii]; });
    g[0][ii] = pow_at(q[idT][0][ii], DIR == 2 ? bg.ptG[ii] : bg.pt);
    if constexpr (DIR == 2) sv = -(q[idR][0][ii] - T(bg.rG[ii])) * T(GRAV);
  });
  if constexpr (DIR == 2) {
    sfor<NA>([&](auto kt) { sfor<NGLL>([&](auto ii) { src[kt][ii] = sv; }); });
  }
  sfor<NA - 1>([&](auto kt) {
    constexpr double rkt1 = 1.0 / (kt + 1.0);
    const T ndt = T(K.ndt[kt]);
    sfor<NGLL>([&](auto ii) {
      T d[5];
      sfor<5>([&](auto l) { d[l] = T(0.0); });
      sfor<NGLL>([&](auto s) {
        constexpr double D = Deriv<NGLL>::m[s][ii];
        if constexpr (D == 0.0) return;
        d[0] += T(D) * q[NM][kt][s];
        sfor<4>([&](auto m) {
          constexpr int l = m + 1;
          if constexpr (l == NM) {
            if constexpr (kt == 0) {
              if constexpr (DIR == 2) d[l] += T(D) * (f[m][kt][s] + T(C0) * g[kt][s] - T(bg.pG[s]));
              else d[l] += T(D) * (f[m][kt][s] + T(C0) * g[kt][s]);
            } else {
              d[l] += T(D) * (f[m][kt][s] + T(0.5 * C0) * g[kt][s]);
            }
          } else {
            d[l] += T(D) * f[m][kt][s];
          }
        });
      });
      sfor<5>([&](auto l) {
        if constexpr (DIR == 2 && l == idW) q[l][kt + 1][ii] = d[l] * ndt + src[kt][ii] * T(rkt1);
        else q[l][kt + 1][ii] = d[l] * ndt;
      });
      if constexpr (DIR == 2) {
        if (bg.wall_top) q[idW][kt + 1][NGLL - 1] = T(0.0);
        if (bg.wall_bot) q[idW][kt + 1][0] = T(0.0);
        src[kt + 1][ii] = -q[idR][kt + 1][ii] * T(GRAV);
      }
    });
    if constexpr (kt + 2 < NA) {
      sfor<NGLL>([&](auto ii) {
        sfor<4>([&](auto m) {
          constexpr int l = m + 1;
          T tot = T(0.0);
          sfor<kt + 2>([&](auto ir) {
            if constexpr (ir == 0) tot += q[NM][ir][ii] * q[l][kt + 1 - ir][ii];
            else tot += q[NM][ir][ii] * q[l][kt + 1 - ir][ii] - q[idR][ir][ii] * f[m][kt + 1 - ir][ii];
          });
          f[m][kt + 1][ii] = tot * tp.rr[ii];
        });
        T tg = T(0.0);
        sfor<kt + 1>([&](auto ir) {
          constexpr double wgt = kt + 1.0 - ir;
          if constexpr (ir == 0)
            tg += T(wgt) * (T(GAMMA) * g[ir][ii] * q[idT][kt + 1 - ir][ii]);
          else
            tg += T(wgt) * (T(GAMMA) * g[ir][ii] * q[idT][kt + 1 - ir][ii] - q[idT][ir][ii] * g[kt + 1 - ir][ii]);
        });
        g[kt + 1][ii] = (T(GAMMA) * g[0][ii] * q[idT][kt + 1][ii] + tg * T(rkt1)) * tp.rrt[ii];
      });
    }
  });
  sfor<5>([&](auto l) {
    T a0 = q[l][0][0], a1 = q[l][0][NGLL - 1];
    sfor<NA - 1>([&](auto km) {
      constexpr int kt = km + 1;
      a0 += q[l][kt][0] * T(K.tav[kt]);
      a1 += q[l][kt][NGLL - 1] * T(K.tav[kt]);
    });
    eL[l] = a0;
    eR[l] = a1;
  });
  if constexpr (DIR == 2) {
    T sa = T(0.0);
    sfor<NGLL>([&](auto ii) {
      T s0 = src[0][ii];
      sfor<NA - 1>([&](auto km) {
        constexpr int kt = km + 1;
        s0 += src[kt][ii] * T(K.tav[kt]);
      });
      sa += s0 * T(Gll<NGLL>::wts[ii]);
    });
    src_avg = sa;
  } else {
    src_avg = T(0.0);
  }
}

// One cell of a DIR sweep, generic scalar: WENO stage then transform stage
template <int ORD, int NGLL, int DIR, class T>
AWFL_HD void cell_fwd(const T (&st)[5][ORD], const CellBg<NGLL>& bg, bool weno_s, bool weno_w, const SweepK& K,
                      Tape<NGLL, T>& tp, T (&eL)[5], T (&eR)[5], T& src_avg) {
  T q0[5][NGLL];
  cell_recon<ORD, NGLL>(st, weno_s, weno_w, q0);
  cell_dt<NGLL, DIR>(q0, bg, K, tp, eL, eR, src_avg);
}

// Upwind characteristic flux, generic scalar (statements of riemann in
// awfl_cell.hpp); p_face is the background pressure the z flux subtracts.
template <int DIR, class T>
AWFL_HD void riemann_fwd(const T (&L)[5], const T (&R)[5], const FaceRef& fr, double p_face, T (&F)[5]) {
  constexpr int NM = 1 + DIR;
  constexpr int T1 = DIR == 0 ? 2 : 1;
  constexpr int T2 = DIR == 2 ? 2 : 3;
  T r = T(0.5) * (L[0] + R[0]);
  T un, ut1, ut2, t;
  if constexpr (DIR == 1) {
    un = T(0.5) * (L[NM] + R[NM]);
    ut1 = T(0.5) * (L[T1] + R[T1]);
    ut2 = T(0.5) * (L[T2] + R[T2]);
    t = T(0.5) * (L[4] + R[4]);
  } else {
    T iL = rcp_g(L[0]), iR = rcp_g(R[0]);
    un = T(0.5) * (L[NM] * iL + R[NM] * iR);
    ut1 = T(0.5) * (L[T1] * iL + R[T1] * iR);
    ut2 = T(0.5) * (L[T2] * iL + R[T2] * iR);
    t = T(0.5) * (L[4] * iL + R[4] * iR);
  }
  T p = T(C0) * pow_at(r * t, fr.rt);
  T cs = sqrt_g(T(GAMMA) * p * rcp_g(r));
  T it = rcp_g(t);
  T i2c = T(0.5) * rcp_g(cs);
  T w1, w2, w3;
  if (val(un) > 0) {
    T a = L[4] * it;
    w1 = L[0] - a;
    w2 = L[T1] - ut1 * a;
    w3 = L[T2] - ut2 * a;
  } else {
    T a = R[4] * it;
    w1 = R[0] - a;
    w2 = R[T1] - ut1 * a;
    w3 = R[T2] - ut2 * a;
  }
  T w5 = (un * R[0] - R[NM]) * i2c + T(0.5) * R[4] * it;
  T w6 = (L[NM] - un * L[0]) * i2c + T(0.5) * L[4] * it;
  T q1 = w1 + w5 + w6;
  T qn = un * w1 + (un - cs) * w5 + (un + cs) * w6;
  T qt1 = w2 + ut1 * w5 + ut1 * w6;
  T qt2 = w3 + ut2 * w5 + ut2 * w6;
  T q5 = t * w5 + t * w6;
  T iq1 = rcp_g(q1);
  F[0] = qn;
  F[NM] = qn * qn * iq1 + T(C0) * pow_at(q5, fr.q5);
  if constexpr (DIR == 2) F[NM] -= T(p_face);
  F[T1] = qn * qt1 * iq1;
  F[T2] = qn * qt2 * iq1;
  F[4] = qn * q5 * iq1;
}

// ---------------------------------------------------------------------------
// reverse sweeps (FP64)
// ---------------------------------------------------------------------------

// Lb += (dF/dL)^T Fb, Rb += (dF/dR)^T Fb at the state (L, R)
template <int DIR>
AWFL_HD void riemann_adj(const double (&L)[5], const double (&R)[5], const FaceRef& fr, const double (&Fb)[5],
                         double (&Lb)[5], double (&Rb)[5]) {
  constexpr int NM = 1 + DIR;
  constexpr int T1 = DIR == 0 ? 2 : 1;
  constexpr int T2 = DIR == 2 ? 2 : 3;
  // primal
  double r = 0.5 * (L[0] + R[0]);
  double un, ut1, ut2, t, iL = 0, iR = 0;
  if constexpr (DIR == 1) {
    un = 0.5 * (L[NM] + R[NM]);
    ut1 = 0.5 * (L[T1] + R[T1]);
    ut2 = 0.5 * (L[T2] + R[T2]);
    t = 0.5 * (L[4] + R[4]);
  } else {
    iL = rcp(L[0]);
    iR = rcp(R[0]);
    un = 0.5 * (L[NM] * iL + R[NM] * iR);
    ut1 = 0.5 * (L[T1] * iL + R[T1] * iR);
    ut2 = 0.5 * (L[T2] * iL + R[T2] * iR);
    t = 0.5 * (L[4] * iL + R[4] * iR);
  }
  double rt = r * t;
  double p = C0 * pow_at(rt, fr.rt);
  // 1/r, 1/t and 1/(r t) from one reciprocal
  const double irt = rcp(rt);
  double ir = t * irt;
  double cs2 = GAMMA * p * ir;
  double cs, i2c;
  sqrt_halfrsqrt(cs2, cs, i2c);     // as the forward riemann
  double it = r * irt;
  double ics = 2.0 * i2c;
  const bool up = un > 0;
  const double(&U)[5] = up ? L : R;
  double a = U[4] * it;
  double w1 = U[0] - a;
  double w2 = U[T1] - ut1 * a;
  double w3 = U[T2] - ut2 * a;
  double w5 = (un * R[0] - R[NM]) * i2c + 0.5 * R[4] * it;
  double w6 = (L[NM] - un * L[0]) * i2c + 0.5 * L[4] * it;
  double s56 = w5 + w6;
  double q1 = w1 + s56;
  double qn = un * w1 + (un - cs) * w5 + (un + cs) * w6;
  double qt1 = w2 + ut1 * s56;
  double qt2 = w3 + ut2 * s56;
  double q5 = t * s56;
  // 1/q1 and 1/q5 from one reciprocal
  const double iq15 = rcp(q1 * q5);
  double iq1 = q5 * iq15;
  double p5 = C0 * pow_at(q5, fr.q5);
  // reverse
  double qnb = Fb[0] + Fb[NM] * 2.0 * qn * iq1 + Fb[T1] * qt1 * iq1 + Fb[T2] * qt2 * iq1 + Fb[4] * q5 * iq1;
  double iq1b = Fb[NM] * qn * qn + Fb[T1] * qn * qt1 + Fb[T2] * qn * qt2 + Fb[4] * qn * q5;
  double q1b = -iq1b * iq1 * iq1;
  double qt1b = Fb[T1] * qn * iq1;
  double qt2b = Fb[T2] * qn * iq1;
  double q5b = Fb[NM] * GAMMA * p5 * (q1 * iq15) + Fb[4] * qn * iq1;
  double tb = q5b * s56;
  double s56b = q5b * t + qt1b * ut1 + qt2b * ut2 + q1b;
  double w2b = qt1b, w3b = qt2b;
  double ut1b = qt1b * s56, ut2b = qt2b * s56;
  double unb = qnb * (w1 + w5 + w6);
  double w1b = qnb * un + q1b;
  double w5b = qnb * (un - cs) + s56b;
  double w6b = qnb * (un + cs) + s56b;
  double csb = qnb * (w6 - w5);
  // w6, w5
  double i2cb = w6b * (L[NM] - un * L[0]) + w5b * (un * R[0] - R[NM]);
  double itb = w6b * 0.5 * L[4] + w5b * 0.5 * R[4];
  Lb[NM] += w6b * i2c;
  unb += -w6b * L[0] * i2c + w5b * R[0] * i2c;
  Lb[0] += -w6b * un * i2c;
  Lb[4] += w6b * 0.5 * it;
  Rb[0] += w5b * un * i2c;
  Rb[NM] += -w5b * i2c;
  Rb[4] += w5b * 0.5 * it;
  // upwind amplitudes
  double ab = -w1b - w2b * ut1 - w3b * ut2;
  ut1b += -w2b * a;
  ut2b += -w3b * a;
  double(&Ub)[5] = up ? Lb : Rb;
  Ub[0] += w1b;
  Ub[T1] += w2b;
  Ub[T2] += w3b;
  Ub[4] += ab * it;
  itb += ab * U[4];
  // sound speed and pressure
  csb += -i2cb * 0.5 * ics * ics;
  tb += -itb * it * it;
  double cs2b = csb * 0.5 * ics;
  double pb = cs2b * GAMMA * ir;
  double rb = -cs2b * cs2 * ir;
  double rtb = pb * GAMMA * p * irt;
  rb += rtb * t;
  tb += rtb * r;
  // averages
  Lb[0] += 0.5 * rb;
  Rb[0] += 0.5 * rb;
  if constexpr (DIR == 1) {
    Lb[NM] += 0.5 * unb; Rb[NM] += 0.5 * unb;
    Lb[T1] += 0.5 * ut1b; Rb[T1] += 0.5 * ut1b;
    Lb[T2] += 0.5 * ut2b; Rb[T2] += 0.5 * ut2b;
    Lb[4] += 0.5 * tb; Rb[4] += 0.5 * tb;
  } else {
    double iLb = 0.5 * (unb * L[NM] + ut1b * L[T1] + ut2b * L[T2] + tb * L[4]);
    double iRb = 0.5 * (unb * R[NM] + ut1b * R[T1] + ut2b * R[T2] + tb * R[4]);
    Lb[NM] += 0.5 * unb * iL; Rb[NM] += 0.5 * unb * iR;
    Lb[T1] += 0.5 * ut1b * iL; Rb[T1] += 0.5 * ut1b * iR;
    Lb[T2] += 0.5 * ut2b * iL; Rb[T2] += 0.5 * ut2b * iR;
    Lb[4] += 0.5 * tb * iL; Rb[4] += 0.5 * tb * iR;
    Lb[0] += -iLb * iL * iL;
    Rb[0] += -iRb * iR * iR;
  }
}

// Lb += (dF(L, L) / dL)^T Fb at a solid z wall, whose face takes both sides
// from the boundary cell's edge state L (S:1577-1578) and where the wall rows
// of the transform make the normal momentum L[idW] zero.  A consistent flux
// of two equal states is the physical flux (rho w, rho u w, rho v w,
// rho w^2 + p, rho theta w) / rho of L, whose derivative at rho w = 0 is
// nonzero only in d(rho w) and in dp/d(rho theta) = gamma p / (rho theta):
// the same derivative riemann_adj forms term by term, in a few operations.
AWFL_HD void riemann_adj_zwall(const double (&L)[5], const FaceRef& fr, const double (&Fb)[5], double (&Lb)[5]) {
  const double iL = rcp(L[0]);
  const double p5 = C0 * pow_at(L[4], fr.q5);
  Lb[idW] += Fb[0] + (Fb[idU] * L[idU] + Fb[idV] * L[idV] + Fb[idT] * L[idT]) * iL;
  Lb[idT] += Fb[idW] * GAMMA * p5 * rcp(L[4]);
}

// tvb * d tv_eval<N>(a) / da, accumulated into ab
template <int N, int M>
AWFL_HD void tv_adj(const double (&a)[M], double tvb, double (&ab)[M]) {
  sfor<Tv<N>::n>([&](auto t) {
    constexpr int j = Tv<N>::j[t];
    constexpr int k = Tv<N>::k[t];
    constexpr double c = Tv<N>::c[t];
    if constexpr (j == k) {
      ab[j] += 2.0 * c * a[j] * tvb;
    } else {
      ab[j] += c * a[k] * tvb;
      ab[k] += c * a[j] * tvb;
    }
  });
}

// ub += (d reconstruct / du)^T gb: WENO (weno_coefs, then the projection to
// GLL points) or the plain projection of the stencil
template <int ORD, int NGLL>
AWFL_HD void reconstruct_adj(const double (&u)[ORD], bool doweno, const double (&gb)[NGLL], double (&ub)[ORD]) {
  using P = Proj<ORD, NGLL>;
  if (!doweno) {
    sfor<NGLL>([&](auto ii) {
      sfor<ORD>([&](auto s) {
        constexpr double cf = P::s2g[s][ii];
        if constexpr (cf != 0) ub[s] += cf * gb[ii];
      });
    });
    return;
  }
  using W = Weno<ORD>;
  constexpr int hs = W::hs;
  constexpr int NP = hs + 2;
  // primal: candidates, bridge, total variations, weights.  As in the
  // forward (reconstruct), the constant coefficients are never formed: the
  // GLL values are u_hs + sum_{c >= 1} kAvgProj[c][.] aw[c].
  double a[NP][ORD];
  sfor<hs + 1>([&](auto i) {
    a[i][0] = 0.0;
    sfor<hs + 1>([&](auto ii) {
      if constexpr (ii == 0) return;
      double t = 0.0;
      sfor<hs + 1>([&](auto s) {
        constexpr double cf = W::recon[i][s][ii];
        if constexpr (cf != 0) t += cf * u[i + s];
      });
      a[i][ii] = t;
    });
    sfor<ORD - hs - 1>([&](auto m) { a[i][hs + 1 + m] = 0.0; });
  });
  a[hs + 1][0] = 0.0;
  sfor<ORD>([&](auto ii) {
    if constexpr (ii == 0) return;
    double t = 0.0;
    sfor<ORD>([&](auto s) {
      constexpr double cf = kBridge<ORD>.m[s][ii];
      if constexpr (cf != 0) t += cf * u[s];
    });
    a[hs + 1][ii] = t;
  });
  double tv[NP];
  sfor<hs + 1>([&](auto i) {
    double ai[hs + 1];
    sfor<hs + 1>([&](auto ii) { ai[ii] = a[i][ii]; });
    tv[i] = tv_eval<hs + 1>(ai);
  });
  double tvhi = tv_eval<ORD>(a[hs + 1]);
  double lo = 0.0;
  sfor<hs + 1>([&](auto i) { lo += tv[i]; });
  lo *= 1.0 / (hs + 1);
  tv[hs + 1] = lo + (tvhi - lo) * W::sigma;
  // the NP reciprocals 1/(tv^2 + eps) and 1/(d^2 + w1 (1 - 2d)) each from
  // one reciprocal of the product and the products of the others
  double rDn[NP], b[NP], w1[NP], m[NP], w[NP], rden[NP];
  double R1, R2;
  {
    double D[NP], ex[NP], P;
    sfor<NP>([&](auto i) { D[i] = tv[i] * tv[i] + WENO_EPS; });
    excl_products<NP>(D, ex, P);
    // 1/P and 1/S1 = P / (eps P + sum_i d_i ex_i) from one reciprocal
    double N = 0.0;
    sfor<NP>([&](auto i) { N += W::idl[i] * ex[i]; });
    const double T1 = fma(WENO_EPS, P, N);
    const double z = rcp(P * T1);
    const double rP = T1 * z;
    R1 = P * P * z;
    sfor<NP>([&](auto i) {
      constexpr double d = W::idl[i];
      rDn[i] = ex[i] * rP;
      b[i] = d * rDn[i];
    });
  }
  {
    double E[NP], ex[NP], P;
    sfor<NP>([&](auto i) {
      constexpr double d = W::idl[i];
      w1[i] = b[i] * R1;
      E[i] = d * d + w1[i] * (1.0 - 2.0 * d);
    });
    excl_products<NP>(E, ex, P);
    // 1/P and 1/S2 from one reciprocal, as above
    double A[NP], M = 0.0;
    sfor<NP>([&](auto i) {
      constexpr double d = W::idl[i];
      A[i] = w1[i] * (d + d * d - 3.0 * d * w1[i] + w1[i] * w1[i]);
      M += A[i] * ex[i];
    });
    const double T2 = fma(WENO_EPS, P, M);
    const double z = rcp(P * T2);
    const double rP = T2 * z;
    R2 = P * P * z;
    sfor<NP>([&](auto i) {
      rden[i] = ex[i] * rP;
      m[i] = A[i] * rden[i];
    });
  }
  sfor<NP>([&](auto i) { w[i] = m[i] * R2; });
  // reverse: GLL projection (the centre cell's average takes the sum of the
  // GLL adjoints), weighted sum
  double awb[ORD];
  awb[0] = 0.0;
  sfor<NGLL>([&](auto ii) { ub[hs] += gb[ii]; });
  sfor<ORD>([&](auto s) {
    if constexpr (s == 0) return;
    double t = 0.0;
    sfor<NGLL>([&](auto ii) {
      constexpr double cf = kAvgProj<ORD, NGLL>.m[s][ii];
      if constexpr (cf != 0) t += cf * gb[ii];
    });
    awb[s] = t;
  });
  double ab[NP][ORD], wb[NP];
  sfor<NP>([&](auto i) {
    double t = 0.0;
    ab[i][0] = 0.0;
    sfor<ORD>([&](auto ii) {
      if constexpr (ii == 0) {
      } else if constexpr (i == hs + 1 || ii <= hs) {
        t += awb[ii] * a[i][ii];
        ab[i][ii] = w[i] * awb[ii];
      } else {
        ab[i][ii] = 0.0;
      }
    });
    wb[i] = t;
  });
  // second normalisation w = m / (sum m + eps)
  double dot = 0.0;
  sfor<NP>([&](auto i) { dot += wb[i] * w[i]; });
  double w1b[NP];
  sfor<NP>([&](auto i) {
    constexpr double d = W::idl[i];
    double mb = R2 * (wb[i] - dot);
    double x = w1[i];
    // m = num / den: m' = (num' - m (1 - 2d)) / den
    double dnum = d + d * d - 6.0 * d * x + 3.0 * x * x;
    w1b[i] = mb * (dnum - m[i] * (1.0 - 2.0 * d)) * rden[i];
  });
  // first normalisation w1 = b / (sum b + eps), b = idl / (tv^2 + eps)
  double dot1 = 0.0;
  sfor<NP>([&](auto i) { dot1 += w1b[i] * w1[i]; });
  double tvb[NP];
  sfor<NP>([&](auto i) {
    double bb = R1 * (w1b[i] - dot1);
    tvb[i] = -2.0 * bb * b[i] * tv[i] * rDn[i];
  });
  // sigma blend of the high-order indicator
  double tvhib = tvb[hs + 1] * W::sigma;
  double lob = tvb[hs + 1] * (1.0 - W::sigma) * (1.0 / (hs + 1));
  sfor<hs + 1>([&](auto i) {
    double ai[hs + 1], aib[hs + 1];
    sfor<hs + 1>([&](auto ii) {
      ai[ii] = a[i][ii];
      aib[ii] = 0.0;
    });
    tv_adj<hs + 1>(ai, tvb[i] + lob, aib);
    sfor<hs + 1>([&](auto ii) { ab[i][ii] += aib[ii]; });
  });
  tv_adj<ORD>(a[hs + 1], tvhib, ab[hs + 1]);
  // candidate and bridge polynomials (their constant coefficients, never
  // formed, pass nothing back)
  sfor<hs + 1>([&](auto i) {
    sfor<hs + 1>([&](auto ii) {
      if constexpr (ii == 0) return;
      sfor<hs + 1>([&](auto s) {
        constexpr double cf = W::recon[i][s][ii];
        if constexpr (cf != 0) ub[i + s] += cf * ab[i][ii];
      });
    });
  });
  sfor<ORD>([&](auto ii) {
    if constexpr (ii == 0) return;
    sfor<ORD>([&](auto s) {
      constexpr double cf = kBridge<ORD>.m[s][ii];
      if constexpr (cf != 0) ub[s] += cf * ab[hs + 1][ii];
    });
  });
}

// q0b = (d (eL, eR, src_avg) / d q0)^T (eLb, eRb, srcb): reverse of the
// transform stage, recomputing its levels from the GLL values q0
template <int NGLL, int DIR>
AWFL_HD void cell_dt_adj(const double (&q0)[5][NGLL], const CellBg<NGLL>& bg, const SweepK& K,
                         const double (&eLb)[5], const double (&eRb)[5], double srcb, double (&q0b)[5][NGLL]) {
  constexpr int NA = NGLL;
  constexpr int NM = 1 + DIR;
  Tape<NGLL, double> tp;
  double eL[5], eR[5], sa;
  cell_dt<NGLL, DIR>(q0, bg, K, tp, eL, eR, sa);
  const auto& q = tp.q;
  const auto& f = tp.f;
  const auto& g = tp.g;
  double qb[5][NA][NGLL], fb[4][NA][NGLL], gb[NA][NGLL], sb[NA][NGLL], rrb[NGLL], rrtb[NGLL];
  sfor<NA>([&](auto kt) {
    sfor<NGLL>([&](auto ii) {
      sfor<5>([&](auto l) { qb[l][kt][ii] = 0.0; });
      sfor<4>([&](auto m) { fb[m][kt][ii] = 0.0; });
      gb[kt][ii] = 0.0;
      sb[kt][ii] = 0.0;
    });
  });
  sfor<NGLL>([&](auto ii) { rrb[ii] = rrtb[ii] = 0.0; });
  // time averages
  sfor<5>([&](auto l) {
    qb[l][0][0] += eLb[l];
    qb[l][0][NGLL - 1] += eRb[l];
    sfor<NA - 1>([&](auto km) {
      constexpr int kt = km + 1;
      qb[l][kt][0] += eLb[l] * K.tav[kt];
      qb[l][kt][NGLL - 1] += eRb[l] * K.tav[kt];
    });
  });
  if constexpr (DIR == 2) {
    sfor<NGLL>([&](auto ii) {
      double s0b = srcb * Gll<NGLL>::wts[ii];
      sb[0][ii] += s0b;
      sfor<NA - 1>([&](auto km) {
        constexpr int kt = km + 1;
        sb[kt][ii] += s0b * K.tav[kt];
      });
    });
  }
  // transform levels, last first
  sfor<NA - 1>([&](auto km) {
    constexpr int kt = NA - 2 - km;
    constexpr double rkt1 = 1.0 / (kt + 1.0);
    if constexpr (kt + 2 < NA) {
      sfor<NGLL>([&](auto ii) {
        // g[kt+1] = (GAMMA g0 qT[kt+1] + tg/(kt+1)) rrt
        double Gb = gb[kt + 1][ii];
        rrtb[ii] += Gb * g[kt + 1][ii] * q[idT][0][ii];
        double Xb = Gb * tp.rrt[ii];
        gb[0][ii] += Xb * GAMMA * q[idT][kt + 1][ii];
        qb[idT][kt + 1][ii] += Xb * GAMMA * g[0][ii];
        double tgb = Xb * rkt1;
        sfor<kt + 1>([&](auto ir) {
          constexpr double wgt = kt + 1.0 - ir;
          gb[ir][ii] += tgb * wgt * GAMMA * q[idT][kt + 1 - ir][ii];
          qb[idT][kt + 1 - ir][ii] += tgb * wgt * GAMMA * g[ir][ii];
          if constexpr (ir > 0) {
            qb[idT][ir][ii] -= tgb * wgt * g[kt + 1 - ir][ii];
            gb[kt + 1 - ir][ii] -= tgb * wgt * q[idT][ir][ii];
          }
        });
        // f[m][kt+1] = tot rr
        sfor<4>([&](auto m) {
          constexpr int l = m + 1;
          double Fb = fb[m][kt + 1][ii];
          rrb[ii] += Fb * f[m][kt + 1][ii] * q[idR][0][ii];
          double tb = Fb * tp.rr[ii];
          sfor<kt + 2>([&](auto ir) {
            qb[NM][ir][ii] += tb * q[l][kt + 1 - ir][ii];
            qb[l][kt + 1 - ir][ii] += tb * q[NM][ir][ii];
            if constexpr (ir > 0) {
              qb[idR][ir][ii] -= tb * f[m][kt + 1 - ir][ii];
              fb[m][kt + 1 - ir][ii] -= tb * q[idR][ir][ii];
            }
          });
        });
      });
    }
    if constexpr (DIR == 2) {
      sfor<NGLL>([&](auto ii) { qb[idR][kt + 1][ii] -= sb[kt + 1][ii] * GRAV; });
      if (bg.wall_top) qb[idW][kt + 1][NGLL - 1] = 0.0;
      if (bg.wall_bot) qb[idW][kt + 1][0] = 0.0;
    }
    sfor<NGLL>([&](auto ii) {
      double db[5];
      sfor<5>([&](auto l) { db[l] = qb[l][kt + 1][ii] * K.ndt[kt]; });
      if constexpr (DIR == 2) sb[kt][ii] += qb[idW][kt + 1][ii] * rkt1;
      sfor<NGLL>([&](auto s) {
        constexpr double D = Deriv<NGLL>::m[s][ii];
        if constexpr (D == 0.0) return;
        qb[NM][kt][s] += D * db[0];
        sfor<4>([&](auto m) {
          constexpr int l = m + 1;
          fb[m][kt][s] += D * db[l];
          if constexpr (l == NM) gb[kt][s] += D * db[l] * (kt == 0 ? C0 : 0.5 * C0);
        });
      });
    });
  });
  // level 0
  if constexpr (DIR == 2) {
    double svb = 0.0;
    sfor<NGLL>([&](auto ii) { svb += sb[0][ii]; });
    qb[idR][0][NGLL - 1] -= svb * GRAV;
  }
  sfor<NGLL>([&](auto ii) {
    sfor<4>([&](auto m) {
      constexpr int l = m + 1;
      double Fb = fb[m][0][ii];
      qb[NM][0][ii] += Fb * q[l][0][ii] * tp.rr[ii];
      qb[l][0][ii] += Fb * q[NM][0][ii] * tp.rr[ii];
      rrb[ii] += Fb * q[NM][0][ii] * q[l][0][ii];
    });
    qb[idT][0][ii] += gb[0][ii] * GAMMA * g[0][ii] * tp.rrt[ii];
    qb[idR][0][ii] -= rrb[ii] * tp.rr[ii] * tp.rr[ii];
    qb[idT][0][ii] -= rrtb[ii] * tp.rrt[ii] * tp.rrt[ii];
  });
  if constexpr (DIR == 2) {
    if (bg.wall_top) qb[idW][0][NGLL - 1] = 0.0;
    if (bg.wall_bot) qb[idW][0][0] = 0.0;
  }
  sfor<5>([&](auto l) { sfor<NGLL>([&](auto ii) { q0b[l][ii] = qb[l][0][ii]; }); });
}

// stb += (d (eL, eR, src_avg) / d st)^T (eLb, eRb, srcb) for one cell
template <int ORD, int NGLL, int DIR>
AWFL_HD void cell_adj(const double (&st)[5][ORD], const CellBg<NGLL>& bg, bool weno_s, bool weno_w,
                      const SweepK& K, const double (&eLb)[5], const double (&eRb)[5], double srcb,
                      double (&stb)[5][ORD]) {
  double q0[5][NGLL], q0b[5][NGLL];
  cell_recon<ORD, NGLL>(st, weno_s, weno_w, q0);
  cell_dt_adj<NGLL, DIR>(q0, bg, K, eLb, eRb, srcb, q0b);
  sfor<5>([&](auto l) {
    double ub[ORD];
    sfor<ORD>([&](auto s) { ub[s] = 0.0; });
    reconstruct_adj<ORD, NGLL>(st[l], (l == idR || l == idT) ? weno_s : weno_w, q0b[l], ub);
    sfor<ORD>([&](auto s) { stb[l][s] += ub[s]; });
  });
}

#undef AWFL_HD

}  // namespace adj
}  // namespace awfl
