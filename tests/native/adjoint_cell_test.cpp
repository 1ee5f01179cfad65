// Host test of the per-cell / per-face tangent-linear and adjoint numerics
// (awfl_adjoint.hpp), compiled with g++ (no GPU).
//
// For random cell stencils and face states near the hydrostatic background,
// for every sweep direction and (ord, ngll) variant:
//   dot   : <ybar, J xdot> (tangent, Dual numbers) == <J^T ybar, xdot>
//           (hand-written reverse) to ~1e-13 relative;
//   fd    : the tangent matches fourth-order central differences of the
//           production forward (cell_edges / riemann of awfl_cell.hpp) to 1e-8;
//   primal: cell_fwd<double> reproduces cell_edges to 1e-13.
// Prints one line per check and exits nonzero if any fails.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../awflcloud_amd/csrc/awfl_adjoint.hpp"

using namespace awfl;
using adj::Dual;

static int g_fail = 0;

static void check(const char* what, int ord, int ngll, int dir, const char* tag, double err, double tol) {
  bool ok = err <= tol && std::isfinite(err);
  if (!ok) g_fail++;
  std::printf("%-6s ord %d ngll %d dir %d %-14s rel err %.3e (tol %.0e) %s\n", what, ord, ngll, dir, tag, err, tol,
              ok ? "ok" : "FAIL");
}

static PowRef power_ref(double B) {
  long double v = powl((long double)B, (long double)GAMMA);
  PowRef r;
  r.h = (double)v;
  r.l = (double)(v - (long double)r.h);
  r.rb = 1.0 / B;
  r.b = B;
  return r;
}

struct Rng {
  std::mt19937_64 g;
  explicit Rng(unsigned long s) : g(s) {}
  double uni(double a, double b) { return std::uniform_real_distribution<double>(a, b)(g); }
};

static SweepK sweep_constants(double dt, double dsw) {
  SweepK K{};
  K.dt = dt;
  K.rd = 1.0 / dsw;
  double dtmult = dt;
  for (int kt = 0; kt < KMAX; kt++) {
    K.ndt[kt] = -1.0 / (dsw * (kt + 1.0));
    if (kt >= 1) {
      K.tav[kt] = dtmult / (kt + 1.0);
      dtmult *= dt;
    }
  }
  return K;
}

template <int NGLL>
static CellBg<NGLL> make_bg(Rng& R, bool wall_bot, bool wall_top) {
  CellBg<NGLL> bg{};
  double rt0 = 330.0 + 20.0 * R.uni(-1, 1);
  bg.rt = rt0;
  bg.r = rt0 / 300.0;
  bg.pt = power_ref(rt0);
  for (int ii = 0; ii < NGLL; ii++) {
    double rt = rt0 * (1.0 + 0.004 * (0.5 - (double)ii / (NGLL - 1)));
    bg.rtG[ii] = rt;
    bg.rG[ii] = rt / 300.0;
    bg.pG[ii] = C0 * std::pow(rt, GAMMA);
    bg.ptG[ii] = power_ref(rt);
  }
  bg.wall_bot = wall_bot;
  bg.wall_top = wall_top;
  return bg;
}

// deviations: rho' ~ 1e-3, momenta ~ a few kg m^-2 s^-1, (rho theta)' ~ 0.5
template <int ORD>
static void make_stencil(Rng& R, double (&st)[5][ORD]) {
  const double amp[5] = {2e-3, 3.0, 2.0, 2.5, 0.8};
  const double mean[5] = {0.0, 4.0, -1.0, 0.5, 0.0};
  for (int l = 0; l < 5; l++)
    for (int s = 0; s < ORD; s++) st[l][s] = mean[l] + amp[l] * R.uni(-1, 1);
}

template <int ORD, int NGLL, int DIR>
static void test_cell(Rng& R, bool weno, bool wall_bot, bool wall_top, const char* tag) {
  for (int trial = 0; trial < 6; trial++) {
    CellBg<NGLL> bg = make_bg<NGLL>(R, wall_bot, wall_top);
    double dsw = DIR == 2 ? 78.125 : 312.5;
    SweepK K = sweep_constants(0.18, dsw);
    double st[5][ORD], xd[5][ORD];
    make_stencil<ORD>(R, st);
    for (int l = 0; l < 5; l++)
      for (int s = 0; s < ORD; s++) xd[l][s] = R.uni(-1, 1) * (l == 0 ? 1e-3 : l == 4 ? 0.5 : 1.0);
    double eLb[5], eRb[5], srcb = DIR == 2 ? R.uni(-1, 1) : 0.0;
    for (int l = 0; l < 5; l++) {
      eLb[l] = R.uni(-1, 1);
      eRb[l] = R.uni(-1, 1);
    }
    // tangent
    Dual sd[5][ORD];
    for (int l = 0; l < 5; l++)
      for (int s = 0; s < ORD; s++) sd[l][s] = Dual(st[l][s], xd[l][s]);
    adj::Tape<NGLL, Dual> tpd;
    Dual eLd[5], eRd[5], sad;
    adj::cell_fwd<ORD, NGLL, DIR>(sd, bg, weno, weno, K, tpd, eLd, eRd, sad);
    double lhs = srcb * sad.d, lscale = std::fabs(srcb * sad.d);
    for (int l = 0; l < 5; l++) {
      lhs += eLb[l] * eLd[l].d + eRb[l] * eRd[l].d;
      lscale += std::fabs(eLb[l] * eLd[l].d) + std::fabs(eRb[l] * eRd[l].d);
    }
    // reverse
    double stb[5][ORD] = {};
    adj::cell_adj<ORD, NGLL, DIR>(st, bg, weno, weno, K, eLb, eRb, srcb, stb);
    double rhs = 0.0;
    for (int l = 0; l < 5; l++)
      for (int s = 0; s < ORD; s++) rhs += stb[l][s] * xd[l][s];
    check("dot", ORD, NGLL, DIR, tag, std::fabs(lhs - rhs) / lscale, 1e-12);
    // production forward vs generic primal, and central differences
    double eL[5], eR[5], sa;
    cell_edges<ORD, NGLL, DIR>(st, bg, weno, weno, K, eL, eR, sa);
    double perr = 0.0, pscale = 0.0;
    for (int l = 0; l < 5; l++) {
      perr = std::fmax(perr, std::fmax(std::fabs(eL[l] - eLd[l].v), std::fabs(eR[l] - eRd[l].v)));
      pscale = std::fmax(pscale, std::fmax(std::fabs(eL[l]), std::fabs(eR[l])));
    }
    check("primal", ORD, NGLL, DIR, tag, perr / pscale, 1e-13);
    // fourth-order central differences: (8 (f(h) - f(-h)) - (f(2h) - f(-2h))) / (12 h)
    const double h = 1e-4;
    double eLs[4][5], eRs[4][5], sas[4];
    const double mult[4] = {1.0, -1.0, 2.0, -2.0};
    for (int c = 0; c < 4; c++) {
      double sp[5][ORD];
      for (int l = 0; l < 5; l++)
        for (int s = 0; s < ORD; s++) sp[l][s] = st[l][s] + mult[c] * h * xd[l][s];
      cell_edges<ORD, NGLL, DIR>(sp, bg, weno, weno, K, eLs[c], eRs[c], sas[c]);
    }
    auto d4 = [&](double p1, double m1, double p2, double m2) { return (8.0 * (p1 - m1) - (p2 - m2)) / (12.0 * h); };
    double ferr = 0.0, fscale = 0.0;
    for (int l = 0; l < 5; l++) {
      double a = d4(eLs[0][l], eLs[1][l], eLs[2][l], eLs[3][l]);
      double b = d4(eRs[0][l], eRs[1][l], eRs[2][l], eRs[3][l]);
      ferr = std::fmax(ferr, std::fmax(std::fabs(a - eLd[l].d), std::fabs(b - eRd[l].d)));
      fscale = std::fmax(fscale, std::fmax(std::fabs(eLd[l].d), std::fabs(eRd[l].d)));
    }
    if (DIR == 2) {
      ferr = std::fmax(ferr, std::fabs(d4(sas[0], sas[1], sas[2], sas[3]) - sad.d));
      fscale = std::fmax(fscale, std::fabs(sad.d));
    }
    check("fd", ORD, NGLL, DIR, tag, ferr / fscale, 1e-8);
  }
}

template <int DIR>
static void test_riemann(Rng& R) {
  for (int trial = 0; trial < 8; trial++) {
    double rt0 = 330.0 + 20.0 * R.uni(-1, 1), r0 = rt0 / 300.0;
    double L[5], Rs[5], Ld[5], Rd[5];
    for (int side = 0; side < 2; side++) {
      double* q = side ? Rs : L;
      q[0] = r0 + 2e-3 * R.uni(-1, 1);
      q[1] = r0 * (3.0 * R.uni(-1, 1) + (trial % 2 ? 2.0 : -2.0));   // both upwind branches
      q[2] = r0 * 2.0 * R.uni(-1, 1);
      q[3] = r0 * (2.0 * R.uni(-1, 1) + (trial % 2 ? 1.5 : -1.5));
      q[4] = rt0 + 0.8 * R.uni(-1, 1);
    }
    for (int l = 0; l < 5; l++) {
      Ld[l] = R.uni(-1, 1) * (l == 0 ? 1e-3 : 1.0);
      Rd[l] = R.uni(-1, 1) * (l == 0 ? 1e-3 : 1.0);
    }
    double pface = C0 * std::pow(rt0, GAMMA);
    FaceRef fr;
    fr.rt = power_ref(DIR == 1 ? r0 * rt0 : rt0);
    fr.q5 = power_ref(rt0);
    Dual LD[5], RD[5], FD[5];
    for (int l = 0; l < 5; l++) {
      LD[l] = Dual(L[l], Ld[l]);
      RD[l] = Dual(Rs[l], Rd[l]);
    }
    adj::riemann_fwd<DIR>(LD, RD, fr, pface, FD);
    double Fb[5], Lb[5] = {}, Rb[5] = {};
    for (int l = 0; l < 5; l++) Fb[l] = R.uni(-1, 1);
    adj::riemann_adj<DIR>(L, Rs, fr, Fb, Lb, Rb);
    double lhs = 0, lscale = 0, rhs = 0;
    for (int l = 0; l < 5; l++) {
      lhs += Fb[l] * FD[l].d;
      lscale += std::fabs(Fb[l] * FD[l].d);
      rhs += Lb[l] * Ld[l] + Rb[l] * Rd[l];
    }
    check("dot", 0, 0, DIR, "riemann", std::fabs(lhs - rhs) / lscale, 1e-12);
    // production flux (with the same power references) vs the generic primal
    // and central differences
    double F[5];
    riemann<DIR>(L, Rs, fr, F);
    if (DIR == 2) F[3] -= pface;
    double perr = 0, pscale = 0;
    for (int l = 0; l < 5; l++) {
      perr = std::fmax(perr, std::fabs(F[l] - FD[l].v));
      pscale = std::fmax(pscale, std::fabs(F[l]));
    }
    check("primal", 0, 0, DIR, "riemann", perr / pscale, 1e-13);
    const double h = 1e-4;
    const double mult[4] = {1.0, -1.0, 2.0, -2.0};
    double Fs[4][5];
    for (int c = 0; c < 4; c++) {
      double Lp[5], Rp[5];
      for (int l = 0; l < 5; l++) {
        Lp[l] = L[l] + mult[c] * h * Ld[l];
        Rp[l] = Rs[l] + mult[c] * h * Rd[l];
      }
      riemann<DIR>(Lp, Rp, fr, Fs[c]);
    }
    double ferr = 0, fscale = 0;
    for (int l = 0; l < 5; l++) {
      double a = (8.0 * (Fs[0][l] - Fs[1][l]) - (Fs[2][l] - Fs[3][l])) / (12.0 * h);
      ferr = std::fmax(ferr, std::fabs(a - FD[l].d));
      fscale = std::fmax(fscale, std::fabs(FD[l].d));
    }
    check("fd", 0, 0, DIR, "riemann", ferr / fscale, 1e-8);
  }
}

template <int ORD, int NGLL>
static void test_variant(Rng& R) {
  test_cell<ORD, NGLL, 0>(R, true, false, false, "weno");
  test_cell<ORD, NGLL, 1>(R, true, false, false, "weno");
  test_cell<ORD, NGLL, 2>(R, true, false, false, "weno");
  test_cell<ORD, NGLL, 2>(R, true, true, false, "weno wall-bot");
  test_cell<ORD, NGLL, 2>(R, true, false, true, "weno wall-top");
  test_cell<ORD, NGLL, 0>(R, false, false, false, "linear");
  test_cell<ORD, NGLL, 2>(R, false, false, false, "linear");
}

int main() {
  Rng R(20261202);
  test_variant<3, 2>(R);
  test_variant<5, 3>(R);
  test_variant<9, 5>(R);
  test_riemann<0>(R);
  test_riemann<1>(R);
  test_riemann<2>(R);
  std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "PASSED", g_fail);
  return g_fail ? 1 : 0;
}
