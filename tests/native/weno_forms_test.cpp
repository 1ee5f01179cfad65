// Host check of the rewritten WENO forms in awflcloud_amd/csrc/awfl_cell.hpp.
//
// 1. The identity behind AvgProjMap: every candidate polynomial and the
//    bridge have the centre cell's average over the centre cell,
//    a0 + sum_{c even >= 2} a_c / ((c + 1) 2^c) = u_hs, and the ideal weights
//    sum to one.
// 2. BridgeMirror reproduces the bridge map kBridge entry by entry.
// 3. reconstruct() (no constant coefficients, mirrored bridge, grouped TV
//    forms) against a statement-by-statement restatement of
//    WenoLimiter.h:93-172 and S:1813-1819 (all ord coefficients, the bridge as
//    (a_hi - sum idl_i a_i) / idl_hi, every term of the TV forms, the weights
//    with divisions) on random, smooth, stepped and near-constant stencils:
//    relative difference <= 1e-13 of the stencil's scale.
// Prints one line per check and PASSED at the end.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../../awflcloud_amd/csrc/awfl_cell.hpp"

using namespace awfl;

static int g_fail = 0;
static void report(const char* what, int ord, int ngll, double err, double tol) {
  const bool ok = err <= tol;
  if (!ok) g_fail++;
  std::printf("%-44s ord %d ngll %d  err %.3e  tol %.1e  %s\n", what, ord, ngll, err, tol, ok ? "ok" : "FAIL");
}

template <int ORD>
static void check_tables() {
  using W = Weno<ORD>;
  constexpr int hs = W::hs;
  auto kmom = [](int c) { return (c % 2) ? 0.0 : 1.0 / ((c + 1) * std::pow(2.0, c)); };
  double err = 0.0;
  for (int i = 0; i <= hs; i++)
    for (int s = 0; s <= hs; s++) {
      double avg = 0.0;
      for (int c = 0; c <= hs; c++) avg += W::recon[i][s][c] * kmom(c);
      err = std::fmax(err, std::fabs(avg - (i + s == hs ? 1.0 : 0.0)));
    }
  for (int s = 0; s < ORD; s++) {
    double avg = 0.0;
    for (int c = 0; c < ORD; c++) avg += kBridge<ORD>.m[s][c] * kmom(c);
    err = std::fmax(err, std::fabs(avg - (s == hs ? 1.0 : 0.0)));
  }
  double sum = 0.0;
  for (int i = 0; i < hs + 2; i++) sum += W::idl[i];
  err = std::fmax(err, std::fabs(sum - 1.0));
  report("centre-cell average of candidates and bridge", ORD, 0, err, 1e-14);
  // the mirrored bridge against the plain map
  double merr = 0.0, scale = 0.0;
  for (int c = 1; c < ORD; c++) {
    const auto& bm = kBridgeMirror<ORD>;
    if (bm.kind[c] == 0) {
      g_fail++;
      std::printf("bridge column %d of ord %d is not mirror-symmetric FAIL\n", c, ORD);
      continue;
    }
    for (int s = 0; s < ORD; s++) {
      const int k = s - hs;
      double v;
      if (k == 0) v = bm.kind[c] == 1 ? bm.cm[c] : 0.0;
      else if (k > 0) v = bm.cf[c][k];
      else v = bm.kind[c] == 1 ? bm.cf[c][-k] : -bm.cf[c][-k];
      merr = std::fmax(merr, std::fabs(v - kBridge<ORD>.m[s][c]));
      scale = std::fmax(scale, std::fabs(kBridge<ORD>.m[s][c]));
    }
  }
  report("mirrored bridge map vs plain map", ORD, 0, merr / scale, 1e-15);
}

// statement-by-statement restatement of the reference reconstruction
template <int ORD, int NGLL>
static void reconstruct_plain(const double (&u)[ORD], double (&g)[NGLL]) {
  using W = Weno<ORD>;
  using P = Proj<ORD, NGLL>;
  constexpr int hs = W::hs;
  constexpr int NP = hs + 2;
  double a[NP][ORD] = {};
  for (int i = 0; i <= hs; i++)
    for (int c = 0; c <= hs; c++)
      for (int s = 0; s <= hs; s++) a[i][c] += W::recon[i][s][c] * u[i + s];
  double hi[ORD] = {};
  for (int c = 0; c < ORD; c++)
    for (int s = 0; s < ORD; s++) hi[c] += W::recon[hs + 1][s][c] * u[s];
  for (int c = 0; c < ORD; c++) {
    double t = hi[c];
    for (int i = 0; i <= hs; i++) t -= W::idl[i] * a[i][c];
    a[hs + 1][c] = t / W::idl[hs + 1];
  }
  double tv[NP];
  for (int i = 0; i < NP; i++) {
    const int n = i <= hs ? Tv<hs + 1>::n : Tv<ORD>::n;
    double r = 0.0;
    for (int t = 0; t < n; t++) {
      const int j = i <= hs ? Tv<hs + 1>::j[t] : Tv<ORD>::j[t];
      const int k = i <= hs ? Tv<hs + 1>::k[t] : Tv<ORD>::k[t];
      const double c = i <= hs ? Tv<hs + 1>::c[t] : Tv<ORD>::c[t];
      r += c * a[i][j] * a[i][k];
    }
    tv[i] = r;
  }
  double lo = 0.0;
  for (int i = 0; i <= hs; i++) lo += tv[i];
  lo /= (hs + 1);
  tv[hs + 1] = lo + (tv[hs + 1] - lo) * W::sigma;
  double w[NP], sum = 0.0;
  for (int i = 0; i < NP; i++) {
    w[i] = W::idl[i] / (tv[i] * tv[i] + WENO_EPS);
    sum += w[i];
  }
  for (int i = 0; i < NP; i++) w[i] /= (sum + WENO_EPS);
  sum = 0.0;
  for (int i = 0; i < NP; i++) {
    const double d = W::idl[i];
    w[i] = w[i] * (d + d * d - 3.0 * d * w[i] + w[i] * w[i]) / (d * d + w[i] * (1.0 - 2.0 * d));
    sum += w[i];
  }
  for (int i = 0; i < NP; i++) w[i] /= (sum + WENO_EPS);
  double aw[ORD] = {};
  for (int c = 0; c < ORD; c++)
    for (int i = 0; i < NP; i++) aw[c] += w[i] * a[i][c];
  for (int ii = 0; ii < NGLL; ii++) {
    double t = 0.0;
    for (int c = 0; c < ORD; c++) t += P::c2g[c][ii] * aw[c];
    g[ii] = t;
  }
}

template <int ORD, int NGLL>
static void check_reconstruct() {
  std::mt19937_64 rng(12345 + ORD);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  double worst = 0.0;
  for (int trial = 0; trial < 20000; trial++) {
    double u[ORD], scale = 0.0;
    const int kind = trial % 4;
    const double amp = std::pow(10.0, -8.0 + 10.0 * (0.5 + 0.5 * U(rng)));
    const double off = (trial % 3 == 0) ? 300.0 * U(rng) : 0.0;
    for (int s = 0; s < ORD; s++) {
      const double x = s - (ORD - 1) / 2.0;
      double v;
      if (kind == 0) v = U(rng);                                     // random
      else if (kind == 1) v = std::sin(0.3 * x + U(rng)) + 0.1 * x;  // smooth
      else if (kind == 2) v = x > 0.5 * U(rng) ? 1.0 : -0.5;         // step
      else v = 1.0 + 1e-9 * U(rng);                                  // near constant
      u[s] = off + amp * v;
      scale = std::fmax(scale, std::fabs(u[s]));
    }
    double g[NGLL], r[NGLL];
    reconstruct<ORD, NGLL>(u, true, g);
    reconstruct_plain<ORD, NGLL>(u, r);
    for (int ii = 0; ii < NGLL; ii++) worst = std::fmax(worst, std::fabs(g[ii] - r[ii]) / scale);
  }
  report("reconstruct vs plain restatement", ORD, NGLL, worst, 1e-13);
}

int main() {
  check_tables<3>();
  check_tables<5>();
  check_tables<7>();
  check_tables<9>();
  check_reconstruct<3, 2>();
  check_reconstruct<5, 3>();
  check_reconstruct<9, 5>();
  if (g_fail) {
    std::printf("%d checks failed\n", g_fail);
    return 1;
  }
  std::printf("PASSED\n");
  return 0;
}
