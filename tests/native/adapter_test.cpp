// Exercises the header-only Spatial/Temporal adapter through the C ABI:
// the same 3 steps advanced (a) with the canonical applySingleTendency on the
// device, (b) with an arbitrary host lambda through get(), (c) with the fused
// sweeps.  Prints the max relative differences; exit 0 when they agree.
#include <cmath>
#include <cstdio>

#include "awfl/spatial_hip.hpp"

using namespace awfl;
typedef Spatial_euler3d_cons_expl_cart_fv_Agrid<nTimeDerivs, timeAvg, nAder> Spatial;
typedef Temporal_ader<Spatial> Temporal;

static double maxrel(const std::vector<double>& a, const std::vector<double>& b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); i++) {
    num = std::fmax(num, std::fabs(a[i] - b[i]));
    den = std::fmax(den, std::fabs(b[i]));
  }
  return den > 0 ? num / den : num;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  try {
    Temporal A, B, C;
    A.init(argv[1]);
    B.init(argv[1]);
    C.init(argv[1]);
    Spatial::StateArr sa = A.spaceOp.createStateArr(), sb = B.spaceOp.createStateArr(),
                      sc = C.spaceOp.createStateArr();
    Spatial::TracerArr ta = A.spaceOp.createTracerArr(), tb = B.spaceOp.createTracerArr();
    A.spaceOp.initState(sa);
    B.spaceOp.initState(sb);
    C.spaceOp.initState(sc);
    real dt = A.spaceOp.computeTimeStep(0.8, sa);
    for (int s = 0; s < 3; s++) {
      A.timeStep(sa, ta, dt);
      for (int spl = 0; spl < B.spaceOp.numSplit(); spl++) {
        real dtloc = dt;
        Spatial& S = B.spaceOp;
        S.computeTendencies(sb, B.stateTendArr, tb, B.tracerTendArr, dtloc, spl);
        auto applySingleTendency = [&](Spatial::Location const& loc) {
          S.get(sb, loc, spl) += dtloc * S.get(B.stateTendArr, loc, 0, spl);
        };
        S.applyTendencies(applySingleTendency, spl);
      }
      C.timeStepFused(sc, dt);
    }
    std::vector<double> ha = sa.createHostCopy(), hb = sb.createHostCopy(), hc = sc.createHostCopy();
    double eab = maxrel(hb, ha), eac = maxrel(hc, ha);
    std::printf("adapter: device-lambda vs host-lambda %.3e, unfused vs fused %.3e\n", eab, eac);
    // the alternative integrators behind the same Temporal interface: three
    // steps each stay finite and differ from the ADER-DT steps
    Temporal_strang<Spatial> S;
    Temporal_ssprk3<Spatial> R;
    S.init(argv[1]);
    R.init(argv[1]);
    Spatial::StateArr ss = S.spaceOp.createStateArr(), sr = R.spaceOp.createStateArr();
    Spatial::TracerArr ts = S.spaceOp.createTracerArr(), tr = R.spaceOp.createTracerArr();
    S.spaceOp.initState(ss);
    R.spaceOp.initState(sr);
    for (int s = 0; s < 3; s++) {
      S.timeStep(ss, ts, dt);
      R.timeStep(sr, tr, 0.5 * dt);
    }
    std::vector<double> hs_ = ss.createHostCopy(), hr = sr.createHostCopy();
    bool finite = true;
    for (double v : hs_) finite = finite && std::isfinite(v);
    for (double v : hr) finite = finite && std::isfinite(v);
    double es = maxrel(hs_, sa.createHostCopy());
    std::printf("adapter: %s / %s finite %d, differ from ADER-DT by %.3e\n", S.getTemporalName(),
                R.getTemporalName(), (int)finite, es);
    return (eab <= 1e-13 && eac <= 1e-13 && finite && es > 1e-12) ? 0 : 1;
  } catch (const std::exception& e) {
    std::printf("adapter: exception %s\n", e.what());
    return 3;
  }
}
