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
    return (eab <= 1e-13 && eac <= 1e-13) ? 0 : 1;
  } catch (const std::exception& e) {
    std::printf("adapter: exception %s\n", e.what());
    return 3;
  }
}
