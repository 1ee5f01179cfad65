// The adapter built with AWFL_TRACERS: the driver's three tracers
// (driver.cpp:27-29) initialised by initTracers (S:505-560), three steps
// through Temporal_ader::timeStep (computeTendencies, applyTendencies and the
// tracer update of Temporal_ader.h:67-69) against the fused step with
// tracers, and the output with the tracer_<name> mixing ratios
// (S:1781-1787).  Exit 0 when the two paths agree and the uniform tracer's
// mixing ratio is still 1.
#include <cmath>
#include <cstdio>

#include "awfl/spatial_hip.hpp"

using namespace awfl;
typedef Spatial_euler3d_cons_expl_cart_fv_Agrid<nTimeDerivs, timeAvg, nAder> Spatial;
typedef Temporal_ader<Spatial> Temporal;

static double maxdiff(const std::vector<double>& a, const std::vector<double>& b, double& scale) {
  double num = 0;
  scale = 0;
  for (size_t i = 0; i < a.size(); i++) {
    num = std::fmax(num, std::fabs(a[i] - b[i]));
    scale = std::fmax(scale, std::fabs(b[i]));
  }
  return num;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  try {
    Temporal A, B;
    for (Temporal* m : {&A, &B}) {
      m->spaceOp.addTracer("uniform", true, "constant value of 1");
      m->spaceOp.addTracer("theta", true, "replica of theta");
      m->spaceOp.addTracer("block", true, "block in domain center");
      m->init(argv[1]);
    }
    Spatial::StateArr sa = A.spaceOp.createStateArr(), sb = B.spaceOp.createStateArr();
    Spatial::TracerArr ta = A.spaceOp.createTracerArr(), tb = B.spaceOp.createTracerArr();
    A.spaceOp.initState(sa);
    B.spaceOp.initState(sb);
    A.spaceOp.initTracers(ta);
    B.spaceOp.initTracers(tb);
    real etime = 0;
    A.spaceOp.output(sa, ta, etime);
    real dt = A.spaceOp.computeTimeStep(0.8, sa);
    B.spaceOp.computeTimeStep(0.8, sb);
    for (int s = 0; s < 3; s++) {
      A.timeStep(sa, ta, dt);
      B.timeStepFused(sb, tb, dt);
      etime += dt;
    }
    A.spaceOp.output(sa, ta, etime);
    double ss, st;
    double es = maxdiff(sa.createHostCopy(), sb.createHostCopy(), ss);
    double et = maxdiff(ta.createHostCopy(), tb.createHostCopy(), st);
    std::printf("adapter tracers: numTracers %d, unfused vs fused state %.3e tracers %.3e\n", A.spaceOp.numTracers,
                es / ss, et / st);
    return (A.spaceOp.numTracers == 3 && es <= 1e-13 * ss && et <= 1e-13 * st) ? 0 : 1;
  } catch (const std::exception& e) {
    std::printf("adapter tracers: exception %s\n", e.what());
    return 3;
  }
}
