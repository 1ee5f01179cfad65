// driver.cpp -- time loop over the MI355X ADER-WENO step, following the
// reference driver (driver.cpp:11-71): read simTime/outFreq from the YAML
// input, build Temporal_ader<Spatial>, loop computeTimeStep + timeStep with a
// fence-bracketed wall timer, write output every outFreq seconds.
//
//   awfl_driver input.yaml [--fused] [--max-steps N]
//
// --fused advances each step with one fused kernel per split (awfl_time_step);
// the default goes through computeTendencies + applyTendencies per split.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "awfl/spatial_hip.hpp"

using namespace awfl;

typedef Spatial_euler3d_cons_expl_cart_fv_Agrid<nTimeDerivs, timeAvg, nAder> Spatial;
typedef Temporal_ader<Spatial> Temporal;

int main(int argc, char** argv) {
  try {
    if (argc <= 1) endrun("ERROR: Must pass the input YAML filename as a parameter");
    std::string inFile(argv[1]);
    bool fused = false;
    long maxSteps = -1;
    for (int a = 2; a < argc; a++) {
      if (!std::strcmp(argv[a], "--fused")) fused = true;
      else if (!std::strcmp(argv[a], "--max-steps") && a + 1 < argc) maxSteps = std::atol(argv[++a]);
      else endrun(std::string("unknown argument ") + argv[a]);
    }
    FlatYaml config(inFile);
    real simTime = config.num("simTime");
    real outFreq = config.num("outFreq");
    int numOut = 0;

    Temporal model;
    model.spaceOp.addTracer("uniform", true, "constant value of 1");
    model.spaceOp.addTracer("theta", true, "replica of theta");
    model.spaceOp.addTracer("block", true, "block in domain center");
    model.init(inFile);

    Spatial::StateArr state = model.spaceOp.createStateArr();
    Spatial::TracerArr tracers = model.spaceOp.createTracerArr();
    model.spaceOp.initState(state);
    model.spaceOp.initTracers(tracers);

    real etime = 0;
    model.spaceOp.output(state, tracers, etime);
    std::chrono::duration<double, std::milli> timer(0);
    long steps = 0;
    while (etime < simTime && (maxSteps < 0 || steps < maxSteps)) {
      real dt = model.spaceOp.computeTimeStep(0.8, state);
      if (etime + dt > simTime) dt = simTime - etime;
      if (etime == 0) std::cout << "dt: " << dt << "\n\n";
      hip_check(hipDeviceSynchronize(), "fence");
      auto t1 = std::chrono::high_resolution_clock::now();
      if (fused) model.timeStepFused(state, dt);
      else model.timeStep(state, tracers, dt);
      hip_check(hipDeviceSynchronize(), "fence");
      auto t2 = std::chrono::high_resolution_clock::now();
      timer = timer + std::chrono::duration<double, std::milli>(t2 - t1);
      etime += dt;
      steps++;
      if (outFreq > 0 && etime / outFreq >= numOut + 1) {
        std::cout << "Etime , dt: " << etime << " , " << dt << "\n";
        model.spaceOp.output(state, tracers, etime);
        numOut++;
      }
    }
    std::cout << "Elapsed Time: " << etime << "\n";
    std::cout << "Steps: " << steps << "\n";
    std::cout << "Walltime: " << timer.count() / 1000 << "\n";
    std::cout << "cell-updates/s: "
              << (double)model.spaceOp.nx * model.spaceOp.ny * model.spaceOp.nz * steps / (timer.count() / 1000)
              << "\n";
    model.finalize(state, tracers);
  } catch (const std::exception& e) {
    return 1;
  }
  return 0;
}
