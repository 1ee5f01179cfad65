// spatial_hip.hpp -- header-only C++ adapter presenting the C ABI of
// awfl_dycore.h as the reference's Spatial / Temporal concept, so a driver
// written against the reference (driver.cpp, Temporal_ader.h,
// Spatial_euler3d_cons_expl_cart_fv_Agrid.h) compiles against it unchanged in
// its calls:
//
//   Spatial_euler3d_cons_expl_cart_fv_Agrid<nTimeDerivs,timeAvg,nAder>
//     hs, numState, Location, StateArr/TracerArr/StateTendArr/TracerTendArr,
//     addTracer, init(inFile), createStateArr, createTracerArr,
//     createStateTendArr, createTracerTendArr, numSplit, computeTimeStep,
//     initState, initTracers, computeTendencies, output, finalize      (S:14-1795)
//     applyTendencies(applySingleTendency, split), get(...)          (README.md:60-103)
//   Temporal_ader<Spatial>: init, timeStep(state, tracers, dt), finalize (Temporal_ader.h:26-77)
//
// Arrays are device buffers (RAII, hipMalloc).  Errors throw std::runtime_error
// after printing, as endrun() does (const.h:76-79).
//
// Tracers: the reference compiles its tracer transport out (#define
// NO_TRACERS, S:10) while its driver still adds three tracers
// (driver.cpp:27-29).  By default this adapter does the same with
// numTracers = 0; built with -DAWFL_TRACERS it transports the tracers added
// by addTracer (awfl_*_tracers of awfl_dycore.h), initialises them as
// initTracers does (S:505-560) and writes their mixing ratios to the output
// (S:1781-1787).
//
// Build a program against it with any C++17 compiler:
//   g++ -std=c++17 -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -Iinclude app.cpp \
//       -Lawflcloud_amd -lawfl_dycore -L/opt/rocm/lib -lamdhip64
#pragma once
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../awfl_dycore.h"
#include "netcdf_classic.hpp"

namespace awfl {

typedef double real;

// const.h:71-72 defaults (ORD 5, NGLL 3); override with -DORD / -DNGLL as the reference does
#ifndef ORD
#define ORD 5
#endif
#ifndef NGLL
#define NGLL 3
#endif
int constexpr ord = ORD;
int constexpr ngll = NGLL;

[[noreturn]] inline void endrun(const std::string& err) {
  std::cerr << err << std::endl;
  throw std::runtime_error(err);
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) endrun(std::string(what) + ": " + hipGetErrorString(e));
}

// A device array of doubles with a fixed 4-d shape (the yakl real4d the
// reference passes around, S:34-38).  Shallow copies share the buffer.
class Array4 {
 public:
  Array4() = default;
  Array4(const std::string& label, int64_t d0, int64_t d1, int64_t d2, int64_t d3) : label_(label) {
    dims_[0] = d0; dims_[1] = d1; dims_[2] = d2; dims_[3] = d3;
    double* p = nullptr;
    if (size() > 0) {
      hip_check(hipMalloc(&p, sizeof(double) * size()), "hipMalloc");
      hip_check(hipMemset(p, 0, sizeof(double) * size()), "hipMemset");
    }
    buf_ = std::shared_ptr<double>(p, [](double* q) { if (q) (void)hipFree(q); });
  }
  double* data() const { return buf_.get(); }
  int64_t size() const { return dims_[0] * dims_[1] * dims_[2] * dims_[3]; }
  int64_t extent(int i) const { return dims_[i]; }
  const std::string& label() const { return label_; }
  std::vector<double> createHostCopy() const {
    std::vector<double> h(size());
    if (size()) hip_check(hipMemcpy(h.data(), data(), sizeof(double) * size(), hipMemcpyDeviceToHost), "D2H");
    return h;
  }
  void copyFromHost(const std::vector<double>& h) {
    if ((int64_t)h.size() != size()) endrun("copyFromHost: size mismatch");
    if (size()) hip_check(hipMemcpy(data(), h.data(), sizeof(double) * size(), hipMemcpyHostToDevice), "H2D");
  }

 private:
  std::string label_;
  int64_t dims_[4] = {0, 0, 0, 0};
  std::shared_ptr<double> buf_;
};

// Flat "key : value" reader for the reference's YAML input files (the keys of
// S:197-272 and driver.cpp:17-22 are all scalars).
class FlatYaml {
 public:
  explicit FlatYaml(const std::string& path) {
    std::ifstream f(path);
    if (!f) endrun("ERROR: Invalid YAML input file");
    std::string line;
    while (std::getline(f, line)) {
      auto h = line.find('#');
      if (h != std::string::npos) line = line.substr(0, h);
      auto c = line.find(':');
      if (c == std::string::npos) continue;
      std::string k = trim(line.substr(0, c)), v = trim(line.substr(c + 1));
      if (!k.empty()) kv_[k] = v;
    }
  }
  bool has(const std::string& k) const { return kv_.count(k) > 0; }
  std::string str(const std::string& k) const {
    auto it = kv_.find(k);
    if (it == kv_.end()) endrun("ERROR: No " + k + " in input file");
    return it->second;
  }
  double num(const std::string& k) const { return std::stod(str(k)); }
  int integer(const std::string& k) const { return std::stoi(str(k)); }
  bool boolean(const std::string& k) const {
    std::string v = str(k);
    if (v == "true" || v == "True" || v == "1") return true;
    if (v == "false" || v == "False" || v == "0") return false;
    endrun("ERROR: " + k + " is not a boolean");
  }

 private:
  static std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\""), b = s.find_last_not_of(" \t\r\"");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
  }
  std::map<std::string, std::string> kv_;
};

// The canonical applySingleTendency of Temporal_ader: state += dt * tend
// (Temporal_ader.h:63-65).  Passing it to applyTendencies runs on the device.
struct ApplySingleTendency {
  real dt;
};

template <int nTimeDerivs, bool timeAvg, int nAder>
class Spatial_euler3d_cons_expl_cart_fv_Agrid {
 public:
  static_assert(nTimeDerivs == 1, "ERROR: This Spatial class isn't setup to use nTimeDerivs > 1");
  static_assert(nAder == ngll, "nAder must equal ngll (Temporal_ader.h:8)");
  static int constexpr hs = (ord - 1) / 2;
  static int constexpr numState = 5;
  static int constexpr idR = 0, idU = 1, idV = 2, idW = 3, idT = 4;
  static int constexpr BC_PERIODIC = 0, BC_WALL = 1;
  static int constexpr DATA_SPEC_THERMAL = 1, DATA_SPEC_COLLISION = 2, DATA_SPEC_STRAKA = 3, DATA_SPEC_IGW = 4;

  struct Location {
    int l, k, j, i;
  };
  typedef Array4 StateArr;
  typedef Array4 TracerArr;
  typedef Array4 StateTendArr;
  typedef Array4 TracerTendArr;

  int nx = 0, ny = 0, nz = 0;
  real xlen = 0, ylen = 0, zlen = 0, dx = 0, dy = 0, dz = 0;
  int bc_x = 0, bc_y = 0, bc_z = 0, dataSpec = 0;
  bool weno_scalars = true, weno_winds = true, sim2d = false, perturb = false;
  std::string outFile;
  int numTracers = 0;   // 0 unless built with AWFL_TRACERS (the reference's NO_TRACERS, S:10)
  std::vector<std::string> tracerName, tracerDesc;
  std::vector<int> tracerPos;

  Spatial_euler3d_cons_expl_cart_fv_Agrid() = default;
  Spatial_euler3d_cons_expl_cart_fv_Agrid(const Spatial_euler3d_cons_expl_cart_fv_Agrid&) = delete;
  ~Spatial_euler3d_cons_expl_cart_fv_Agrid() {
    if (ctx_) awfl_destroy(ctx_);
  }

  // S:121-125.  Without AWFL_TRACERS the names are kept and the arrays are
  // created with numTracers = 0.
  void addTracer(std::string name, bool posDef = true, std::string desc = "") {
    tracerName.push_back(name);
    tracerDesc.push_back(desc);
    tracerPos.push_back(posDef ? 1 : 0);
  }

  // S:192-340
  void init(std::string inFile) {
    FlatYaml c(inFile);
    nx = c.integer("nx"); ny = c.integer("ny"); nz = c.integer("nz");
    xlen = c.num("xlen"); ylen = c.num("ylen"); zlen = c.num("zlen");
    bc_x = bc(c.str("bc_x"), "bc_x"); bc_y = bc(c.str("bc_y"), "bc_y"); bc_z = bc(c.str("bc_z"), "bc_z");
    weno_scalars = c.boolean("weno_scalars");
    weno_winds = c.boolean("weno_winds");
    std::string d = c.str("initData");
    if (d == "thermal") dataSpec = DATA_SPEC_THERMAL;
    else if (d == "collision") dataSpec = DATA_SPEC_COLLISION;
    else if (d == "straka") dataSpec = DATA_SPEC_STRAKA;
    else if (d == "igw") dataSpec = DATA_SPEC_IGW;
    else endrun("ERROR: Invalid dataSpec");
    outFile = c.str("outFile");
    perturb = c.boolean("perturb");
    sim2d = ny == 1;
    dx = xlen / nx; dy = ylen / ny; dz = zlen / nz;
    awfl_config cfg{};
    cfg.nx = nx; cfg.ny = ny; cfg.nz = nz;
    cfg.xlen = xlen; cfg.ylen = ylen; cfg.zlen = zlen;
    cfg.bc_x = bc_x; cfg.bc_y = bc_y; cfg.bc_z = bc_z;
    cfg.weno_scalars = weno_scalars; cfg.weno_winds = weno_winds;
    cfg.ord = ord; cfg.ngll = ngll; cfg.data_spec = dataSpec; cfg.precision = AWFL_FP64;
    if (awfl_create(&cfg, &ctx_) != AWFL_OK) endrun(std::string("awfl_create: ") + awfl_last_error(nullptr));
#ifdef AWFL_TRACERS
    numTracers = (int)tracerName.size();   // S:318
    if (numTracers > 0) check(awfl_set_tracers(ctx_, numTracers, tracerPos.data()));
#endif
  }

  StateArr createStateArr() const { return StateArr("stateArr", numState, nz + 2 * hs, ny + 2 * hs, nx + 2 * hs); }
  TracerArr createTracerArr() const { return TracerArr("tracerArr", numTracers, nz + 2 * hs, ny + 2 * hs, nx + 2 * hs); }
  StateTendArr createStateTendArr() const { return StateTendArr("stateTendArr", numState, nz, ny, nx); }
  TracerTendArr createTracerTendArr() const { return TracerTendArr("tracerTendArr", numTracers, nz, ny, nx); }
  int numSplit() const { return awfl_num_split(ctx_); }

  // S:159-187 (computed once, then cached)
  real computeTimeStep(real cfl, StateArr const& state) {
    double dt = 0;
    check(awfl_compute_time_step(ctx_, state.data(), cfl, &dt));
    return dt;
  }

  // S:345-500: the four data specs of the reference
  void initState(StateArr& state) {
    awfl_init_spec s{};
    if (dataSpec == DATA_SPEC_THERMAL) {
      add_ellipsoid(s, 0, xlen / 2, ylen / 2, 2000, 2000, 2000, 2000, 2, 1);
    } else if (dataSpec == DATA_SPEC_COLLISION) {
      add_ellipsoid(s, 0, xlen / 2, ylen / 2, 2000, 2000, 2000, 2000, 20, 1);
      add_ellipsoid(s, 0, xlen / 2, ylen / 2, 8000, 2000, 2000, 2000, -20, 1);
    } else if (dataSpec == DATA_SPEC_STRAKA) {
      add_ellipsoid(s, 1, xlen / 2, ylen / 2, 3000, 4000, 2000, 2000, -15, 1);
    } else {
      add_ellipsoid(s, 2, xlen / 3, ylen / 3, 0, 5000, 5000, 1, 0.01, 1);
      s.u0 = 20;
      s.v0 = 20;
      s.y2d = ylen / 3;
    }
    if (perturb) {           // S:488-499 multiplies by 1 + rand()/RAND_MAX * 1e-7
      s.perturb_amp = 1e-7;
      s.perturb_seed = 1;
    }
    check(awfl_init_state(ctx_, state.data(), &s));
  }
  // S:505-560
  void initTracers(TracerArr& tracers) {
    if (numTracers > 0) check(awfl_init_tracers(ctx_, tracers.data()));
  }

  // S:565-594 (with the tracer blocks when tracers are transported)
  void computeTendencies(StateArr& state, StateTendArr& tend, TracerArr& tracers, TracerTendArr& ttend, real& dt,
                         int split) {
    if (numTracers > 0)
      check(awfl_compute_tendencies_tracers(ctx_, state.data(), tracers.data(), tend.data(), ttend.data(), dt, split));
    else
      check(awfl_compute_tendencies(ctx_, state.data(), tend.data(), dt, split));
    bound_state_ = &state;
    bound_tend_ = &tend;
    bound_tracers_ = &tracers;
    bound_ttend_ = &ttend;
  }

  // tracers += dt * tracerTend of the last computeTendencies call
  // (Temporal_ader.h:67-69)
  void applyTracerTendencies(real dt) {
    if (numTracers == 0) return;
    if (!bound_tracers_ || !bound_ttend_) endrun("applyTracerTendencies called before computeTendencies");
    check(awfl_apply_tracer_tendencies(ctx_, bound_tracers_->data(), bound_ttend_->data(), dt));
  }

  // README.md:60-103 contract: apply the tendency of the last computeTendencies
  // call through applySingleTendency at every Location of the split.  The
  // canonical lambda runs on the device; any other callable(Location const&)
  // runs over host mirrors obtained through get().
  template <class F>
  void applyTendencies(F const& applySingleTendency, int split) {
    if (!bound_state_ || !bound_tend_) endrun("applyTendencies called before computeTendencies");
    if constexpr (std::is_same<F, ApplySingleTendency>::value) {
      check(awfl_apply_tendencies(ctx_, bound_state_->data(), bound_tend_->data(), applySingleTendency.dt, split));
    } else {
      host_state_ = bound_state_->createHostCopy();
      host_tend_ = bound_tend_->createHostCopy();
      for (int l = 0; l < numState; l++)
        for (int k = 0; k < nz; k++)
          for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) applySingleTendency(Location{l, k, j, i});
      bound_state_->copyFromHost(host_state_);
      host_state_.clear();
      host_tend_.clear();
    }
  }
  // host views used by a host-side applySingleTendency (README.md:69-84)
  real& get(StateArr const&, Location const& loc, int) {
    return host_state_[(((size_t)loc.l * (nz + 2 * hs) + hs + loc.k) * (ny + 2 * hs) + hs + loc.j) * (nx + 2 * hs) +
                       hs + loc.i];
  }
  real& get(StateTendArr const&, Location const& loc, int, int) {
    return host_tend_[(((size_t)loc.l * nz + loc.k) * ny + loc.j) * nx + loc.i];
  }

  // S:1704-1791: appends the diagnostics at elapsed time etime to the netCDF
  // file outFile (created when etime == 0): coordinates x, y, z, background
  // hyDens/hyPressure/hyTheta, the unlimited t, and dens_pert, u, v, w,
  // pot_temp_pert, pressure_pert on (t, z, y, x), computed on the device
  // (awfl_diagnostics), then tracer_<name> = tracer / rho per tracer
  // (S:1781-1787); classic 64-bit-offset format (netcdf_classic.hpp)
  void output(StateArr const& state, TracerArr const& tracers, real etime) {
    static const char* names[6] = {"dens_pert", "u", "v", "w", "pot_temp_pert", "pressure_pert"};
    const size_t n = (size_t)nx * ny * nz;
    double* d_out = nullptr;
    if (hipMalloc(&d_out, 6 * n * sizeof(double)) != hipSuccess) endrun("output: device allocation failed");
    int rc = awfl_diagnostics(ctx_, state.data(), d_out);
    std::vector<double> h(6 * n);
    hipError_t e = rc == AWFL_OK ? hipMemcpy(h.data(), d_out, 6 * n * sizeof(double), hipMemcpyDeviceToHost)
                                 : hipSuccess;
    (void)hipFree(d_out);
    check(rc);
    if (e != hipSuccess) endrun("output: copy to host failed");
    std::vector<double> hq;            // mixing ratios, [numTracers][nz][ny][nx]
    if (numTracers > 0) {
      std::vector<double> ht = tracers.createHostCopy(), hd(nz + 2 * hs);
      check(awfl_get_background(ctx_, 0, hd.data(), (int64_t)hd.size()));
      hq.resize((size_t)numTracers * n);
      for (int t = 0; t < numTracers; t++)
        for (int k = 0; k < nz; k++)
          for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
              const size_t c = ((size_t)k * ny + j) * nx + i;
              const double r = h[c] + hd[hs + k];
              hq[t * n + c] = ht[(((size_t)t * (nz + 2 * hs) + hs + k) * (ny + 2 * hs) + hs + j) * (nx + 2 * hs) + hs + i] / r;
            }
    }
    nc::ClassicFile f;
    int64_t rec = 0;
    if (etime == 0) {
      std::vector<nc::Dim> dims = {{"x", nx}, {"y", ny}, {"z", nz}, {"t", 0}};
      std::vector<nc::Var> vars = {{"x", {0}}, {"y", {1}}, {"z", {2}}, {"hyDens", {2}}, {"hyPressure", {2}},
                                   {"hyTheta", {2}}, {"t", {3}}};
      for (const char* v : names) vars.push_back({v, {3, 2, 1, 0}});
      for (int t = 0; t < numTracers; t++) vars.push_back({"tracer_" + tracerName[t], {3, 2, 1, 0}});
      f.create(outFile, dims, vars);
      std::vector<double> c(std::max(nx, std::max(ny, nz)));
      for (int i = 0; i < nx; i++) c[i] = (i + 0.5) * dx;
      f.write_fixed("x", c.data());
      for (int j = 0; j < ny; j++) c[j] = (j + 0.5) * dy;
      f.write_fixed("y", c.data());
      for (int k = 0; k < nz; k++) c[k] = (k + 0.5) * dz;
      f.write_fixed("z", c.data());
      const int which[3] = {0, 1, 2};
      const char* bnames[3] = {"hyDens", "hyPressure", "hyTheta"};
      std::vector<double> b(nz + 2 * hs);
      for (int w = 0; w < 3; w++) {
        check(awfl_get_background(ctx_, which[w], b.data(), (int64_t)b.size()));
        f.write_fixed(bnames[w], b.data() + hs);
      }
    } else {
      f.open(outFile);
      rec = f.num_records();
    }
    double t = etime;
    f.write_record("t", rec, &t);
    for (int v = 0; v < 6; v++) f.write_record(names[v], rec, h.data() + v * n);
    for (int t = 0; t < numTracers; t++) f.write_record("tracer_" + tracerName[t], rec, hq.data() + t * n);
  }

  void finalize(StateArr const&, TracerArr const&) {}
  const char* getName() { return "ADER-WENO Euler 3-D, fused MI355X sweeps"; }

  awfl_ctx* context() const { return ctx_; }

 private:
  static int bc(const std::string& v, const char* key) {
    if (v == "periodic") return BC_PERIODIC;
    if (v == "wall") return BC_WALL;
    endrun(std::string("Invalid ") + key);
  }
  static void add_ellipsoid(awfl_init_spec& s, int shape, double x0, double y0, double z0, double xr, double yr,
                            double zr, double amp, double pwr) {
    int b = s.nbub++;
    s.shape[b] = shape; s.x0[b] = x0; s.y0[b] = y0; s.z0[b] = z0;
    s.xr[b] = xr; s.yr[b] = yr; s.zr[b] = zr; s.amp[b] = amp; s.pwr[b] = pwr;
  }
  void check(int rc) const {
    if (rc < 0) endrun(std::string("awfl: ") + awfl_last_error(ctx_));
  }
  awfl_ctx* ctx_ = nullptr;
  StateArr* bound_state_ = nullptr;
  StateTendArr* bound_tend_ = nullptr;
  TracerArr* bound_tracers_ = nullptr;
  TracerTendArr* bound_ttend_ = nullptr;
  std::vector<double> host_state_, host_tend_;
};

// Temporal_ader.h:6-8, 26-77 with the split update routed through
// applyTendencies (the README's applySingleTendency contract).
int constexpr nTimeDerivs = 1;
bool constexpr timeAvg = true;
int constexpr nAder = ngll;

template <class Spatial>
class Temporal_ader {
 public:
  typedef typename Spatial::StateTendArr StateTendArr;
  typedef typename Spatial::StateArr StateArr;
  typedef typename Spatial::TracerTendArr TracerTendArr;
  typedef typename Spatial::TracerArr TracerArr;
  typedef typename Spatial::Location Location;

  StateTendArr stateTendArr;
  TracerTendArr tracerTendArr;
  Spatial spaceOp;

  void init(std::string inFile) {
    spaceOp.init(inFile);
    stateTendArr = spaceOp.createStateTendArr();
    tracerTendArr = spaceOp.createTracerTendArr();
  }

  void timeStep(StateArr& stateArr, TracerArr& tracerArr, real dt) {
    for (int spl = 0; spl < spaceOp.numSplit(); spl++) {
      real dtloc = dt;
      spaceOp.computeTendencies(stateArr, stateTendArr, tracerArr, tracerTendArr, dtloc, spl);
      spaceOp.applyTendencies(ApplySingleTendency{dtloc}, spl);
      spaceOp.applyTracerTendencies(dtloc);
    }
  }

  // all three splits as fused sweeps (one kernel per split, in place)
  void timeStepFused(StateArr& stateArr, real dt) {
    if (spaceOp.numTracers > 0) endrun("timeStepFused: tracers are transported; pass them");
    if (awfl_time_step(spaceOp.context(), stateArr.data(), dt) < 0)
      endrun(std::string("awfl_time_step: ") + awfl_last_error(spaceOp.context()));
  }
  void timeStepFused(StateArr& stateArr, TracerArr& tracerArr, real dt) {
    const int rc = spaceOp.numTracers > 0
                       ? awfl_time_step_tracers(spaceOp.context(), stateArr.data(), tracerArr.data(), dt)
                       : awfl_time_step(spaceOp.context(), stateArr.data(), dt);
    if (rc < 0) endrun(std::string("time step: ") + awfl_last_error(spaceOp.context()));
  }

  void finalize(StateArr& state, TracerArr& tracers) { spaceOp.finalize(state, tracers); }
  const char* getTemporalName() { return "ADER-DT"; }
};

// The alternative integrators of qjrmsCode/TimeIntegrator.cpp over the same
// Spatial operator, with the Temporal interface of Temporal_ader:
//   Temporal_strang: x, y sweeps with dt and z as two dt/2 sweeps (:32-65)
//   Temporal_ssprk3: SSPRK3 on the semi-discrete operator (:68-110)
template <class Spatial, int (*STEP)(awfl_ctx*, double*, double)>
class Temporal_alt {
 public:
  typedef typename Spatial::StateArr StateArr;
  typedef typename Spatial::TracerArr TracerArr;
  typedef typename Spatial::Location Location;

  Spatial spaceOp;

  void init(std::string inFile) { spaceOp.init(inFile); }
  void timeStep(StateArr& stateArr, TracerArr&, real dt) {
    if (spaceOp.numTracers > 0) endrun("this integrator does not transport tracers (Temporal_ader does)");
    if (STEP(spaceOp.context(), stateArr.data(), dt) < 0)
      endrun(std::string("time step: ") + awfl_last_error(spaceOp.context()));
  }
  void finalize(StateArr& state, TracerArr& tracers) { spaceOp.finalize(state, tracers); }
};

template <class Spatial>
class Temporal_strang : public Temporal_alt<Spatial, awfl_time_step_strang> {
 public:
  const char* getTemporalName() { return "ADER-DT, z as two half steps"; }
};

template <class Spatial>
class Temporal_ssprk3 : public Temporal_alt<Spatial, awfl_time_step_ssprk3> {
 public:
  const char* getTemporalName() { return "SSPRK3, semi-discrete"; }
};

}  // namespace awfl
