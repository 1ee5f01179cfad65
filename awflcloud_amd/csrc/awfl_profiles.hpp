// awfl_profiles.hpp -- hydrostatic background profiles of the initial data
// (Profiles.h:8-47): constant potential temperature and constant
// Brunt-Vaisala frequency, density and pressure at height z; and the linear
// ellipsoid of Profiles.h:49-58.
#pragma once
#include <cmath>

#include "awfl_cell.hpp"

namespace awfl {

constexpr double CP = 1004.;
constexpr double RD = 287.;
constexpr double P0 = 1.0e5;

__host__ __device__ inline double const_theta_density(double t0, double z) {
  double exner = 1. - GRAV * z / (CP * t0);
  double p = pow(exner, CP / RD) * P0;
  double rt = pow(p / C0, 1. / GAMMA);
  return rt / t0;
}
__host__ __device__ inline double const_theta_pressure(double t0, double z) {
  double r = const_theta_density(t0, z);
  return C0 * pow(r * t0, GAMMA);
}
__host__ __device__ inline double bvf_theta(double t0, double bvf, double z) { return t0 * exp(bvf * bvf * z / GRAV); }
__host__ __device__ inline double bvf_density(double t0, double bvf, double z) {
  double t = bvf_theta(t0, bvf, z);
  double exner = 1. - GRAV * GRAV / (CP * bvf * bvf) * (t - t0) / (t * t0);
  double p = pow(exner, CP / RD) * P0;
  double rt = pow(p / C0, 1. / GAMMA);
  return rt / t;
}
__host__ __device__ inline double bvf_pressure(double t0, double bvf, double z) {
  double t = bvf_theta(t0, bvf, z);
  double r = bvf_density(t0, bvf, z);
  return C0 * pow(r * t, GAMMA);
}

// Profiles.h:49-58
__host__ __device__ inline double ellipsoid_linear(double x, double y, double z, double x0, double y0, double z0,
                                                   double xrad, double yrad, double zrad, double amp) {
  double xn = (x - x0) / xrad, yn = (y - y0) / yrad, zn = (z - z0) / zrad;
  double dist = sqrt(xn * xn + yn * yn + zn * zn);
  return amp * fmax(1. - dist, 0.);
}

}  // namespace awfl
