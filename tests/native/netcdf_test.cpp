// Writes a small netCDF classic (CDF-2) file with include/awfl/netcdf_classic.hpp:
// fixed coordinates, an unlimited t and two record variables, the second
// record appended after reopening the file.  tests/test_netcdf.py reads it
// back with scipy.io.netcdf_file.
#include <cstdio>
#include <vector>

#include "../../include/awfl/netcdf_classic.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const int nx = 5, ny = 3, nz = 2;
  using namespace awfl::nc;
  std::vector<Dim> dims = {{"x", nx}, {"y", ny}, {"z", nz}, {"t", 0}};
  std::vector<Var> vars = {{"x", {0}}, {"hyDens", {2}}, {"t", {3}}, {"a", {3, 2, 1, 0}}, {"b", {3, 2, 1, 0}}};
  std::vector<double> x(nx), h(nz), a(nx * ny * nz), b(nx * ny * nz);
  for (int i = 0; i < nx; i++) x[i] = i + 0.5;
  for (int k = 0; k < nz; k++) h[k] = 1.0 / (k + 1);
  {
    ClassicFile f;
    f.create(argv[1], dims, vars);
    f.write_fixed("x", x.data());
    f.write_fixed("hyDens", h.data());
    for (int r = 0; r < 1; r++) {
      double t = 0.0;
      for (int n = 0; n < nx * ny * nz; n++) {
        a[n] = n + 1000.0 * r;
        b[n] = -n * 0.25 + r;
      }
      f.write_record("t", r, &t);
      f.write_record("a", r, a.data());
      f.write_record("b", r, b.data());
    }
  }
  {
    ClassicFile f;
    f.open(argv[1]);
    long r = (long)f.num_records();
    double t = 1.5;
    for (int n = 0; n < nx * ny * nz; n++) {
      a[n] = n + 1000.0 * r;
      b[n] = -n * 0.25 + r;
    }
    f.write_record("t", r, &t);
    f.write_record("a", r, a.data());
    f.write_record("b", r, b.data());
  }
  std::printf("ok\n");
  return 0;
}
