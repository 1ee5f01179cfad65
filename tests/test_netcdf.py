"""The netCDF classic writer used by the adapter's output (include/awfl/
netcdf_classic.hpp), built for the host with g++; the file is read back with
scipy.io.netcdf_file, an independent reader of the classic format."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "netcdf_test.cpp")


def test_netcdf_classic_roundtrip():
    from scipy.io import netcdf_file
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "netcdf_test")
    subprocess.run(["g++", "-O2", "-std=c++17", SRC, "-o", exe], check=True)
    path = os.path.join(d, "out.nc")
    subprocess.run([exe, path], check=True, timeout=60)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions["x"] == 5 and f.dimensions["z"] == 2 and f.dimensions["t"] is None
        assert np.array_equal(f.variables["x"][:], np.arange(5) + 0.5)
        assert np.array_equal(f.variables["hyDens"][:], [1.0, 0.5])
        assert np.array_equal(f.variables["t"][:], [0.0, 1.5])
        a = f.variables["a"][:]
        b = f.variables["b"][:]
        assert a.shape == (2, 2, 3, 5) and a.dtype.kind == "f" and a.dtype.itemsize == 8
        n = np.arange(30).reshape(2, 3, 5)
        assert np.array_equal(a[0], n) and np.array_equal(a[1], n + 1000.0)
        assert np.array_equal(b[0], -n * 0.25) and np.array_equal(b[1], -n * 0.25 + 1)
