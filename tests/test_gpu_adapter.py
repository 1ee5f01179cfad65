"""The C++ Spatial/Temporal adapter (include/awfl/spatial_hip.hpp) and the
C++ driver run on the device: device and host applySingleTendency paths and
the fused step agree; the driver completes a short run."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_adapter_paths_agree(tmp_path):
    _gpu()
    exe = os.path.join(ROOT, "tests", "native", "adapter_test")
    if not os.path.exists(exe):
        pytest.fail("adapter_test not built (python -m awflcloud_amd.build)")
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "native", "adapter_small.yaml")], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


@pytest.mark.parametrize("fused", [False, True])
def test_driver_runs(tmp_path, fused):
    _gpu()
    exe = os.path.join(ROOT, "awflcloud_amd", "awfl_driver")
    if not os.path.exists(exe):
        pytest.fail("awfl_driver not built (python -m awflcloud_amd.build)")
    args = [exe, os.path.join(ROOT, "awflcloud_amd", "driver", "thermal_64x64x32.yaml"), "--max-steps", "10"]
    if fused:
        args.append("--fused")
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "Walltime" in r.stdout and "Steps: 10" in r.stdout
    # netCDF output (S:1704-1791): records at t = 0 and every outFreq = 3.6 s
    from scipy.io import netcdf_file
    import numpy as np
    with netcdf_file(str(tmp_path / "thermal_64x64x32.nc"), "r", mmap=False) as f:
        assert (f.dimensions["x"], f.dimensions["y"], f.dimensions["z"]) == (64, 64, 32)
        t = f.variables["t"][:].copy()
        assert len(t) == 3 and t[0] == 0.0 and 3.6 <= t[1] < 3.6 + 0.8 and 7.2 <= t[2] < 7.2 + 0.8
        for name in ("dens_pert", "u", "v", "w", "pot_temp_pert", "pressure_pert"):
            assert f.variables[name].shape == (3, 32, 64, 64)
        assert np.allclose(f.variables["x"][:], (np.arange(64) + 0.5) * 312.5)
        assert np.allclose(f.variables["z"][:], (np.arange(32) + 0.5) * 312.5)
        assert np.all(np.diff(f.variables["hyDens"][:]) < 0)
        th0 = f.variables["pot_temp_pert"][0]
        assert 1.5 < th0.max() <= 2.0 + 2e-3 and abs(th0.min()) <= 2e-3
        assert np.abs(f.variables["w"][0]).max() == 0.0 and np.abs(f.variables["w"][2]).max() > 0.0


def test_adapter_with_tracers(tmp_path):
    """Adapter built with AWFL_TRACERS: the driver's three tracers through
    Temporal_ader::timeStep and the fused step, and their mixing ratios in
    the netCDF output (S:1781-1787)."""
    _gpu()
    exe = os.path.join(ROOT, "tests", "native", "adapter_tracers_test")
    if not os.path.exists(exe):
        pytest.fail("adapter_tracers_test not built (python -m awflcloud_amd.build)")
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "native", "adapter_small.yaml")], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    from scipy.io import netcdf_file
    import numpy as np
    with netcdf_file(str(tmp_path / "adapter_small.nc"), "r", mmap=False) as f:
        for name in ("tracer_uniform", "tracer_theta", "tracer_block"):
            assert f.variables[name].shape == (2, 16, 24, 32)
        q = f.variables["tracer_uniform"][:].copy()
        # rh * 1 over the constant-theta background: a mixing ratio of 1 that
        # transport keeps
        assert np.abs(q - 1.0).max() <= 1e-12
        b = f.variables["tracer_block"][0]
        assert b.max() == pytest.approx(1.0, abs=1e-12) and b.min() == 0.0
        th = f.variables["tracer_theta"][0]
        assert 1.0 < th.max() <= 2.0
