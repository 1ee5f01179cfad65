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
    out = tmp_path / "thermal_64x64x32.bin"
    assert out.exists() and out.stat().st_size == 8 * (1 + 6 * 64 * 64 * 32)
