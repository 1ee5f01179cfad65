"""Tangent-linear and adjoint numerics of one cell and one face
(awflcloud_amd/csrc/awfl_adjoint.hpp), built for the host with g++.

tests/native/adjoint_cell_test.cpp checks, for every sweep direction, the
(ord, ngll) variants (3, 2), (5, 3), (9, 5), WENO on and off and z walls:
  * the dot-product identity <ybar, J xdot> = <J^T ybar, xdot> between the
    Dual-number tangent and the hand-written reverse (<= 1e-12 relative);
  * the tangent against fourth-order central differences of the production
    FP64 forward (cell_edges / riemann) (<= 1e-8 relative);
  * the generic primal against the production forward (<= 1e-13)."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "adjoint_cell_test.cpp")


def test_cell_and_face_adjoint_host():
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "adjoint_cell_test")
    subprocess.run(["g++", "-O2", "-std=c++17", SRC, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    fails = [ln for ln in r.stdout.splitlines() if ln.endswith("FAIL")]
    assert r.returncode == 0 and not fails, "\n".join(fails[:20]) + r.stdout[-400:]
    assert "PASSED" in r.stdout
