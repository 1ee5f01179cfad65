"""The rewritten WENO forms of awflcloud_amd/csrc/awfl_cell.hpp, built for the
host with g++ (tests/native/weno_forms_test.cpp): the centre-cell average
identity the reconstruction relies on (every candidate and the bridge, ideal
weights summing to one), the mirrored bridge map against the plain one, and
reconstruct() against a statement-by-statement restatement of the reference
reconstruction (WenoLimiter.h:93-172, S:1813-1819) on random, smooth, stepped
and near-constant stencils at (ord, ngll) = (3, 2), (5, 3), (9, 5)."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "weno_forms_test.cpp")


def test_weno_forms_host():
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "weno_forms_test")
    subprocess.run(["g++", "-O2", "-std=c++17", SRC, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    fails = [ln for ln in r.stdout.splitlines() if ln.endswith("FAIL")]
    assert r.returncode == 0 and not fails, "\n".join(fails) + r.stdout[-400:]
    assert "PASSED" in r.stdout
