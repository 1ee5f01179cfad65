"""The kernels' FP64 helpers (awfl_fastmath.hpp), built for the host and
checked against 200-bit mpmath: pow_gamma(x) = x^GAMMA within 1 ulp over the
range of rho*theta and far beyond, rcp(x) within 1 ulp.  The GPU test checks
the device build of the same routine against these host results."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "fastmath_probe.cpp")
GAMMA = 1.40027894002789400278940027894
C0 = 27.5629410929725921310572974482


@pytest.fixture(scope="module")
def probe():
    d = tempfile.mkdtemp()
    so = os.path.join(d, "probe.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", SRC, "-o", so], check=True)
    return C.CDLL(so)


def _call(L, name, x):
    y = np.zeros_like(x)
    D = C.POINTER(C.c_double)
    getattr(L, name)(x.ctypes.data_as(D), y.ctypes.data_as(D), C.c_long(x.size))
    return y


def host_pow_gamma(x):
    d = tempfile.mkdtemp()
    so = os.path.join(d, "probe.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", SRC, "-o", so], check=True)
    return _call(C.CDLL(so), "probe_pow_gamma", np.ascontiguousarray(x, dtype=np.float64))


def test_pow_gamma_within_one_ulp(probe):
    from mpmath import mp, mpf, power
    mp.prec = 200
    rng = np.random.default_rng(0)
    x = np.concatenate([np.exp(rng.uniform(np.log(1e-3), np.log(1e8), 4000)), rng.uniform(50, 500, 4000),
                        [1.0, 0.5, 2.0, 348.0, 1e5, 0.7071067811865476, 1.4142135623730951]])
    y = _call(probe, "probe_pow_gamma", x)
    g = mpf(GAMMA)
    worst = 0.0
    for xi, yi in zip(x, y):
        ex = power(mpf(float(xi)), g)
        worst = max(worst, float(abs(mpf(float(yi)) - ex) / mpf(float(np.spacing(float(ex))))))
    assert worst < 1.0, worst


def test_rcp_within_one_ulp(probe):
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-30, 30, 10000)) * rng.choice([-1.0, 1.0], 10000)
    r = _call(probe, "probe_rcp", x)
    assert np.all(np.abs(r - 1.0 / x) <= np.spacing(np.abs(1.0 / x)))
