"""Attribution of the HIP step's difference from the CPU oracle.

The production sweeps differ from the oracle after one step of the headline
workload (256x256x128 bubble, WENO5/ADER3) by up to ~2e-11 of a variable's
perturbation maximum, because a tendency is a difference of nearly equal
face fluxes (the ~1e5 Pa pressure cancels) and magnifies every rounding
difference about a thousandfold.  The parity builds of the FP64 sweeps
(awflcloud_amd/build.py PARITY_BUILDS) switch the production arithmetic back
toward the reference's one piece at a time:

  div       IEEE division and sqrt in place of refined reciprocals and the
            rsq-based sound speed
  pow       the general pow_gamma in place of the near-background series
  contract  no FMA contraction
  bridge    the WENO bridge polynomial by its three statements
  fast_all  div + pow + bridge + no contraction
  ref_libm  the cell and face bodies in the reference's operation order, IEEE
            division and sqrt, no contraction, the device library's pow
  ref_pg    the same with pow_gamma (the double-double method) as the power

Each is compared with the oracle using the C library's pow (mode 0) and with
the oracle using pow_gamma's method (mode 1, pinned in
test_oracle_pow_mode.py).  ref_pg against oracle mode 1 has every operation
and every power function in common: it must agree to the last bit, which
shows that nothing but the arithmetic rewrites and the power function
separates the production step from the reference's arithmetic.  The figures
are printed for every build; the production build keeps its own gates in
test_gpu_parity.py."""
import os

import numpy as np
import pytest

from helpers import bubble_dicts, full_rel_err, per_var_rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BUILDS = ["production", "div", "pow", "contract", "bridge", "fast_all", "ref_libm", "ref_pg"]


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _libpath(name):
    from awflcloud_amd import build
    from awflcloud_amd.dycore import LIB_PATH
    return LIB_PATH if name == "production" else build.parity_lib(name)


def _oracle_steps(host, nx, ny, nz, bc, dt, nsteps, mode):
    import pyoracle
    O = pyoracle.Oracle(nx, ny, nz, bc=bc, nthreads=0)
    O.set_dt_cache(dt)
    O.dim_switch = 1
    h = host.copy()
    pyoracle.set_pow_mode(mode)
    try:
        for _ in range(nsteps):
            O.time_step(h, dt)
    finally:
        pyoracle.set_pow_mode(0)
    hs = O.hs
    rbar = O.background(0)[hs:hs + nz][:, None, None]
    rtbar = O.background(3)[hs:hs + nz][:, None, None]
    return O.interior(h).copy(), rbar, rtbar


def _gpu_steps(name, st0, nx, ny, nz, bc, dt, nsteps):
    from awflcloud_amd import Dycore
    D = Dycore(nx, ny, nz, bc=bc, library=_libpath(name))
    D.set_time_step(dt)
    D.dim_switch = 1
    st = st0.clone()
    for _ in range(nsteps):
        D.time_step(st, dt)
    out = D.interior(st).cpu().numpy()
    D.close()
    return out


@pytest.fixture(scope="module")
def headline():
    """one step of the 256x256x128 bubble: initial state, dt and both oracles"""
    _need_gpu()
    from awflcloud_amd import Dycore, bubble_spec
    nx, ny, nz, bc = 256, 256, 128, (0, 0, 1)
    D = Dycore(nx, ny, nz, bc=bc)
    st0 = D.init_state(bubble_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=1))
    dt = D.compute_time_step(st0, 0.8)
    host = st0.cpu().numpy().copy()
    D.close()
    ref = {m: _oracle_steps(host, nx, ny, nz, bc, dt, 1, m) for m in (0, 1)}
    return st0, dt, ref


@pytest.mark.parametrize("name", BUILDS)
def test_build_step_against_oracle(name, headline, figures):
    st0, dt, ref = headline
    if name != "production" and not os.path.exists(_libpath(name)):
        pytest.fail("parity build %s is missing; run `python -m awflcloud_amd.build`" % name)
    got = _gpu_steps(name, st0, 256, 256, 128, (0, 0, 1), dt, 1)
    out = {}
    for mode in (0, 1):
        r, rbar, rtbar = ref[mode]
        out[mode] = (per_var_rel_err(got, r), full_rel_err(got, r, rbar, rtbar))
    figures("%-10s step 1 of 256x256x128; of each variable's perturbation max: vs oracle(libm pow) %s, "
            "vs oracle(pow_gamma) %s; rho, rho*theta of the full field vs oracle(libm) %.1e %.1e"
            % (name, ["%.1e" % e for e in out[0][0]], ["%.1e" % e for e in out[1][0]],
               out[0][1][0], out[0][1][4]))
    # every build keeps the per-step bar on the full fields
    assert max(out[0][1][0], out[0][1][4]) <= 1e-12
    assert max(out[0][0]) <= 1e-10
    if name == "ref_pg":
        # the same operations and the same power function as oracle mode 1
        assert max(out[1][0]) == 0.0, out[1][0]


@pytest.mark.parametrize("bc", [(1, 1, 1), (0, 0, 0), (1, 0, 1)])
def test_ref_pg_bitwise_over_steps(bc, figures):
    """the reference-order build with pow_gamma against oracle mode 1 over ten
    steps of a bubble in wind, with walls and periodic axes in every
    direction: bit for bit"""
    _need_gpu()
    from awflcloud_amd import Dycore, bubble_spec
    nx, ny, nz = 40, 36, 28
    D = Dycore(nx, ny, nz, bc=bc, library=_libpath("ref_pg"))
    st0 = D.init_state(bubble_spec(bubble_dicts(x0=9000.0, y0=11000.0), noise_amp=1e-3, noise_seed=3,
                                   u0=4.0, v0=-3.0))
    dt = D.compute_time_step(st0, 0.8)
    host = st0.cpu().numpy().copy()
    D.close()
    r, _, _ = _oracle_steps(host, nx, ny, nz, bc, dt, 10, 1)
    got = _gpu_steps("ref_pg", st0, nx, ny, nz, bc, dt, 10)
    diff = int(np.count_nonzero(got != r))
    figures("bc %s: %d of %d values differ after 10 steps" % (bc, diff, got.size))
    assert diff == 0
