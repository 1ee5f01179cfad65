"""Alternative time integrators (awfl_time_step_strang / _ssprk3) on the
GPU against the oracle's restatement of qjrmsCode/TimeIntegrator.cpp over the
same spatial operator (pyoracle.Oracle.strang_step / ssprk3_step), on the
same bytes.  Gates as for the ADER-DT step: density and rho*theta relative to
the full field <= 1e-12 per step; every variable relative to its own
perturbation max <= 1e-10 after the steps."""
import numpy as np
import pytest

from helpers import bubble_dicts, full_rel_err, per_var_rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("method", ["strang", "ssprk3"])
@pytest.mark.parametrize("bc,n", [((0, 0, 1), (32, 24, 16)), ((1, 1, 1), (28, 20, 16)), ((0, 0, 1), (40, 1, 16))])
def test_integrator_matches_oracle(method, bc, n):
    _need_gpu()
    import pyoracle
    from awflcloud_amd import Dycore, bubble_spec
    nx, ny, nz = n
    print("case", method, bc, n, flush=True)
    D = Dycore(nx, ny, nz, bc=bc)
    O = pyoracle.Oracle(nx, ny, nz, bc=bc)
    # ny == 1: y2d = 0 puts the 2-D plane at ylen/2, through the bubble
    st = D.init_state(bubble_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=9, u0=4.0, y2d=0.0))
    host = st.cpu().numpy().copy()
    dt = D.compute_time_step(st, 0.8 if method == "strang" else 0.3)
    hs, nzz = O.hs, O.nz
    rbar = O.background(0)[hs:hs + nzz][:, None, None]
    rtbar = O.background(3)[hs:hs + nzz][:, None, None]
    D.dim_switch = 1
    for s in range(3):
        if method == "strang":
            D.time_step_strang(st, dt)
            O.strang_step(host, dt, s % 2 == 0)
        else:
            D.time_step_ssprk3(st, dt)
            O.ssprk3_step(host, dt)
        g = D.interior(st).cpu().numpy()
        f = full_rel_err(g, O.interior(host), rbar, rtbar)
        assert max(f[0], f[4]) <= 1e-12, (s, f)
    e = per_var_rel_err(g, O.interior(host))
    print("3 steps pert", ["%.2e" % v for v in e], "full", ["%.2e" % v for v in f], flush=True)
    assert max(e) <= 1e-10
    assert bool(np.isfinite(g).all())
