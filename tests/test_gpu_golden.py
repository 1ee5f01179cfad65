"""The GPU path against the committed oracle fixtures (tests/golden/), through
the C ABI, without running the oracle:

  * 12 x 10 x 8 boxes (periodic x/y + wall z; walls everywhere) from the
    fixture's initial state: the tendency of every direction (<= 5e-12 of
    each variable's largest value), the states after 1, 2, 3 steps (density
    and rho*theta <= 1e-13 of the full field, every variable <= 1e-10 of its
    perturbation maximum), three tracers' tendencies and their values after
    two steps (<= 1e-11 of each tracer's scale);
  * the 64 x 64 x 32 thermal bubble initialised on the device: per-variable
    L2 and Linf norms of the state at steps 1, 10 and 100 (<= 1e-9).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "oracle_steps_12x10x8.npz")
BUBBLE = dict(shape=1, x0=10000.0, y0=10000.0, z0=2000.0, xr=2000.0, yr=2000.0, zr=2000.0, amp=2.0, pwr=1.0)


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _rel(a, b):
    return max(float(np.abs(a[l] - b[l]).max() / max(np.abs(b[l]).max(), 1e-300)) for l in range(a.shape[0]))


@pytest.mark.parametrize("tag,bc", [("per", (0, 0, 1)), ("wall", (1, 1, 1))])
def test_small_boxes_match_fixtures(tag, bc):
    _need_gpu()
    from awflcloud_amd import Dycore
    F = np.load(FIX)
    g = lambda k: F["%s_%s" % (tag, k)]  # noqa: E731
    D = Dycore(12, 10, 8, bc=bc)
    hs = D.hs
    I = lambda a: a[:, hs:hs + 8, hs:hs + 10, hs:hs + 12]  # noqa: E731
    st0 = torch.from_numpy(g("state0").copy()).cuda()
    dt = float(g("dt"))
    D.set_time_step(dt)
    tend = D.new_tend()
    for d in range(3):
        D.dim_switch = 1
        D.compute_tendencies(st0, tend, dt, d)
        err = _rel(tend.cpu().numpy(), g("tend%d" % d))
        print("%s dir %d tendency err %.2e" % (tag, d, err), flush=True)
        assert err <= 5e-12
    st = st0.clone()
    D.dim_switch = 1
    for s in range(1, 4):
        D.time_step(st, dt)
        got, want = I(st.cpu().numpy()), I(g("state%d" % s))
        # density and rho*theta relative to the full fields (perturbation + background)
        rb = D.background(0)[hs:hs + 8]
        rtb = D.background(3)[hs:hs + 8]
        e_r = float(np.abs(got[0] - want[0]).max() / np.abs(want[0] + rb[:, None, None]).max())
        e_t = float(np.abs(got[4] - want[4]).max() / np.abs(want[4] + rtb[:, None, None]).max())
        e_all = _rel(got, want)
        print("%s step %d: rho %.2e rho*theta %.2e per-variable %.2e" % (tag, s, e_r, e_t, e_all), flush=True)
        assert e_r <= 1e-13 and e_t <= 1e-13 and e_all <= 1e-10
    # tracers
    D.set_tracers([0, 1, 1])
    tr0 = torch.from_numpy(g("tracers0").copy()).cuda()
    ttend = torch.zeros(D.tracer_tend_shape, dtype=torch.float64, device="cuda")
    for d in range(3):
        D.dim_switch = 1
        D.compute_tendencies_tracers(st0, tr0, tend, ttend, dt, d)
        err = _rel(ttend.cpu().numpy(), g("ttend%d" % d))
        print("%s dir %d tracer tendency err %.2e" % (tag, d, err), flush=True)
        assert err <= 1e-11
    st, tr = st0.clone(), tr0.clone()
    D.dim_switch = 1
    for _ in range(2):
        D.time_step_tracers(st, tr, dt)
    err = _rel(I(tr.cpu().numpy()), I(g("tracers2")))
    print("%s tracers after 2 steps err %.2e" % (tag, err), flush=True)
    assert err <= 1e-11


def test_c1_norms_match_fixture():
    _need_gpu()
    from awflcloud_amd import Dycore, bubble_spec
    want = json.load(open(os.path.join(ROOT, "tests", "golden", "oracle_c1_norms.json")))
    D = Dycore(64, 64, 32, bc=(0, 0, 1))
    st = D.init_state(bubble_spec([BUBBLE], noise_amp=1e-3, noise_seed=1))
    dt = D.compute_time_step(st, 0.8)
    assert abs(dt - want["dt"]) <= 1e-15 * want["dt"]
    D.dim_switch = 1
    done = 0
    for target in (1, 10, 100):
        for _ in range(target - done):
            D.time_step(st, want["dt"])
        done = target
        I = D.interior(st)
        w = want["steps"][str(target)]
        l2 = [float(torch.sqrt(torch.mean(I[l] ** 2))) for l in range(5)]
        li = [float(I[l].abs().max()) for l in range(5)]
        e2 = max(abs(a - b) / b for a, b in zip(l2, w["l2"]))
        ei = max(abs(a - b) / b for a, b in zip(li, w["linf"]))
        print("step %d: L2 err %.2e Linf err %.2e" % (target, e2, ei), flush=True)
        assert e2 <= 1e-9 and ei <= 1e-9
