"""The alternative time integrators of the oracle (qjrmsCode/TimeIntegrator.cpp
restated over the Agrid spatial operator): the half-step Strang-type
splitting and SSPRK3 on the semi-discrete operator.

Pinned by properties the mathematics fixes:
  * SSPRK3 is third order in time: on a smooth problem the one-interval error
    against a fine-step solution drops ~8x when dt halves;
  * the semi-discrete tendency of a horizontally uniform state has no x or y
    part, and both integrators keep the hydrostatic rest state at rest;
  * mass is conserved with periodic x/y (flux form);
  * the splitting with its two z half steps differs from the full-step z
    sweep (the half steps use dt/2 in the transform and in the update)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pyoracle  # noqa: E402
from helpers import bubble_dicts  # noqa: E402


def _state(O, amp=2.0):
    st = O.init_state(pyoracle.make_spec(bubble_dicts(amp=amp, r=2500.0, z0=4000.0), u0=5.0, v0=-3.0))
    return st


def test_ssprk3_third_order_in_time():
    O = pyoracle.Oracle(16, 12, 16, ord=5, ngll=3)
    st0 = _state(O)
    T = 1.0
    ref = st0.copy()
    for _ in range(128):
        O.ssprk3_step(ref, T / 128)
    errs = []
    for n in (1, 2, 4, 8):
        s = st0.copy()
        for _ in range(n):
            O.ssprk3_step(s, T / n)
        errs.append(np.abs(O.interior(s) - O.interior(ref))[1:4].max())
    ratios = [errs[i] / errs[i + 1] for i in range(3)]
    assert all(6.0 < r < 10.0 for r in ratios), (errs, ratios)


def test_rest_state_and_mass():
    for step in ("strang", "ssprk3"):
        O = pyoracle.Oracle(16, 12, 10)
        rest = O.init_state(pyoracle.make_spec([]))
        dt = O.compute_time_step(rest, 0.5)
        s = rest.copy()
        for n in range(3):
            if step == "strang":
                O.strang_step(s, dt, n % 2 == 0)
            else:
                O.ssprk3_step(s, dt)
        assert np.abs(O.interior(s)).max() < 1e-9, step
        st = _state(O)
        m0 = O.interior(st)[0].sum()
        for n in range(3):
            if step == "strang":
                O.strang_step(st, dt, n % 2 == 0)
            else:
                O.ssprk3_step(st, dt)
        m1 = O.interior(st)[0].sum()
        assert abs(m1 - m0) <= 1e-12 * np.abs(O.interior(st)[0]).sum() + 1e-14, (step, m0, m1)


def test_half_steps_differ_from_full_z_step():
    O = pyoracle.Oracle(12, 10, 12)
    st = _state(O)
    dt = O.compute_time_step(st, 0.5)
    a, b = st.copy(), st.copy()
    O.sweep(a, dt / 2, 2)
    O.sweep(a, dt / 2, 2)
    O.sweep(b, dt, 2)
    d = np.abs(O.interior(a) - O.interior(b))[3].max()
    assert d > 1e-9 * np.abs(O.interior(b)[3]).max()
