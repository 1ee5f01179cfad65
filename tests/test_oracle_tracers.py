"""Tracer transport of the CPU oracle (oracle_*_tr), pinned by what the
mathematics fixes.  The reference compiles its tracer kernels out (#define
NO_TRACERS, S:10), so there is no reference output to compare with; these
checks each fail on a plausible mistake in the restatement:

  * a tracer of uniform mixing ratio q0 (mass q0 * rho) keeps q0 to rounding:
    its transform, edges and upwind flux are q0 times the density's, so a
    wrong level, index or upwind edge breaks it;
  * tracer mass is conserved (periodic x/y, walls in z carry no mass flux);
  * a blob in a uniform wind over the rest state moves with the wind: its
    centroid lands on the exact shift and its shape stays within the
    scheme's diffusion (a sign or direction error moves it elsewhere);
  * the positivity limiter keeps positive tracers non-negative where the
    unlimited scheme undershoots;
  * tracers do not change the state.
"""
import numpy as np
import pytest

from helpers import bubble_dicts

P = pytest.importorskip("pyoracle")


def _setup(n=(24, 20, 12), bc=(0, 0, 1), ord=5, ngll=3, pos=(0, 1, 1), u0=5.0, v0=-3.0):
    O = P.Oracle(*n, bc=bc, ord=ord, ngll=ngll)
    st = O.init_state(P.make_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=5, u0=u0, v0=v0))
    dt = O.compute_time_step(st, 0.8)
    O.set_tracers(list(pos))
    return O, st, dt


def _fields(O, st):
    rho = O.density(st)
    x = (np.arange(O.nx) + 0.5) * O.dx
    y = (np.arange(O.ny) + 0.5) * O.dy
    z = (np.arange(O.nz) + 0.5) * O.dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    blob = np.exp(-(((X - 8000) / 2500) ** 2 + ((Y - 9000) / 2500) ** 2 + ((Z - 4000) / 1500) ** 2))
    blob[blob < 0.05] = 0.0
    tr = np.zeros(O.tracer_shape())
    TI = O.tracer_interior(tr)
    TI[0] = 0.25 * rho
    TI[1] = blob * rho
    TI[2] = (blob > 0.3) * rho
    return tr


CASES = [
    dict(n=(24, 20, 12), bc=(0, 0, 1)),
    dict(n=(24, 20, 12), bc=(1, 1, 1)),
    dict(n=(24, 20, 12), bc=(0, 0, 0)),
    dict(n=(20, 16, 12), bc=(0, 1, 1), ord=3, ngll=2),
    dict(n=(32, 1, 16), bc=(0, 0, 1)),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c["n"])) + "-bc" +
                         "".join(map(str, c["bc"])) + "-o%d" % c.get("ord", 5))
def test_uniform_ratio_mass_and_state(case):
    O, st, dt = _setup(**case)
    tr = _fields(O, st)
    TI = O.tracer_interior
    m0 = TI(tr).sum(axis=(1, 2, 3))
    ref = st.copy()
    O2, _, _ = _setup(**case)
    O2.set_dt_cache(dt)
    for _ in range(8):
        O.time_step_tr(st, tr, dt)
        O2.time_step(ref, dt)
    # tracers leave the state untouched
    assert np.array_equal(O.interior(st), O2.interior(ref))
    q = TI(tr)[0] / O.density(st)
    assert np.abs(q - 0.25).max() <= 1e-14
    m1 = TI(tr).sum(axis=(1, 2, 3))
    periodic_xy = case["bc"][0] == 0 and (case["bc"][1] == 0 or case["n"][1] == 1)
    if periodic_xy:
        assert np.all(np.abs(m1 - m0) <= 1e-13 * np.abs(m0))
    else:
        # x/y walls pass the Riemann mass flux of the wall cell's edge state
        # (its normal momentum is not zeroed there): density and the
        # uniform-ratio tracer change together
        assert abs(m1[0] - 0.25 * O.density(st).sum()) <= 1e-13 * m1[0]


def test_positivity_limiter():
    O, st, dt = _setup(n=(32, 24, 16))
    tr = _fields(O, st)
    for _ in range(20):
        O.time_step_tr(st, tr, dt)
    TI = O.tracer_interior(tr)
    assert TI[1].min() >= -1e-9 * TI[1].max()
    assert TI[2].min() >= -1e-8 * TI[2].max()
    # without tracerPos the step function undershoots visibly
    O2, st2, _ = _setup(n=(32, 24, 16), pos=(0, 0, 0))
    O2.set_dt_cache(dt)
    tr2 = _fields(O2, st2)
    for _ in range(20):
        O2.time_step_tr(st2, tr2, dt)
    assert O2.tracer_interior(tr2)[2].min() < -1e-6


@pytest.mark.parametrize("wind", [(10.0, 0.0), (0.0, -10.0)])
def test_blob_moves_with_uniform_wind(wind):
    u0, v0 = wind
    O = P.Oracle(64, 64, 6, bc=(0, 0, 1))
    st = O.init_state(P.make_spec([], u0=u0, v0=v0))
    st0 = st.copy()
    dt = O.compute_time_step(st, 0.8)
    O.set_tracers([0])
    tr = np.zeros(O.tracer_shape())
    x = (np.arange(O.nx) + 0.5) * O.dx
    y = (np.arange(O.ny) + 0.5) * O.dy
    z = (np.arange(O.nz) + 0.5) * O.dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")

    def prof(sx, sy):
        ddx = (X - 10000 - sx + 10000) % 20000 - 10000
        ddy = (Y - 10000 - sy + 10000) % 20000 - 10000
        return np.exp(-((ddx / 2000) ** 2 + (ddy / 2000) ** 2))

    O.tracer_interior(tr)[0] = prof(0, 0) * O.density(st)
    n = 40
    for _ in range(n):
        O.time_step_tr(st, tr, dt)
    # the rest state with wind is steady to rounding
    assert np.abs(O.interior(st) - O.interior(st0)).max() <= 1e-11
    q = O.tracer_interior(tr)[0] / O.density(st)
    sx, sy = u0 * n * dt, v0 * n * dt
    assert np.abs(q - prof(sx, sy)).max() <= 1e-4
    w = q.sum(axis=0)
    cx = (w.sum(axis=0) * x).sum() / w.sum()
    cy = (w.sum(axis=1) * y).sum() / w.sum()
    assert abs(cx - (10000 + sx)) <= 1e-6 * O.dx and abs(cy - (10000 + sy)) <= 1e-6 * O.dy


def test_split_order_and_zero_tracers_match_state_only_step():
    """With tracers the split order and dimSwitch follow the state-only
    step; zero tracers stay zero."""
    O, st, dt = _setup()
    tr = np.zeros(O.tracer_shape())
    for _ in range(3):
        O.time_step_tr(st, tr, dt)
    assert O.dim_switch == 0
    assert np.abs(tr).max() == 0.0
