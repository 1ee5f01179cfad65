"""Pin the CPU oracle against what the mathematics fixes.

The reference ships no tests or golden vectors for the step (its only stored
known answer, the derivative operator of sage/derivatives.ipynb, is pinned in
test_tables.py) and cannot be built here (its YAKL/yaml-cpp/netCDF
dependencies are absent).  Each check below is chosen so that a plausible
restatement mistake -- a dropped term, a wrong sign or index, a transposed
table, a swapped upwind branch -- fails it:

* WENO: constants and degree-<=hs data reproduced exactly, translation
  invariance, mirror symmetry, 5th-order convergence, bounded at a step.
* ADER differential transform: equals the Taylor coefficients of the
  collocation ODE system obtained independently by Picard iteration with
  power-series arithmetic (series division, exp(gamma*log) for the
  pressure), including the z-sweep gravity source, its level-0 broadcast
  (S:1439) and the wall conditions.
* Riemann: consistency F(q, q) = f(q) for all three directions (incl. the
  y-direction averaging of conservative variables), upwinding of the shear
  wave by the sign of the normal velocity, strict '>' at zero.
* Whole sweeps/steps: hydrostatic rest state stays at rest, mass is conserved
  to round-off, an entropy wave translates exactly at 3rd-order accuracy,
  the CFL time step matches its closed form, and the initial-data quadrature
  reproduces the background for a uniform theta'.
"""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

C0 = 27.5629410929725921310572974482
GAMMA = 1.40027894002789400278940027894
GRAV = 9.81

ORDERS = [(3, 2), (5, 3), (7, 4), (9, 5)]


def oracle(ord=5, ngll=3, nx=8, ny=8, nz=8, **kw):
    return pyoracle.Oracle(nx, ny, nz, ord=ord, ngll=ngll, **kw)


def gll_pts(n):
    sys.path.insert(0, os.path.join(ROOT, "awflcloud_amd", "csrc"))
    import gen_tables
    return np.array(gen_tables.build()["gll_pts_%d" % n])


# ---------------------------------------------------------------------------
# WENO
# ---------------------------------------------------------------------------
def cell_avgs(f_int, centres, h):
    """cell averages of f over [c-h/2, c+h/2] from its antiderivative."""
    return np.array([(f_int(c + h / 2) - f_int(c - h / 2)) / h for c in centres])


@pytest.mark.parametrize("o,n", ORDERS)
def test_weno_constant_and_translation(o, n):
    O = oracle(o, n)
    rng = np.random.default_rng(1)
    for c in (0.0, 1.0, -3.25, 1234.5):
        aw = O.weno_coefs(np.full(o, c))
        assert abs(aw[0] - c) <= 1e-15 * max(1, abs(c)) and np.all(np.abs(aw[1:]) <= 1e-13 * max(1, abs(c)))
    u = rng.standard_normal(o)
    g0 = O.reconstruct(u)
    g1 = O.reconstruct(u + 7.0)
    assert np.allclose(g1 - g0, 7.0, atol=1e-13)


@pytest.mark.parametrize("o,n", ORDERS)
def test_weno_mirror_symmetry(o, n):
    O = oracle(o, n)
    rng = np.random.default_rng(2)
    for _ in range(20):
        u = rng.standard_normal(o)
        g = O.reconstruct(u)
        gr = O.reconstruct(u[::-1].copy())
        assert np.allclose(gr, g[::-1], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("o,n", ORDERS)
def test_weno_exact_for_low_degree(o, n):
    """Cell averages of a degree-hs polynomial: every candidate is exact, so
    the reconstruction is exact whatever the weights."""
    hs = (o - 1) // 2
    O = oracle(o, n)
    rng = np.random.default_rng(3)
    coef = rng.standard_normal(hs + 1)
    P = np.polynomial.Polynomial(coef)
    Pint = P.integ()
    u = cell_avgs(Pint, np.arange(-hs, hs + 1, dtype=float), 1.0)
    g = O.reconstruct(u)
    assert np.allclose(g, P(gll_pts(n)), rtol=0, atol=1e-12)


@pytest.mark.parametrize("o,n", [(5, 3), (9, 5)])
def test_weno_convergence_order(o, n):
    """Edge value of a smooth profile converges at (close to) order ord."""
    hs = (o - 1) // 2
    O = oracle(o, n)
    errs = []
    hs_list = [1 / 8, 1 / 16, 1 / 32]
    for h in hs_list:
        x0 = 0.3
        centres = x0 + h * np.arange(-hs, hs + 1)
        u = cell_avgs(lambda x: -np.cos(2 * np.pi * x) / (2 * np.pi) + x * x * x / 3, centres, h)
        g = O.reconstruct(u)
        exact = np.sin(2 * np.pi * (x0 + h * gll_pts(n))) + (x0 + h * gll_pts(n)) ** 2
        errs.append(np.abs(g - exact).max())
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    assert min(rates) > o - 1.0, (errs, rates)


@pytest.mark.parametrize("o,n", ORDERS)
def test_weno_step_is_non_oscillatory(o, n):
    O = oracle(o, n)
    hs = (o - 1) // 2
    for jump_at in range(1, o):
        u = np.where(np.arange(o) < jump_at, 0.0, 1.0)
        g = O.reconstruct(u)
        assert g.min() > -0.02 and g.max() < 1.02, (jump_at, g)
    # the linear (weno off) reconstruction is exact for degree < ord data
    P = np.polynomial.Polynomial(np.random.default_rng(4).standard_normal(o))
    u = cell_avgs(P.integ(), np.arange(-hs, hs + 1, dtype=float), 1.0)
    assert np.allclose(O.reconstruct(u, doweno=False), P(gll_pts(n)), atol=1e-10)


# ---------------------------------------------------------------------------
# ADER differential transform vs Picard iteration with power series
# ---------------------------------------------------------------------------
class Series:
    """Truncated power series in t, length N, float64 coefficients."""

    def __init__(self, c, N):
        c = np.asarray(c, dtype=float)
        self.c = np.zeros(N)
        self.c[:min(N, c.size)] = c[:N]
        self.N = N

    def __add__(s, o):
        return Series(s.c + (o.c if isinstance(o, Series) else np.r_[o, np.zeros(s.N - 1)]), s.N)

    __radd__ = __add__

    def __sub__(s, o):
        return s + (-1.0) * o

    def __mul__(s, o):
        if not isinstance(o, Series):
            return Series(s.c * o, s.N)
        return Series(np.convolve(s.c, o.c)[:s.N], s.N)

    __rmul__ = __mul__

    def __truediv__(s, o):
        q = np.zeros(s.N)
        for k in range(s.N):
            q[k] = (s.c[k] - np.dot(q[:k], o.c[k:0:-1])) / o.c[0]
        return Series(q, s.N)

    def log(s):
        # log(a0 (1 + x)) with x = s/a0 - 1 (x0 = 0): series of log(1+x)
        x = Series(s.c / s.c[0], s.N) - 1.0
        out = Series([math.log(s.c[0])], s.N)
        p = Series([1.0], s.N)
        for k in range(1, s.N):
            p = p * x
            out = out + ((-1.0) ** (k + 1) / k) * p
        return out

    def exp(s):
        x = Series(np.r_[0.0, s.c[1:]], s.N)
        out = Series([1.0], s.N)
        p = Series([1.0], s.N)
        for k in range(1, s.N):
            p = p * x * (1.0 / k)
            out = out + p
        return out * math.exp(s.c[0])

    def __pow__(s, g):
        return (s.log() * g).exp()


def picard(q0, D, dx, direction, N, hyP=None, hyR=None, wall_bot=False, wall_top=False):
    """Taylor coefficients (N levels) of dq/dt = -(1/dx) D^T F(q) [+ gravity]."""
    nm = 1 + direction
    ng = q0.shape[1]
    S = [[Series([q0[l, i]], N) for i in range(ng)] for l in range(5)]
    src0 = None
    if direction == 2:
        src0 = -(q0[0, ng - 1] - hyR[ng - 1]) * GRAV   # S:1439 broadcast of the last point
    for _ in range(N):
        F = [[None] * ng for _ in range(5)]
        for i in range(ng):
            r = S[0][i]
            qn = S[nm][i]
            F[0][i] = qn
            for l in range(1, 5):
                F[l][i] = qn * S[l][i] / r
            F[nm][i] = F[nm][i] + C0 * (S[4][i] ** GAMMA)
            if direction == 2:
                F[nm][i] = F[nm][i] - hyP[i]
        new = [[None] * ng for _ in range(5)]
        for l in range(5):
            for i in range(ng):
                d = np.zeros(N)
                for s in range(ng):
                    d += D[s][i] * F[l][s].c
                c = np.zeros(N)
                c[0] = q0[l, i]
                rhs = -d / dx
                if direction == 2 and l == 3:
                    src = -(S[0][i].c - np.r_[hyR[i], np.zeros(N - 1)]) * GRAV
                    src[0] = src0
                    rhs = rhs + src
                c[1:] = rhs[:-1] / np.arange(1, N)
                if direction == 2 and l == 3 and ((wall_bot and i == 0) or (wall_top and i == ng - 1)):
                    c[1:] = 0.0
                new[l][i] = Series(c, N)
        S = new
    return np.array([[S[l][i].c for i in range(ng)] for l in range(5)])   # [l][ii][kt]


def dt_arrays(q0, direction, hyR=None):
    """Level-0 arrays in the oracle's diff_transform layout [var][kt][ii]."""
    nm = 1 + direction
    ng = q0.shape[1]
    nv = 11 if direction == 2 else 10
    a = np.zeros((nv, ng, ng))
    a[:5, 0, :] = q0
    for ii in range(ng):
        r = q0[0, ii]
        vel = q0[1:, ii] / r
        for m in range(4):
            a[5 + m, 0, ii] = r * vel[nm - 1] * vel[m]
        a[9, 0, ii] = (r * vel[3]) ** GAMMA
    if direction == 2:
        a[10, :, :] = -(q0[0, ng - 1] - hyR[ng - 1]) * GRAV
    return a


@pytest.mark.parametrize("o,n", ORDERS)
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_differential_transform_matches_picard_series(o, n, direction):
    nz = 6
    O = oracle(o, n, nz=nz, bc=(0, 0, 1))
    D = O.table(3).reshape(n, n)
    rng = np.random.default_rng(10 * o + direction)
    dx = 150.0
    k_cell = 2
    hyR = O.background(4).reshape(nz, n)[k_cell]
    hyP = O.background(5).reshape(nz, n)[k_cell]
    hyRT = O.background(7).reshape(nz, n)[k_cell]
    q0 = np.zeros((5, n))
    q0[0] = hyR * (1 + 0.01 * rng.standard_normal(n))
    q0[1:4] = q0[0] * rng.uniform(-20, 20, (3, n))
    q0[4] = hyRT * (1 + 0.01 * rng.standard_normal(n))
    got = O.diff_transform(direction, dt_arrays(q0, direction, hyR), dx, k_cell)
    ref = picard(q0, D, dx, direction, n, hyP=hyP, hyR=hyR)
    for l in range(5):
        for kt in range(1, n):
            g = got[l, kt, :]
            e = ref[l, :, kt]
            scale = np.abs(ref[l, :, kt]).max() + 1e-300
            assert np.abs(g - e).max() <= 1e-8 * scale, (l, kt, g, e)


def test_differential_transform_z_wall():
    n, nz = 3, 6
    O = oracle(5, n, nz=nz, bc=(0, 0, 1))
    D = O.table(3).reshape(n, n)
    for k_cell, bot, top in [(0, True, False), (nz - 1, False, True)]:
        hyR = O.background(4).reshape(nz, n)[k_cell]
        hyP = O.background(5).reshape(nz, n)[k_cell]
        hyRT = O.background(7).reshape(nz, n)[k_cell]
        rng = np.random.default_rng(k_cell)
        q0 = np.zeros((5, n))
        q0[0] = hyR * (1 + 0.01 * rng.standard_normal(n))
        q0[1:4] = q0[0] * rng.uniform(-5, 5, (3, n))
        q0[4] = hyRT * (1 + 0.01 * rng.standard_normal(n))
        if bot:
            q0[3, 0] = 0.0
        if top:
            q0[3, -1] = 0.0
        got = O.diff_transform(2, dt_arrays(q0, 2, hyR), 100.0, k_cell)
        ref = picard(q0, D, 100.0, 2, n, hyP=hyP, hyR=hyR, wall_bot=bot, wall_top=top)
        assert np.allclose(got[:5, 1:, :], np.transpose(ref[:, :, 1:], (0, 2, 1)), rtol=1e-8, atol=1e-12)
        edge = 0 if bot else n - 1
        assert np.all(got[3, :, edge] == 0.0)


# ---------------------------------------------------------------------------
# Riemann
# ---------------------------------------------------------------------------
def phys_flux(q, direction):
    nm = 1 + direction
    un = q[nm] / q[0]
    f = un * q
    f[nm] += C0 * q[4] ** GAMMA
    return f


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_riemann_consistency(direction):
    O = oracle(nz=8)
    rng = np.random.default_rng(5)
    for _ in range(50):
        r = rng.uniform(0.5, 1.3)
        q = np.array([r, r * rng.uniform(-30, 30), r * rng.uniform(-30, 30), r * rng.uniform(-30, 30),
                      r * rng.uniform(280, 320)])
        f = O.riemann(direction, q, q, kface=8)
        e = phys_flux(q, direction)
        if direction == 2:    # face nz subtracts hyPressureGLL(nz-1, ngll-1) (S:1629)
            e[3] -= O.background(5).reshape(8, 3)[-1, -1]
        assert np.allclose(f, e, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_riemann_upwinds_shear_wave(direction):
    """L and R differing only in a tangential momentum: the flux of that
    momentum is the upwind side's physical flux (sign of u_n, strict '>')."""
    O = oracle()
    nm = 1 + direction
    tang = [l for l in (1, 2, 3) if l != nm]
    base = np.array([1.1, 0.0, 0.0, 0.0, 1.1 * 300.0])
    for un, side in [(12.0, "L"), (-12.0, "R"), (0.0, "R")]:
        L = base.copy()
        L[nm] = 1.1 * un
        R = L.copy()
        L[tang[0]] = 1.1 * 7.0
        R[tang[0]] = 1.1 * -4.0
        f = O.riemann(direction, L, R, kface=0)
        up = L if side == "L" else R
        e = phys_flux(up, direction)
        assert abs(f[tang[0]] - e[tang[0]]) <= 1e-12 * max(1.0, abs(e[tang[0]])), (direction, un)


def test_riemann_y_averages_conservatives():
    """The y flux averages conservative variables (S:1210-1214): with unequal
    densities the characteristic speeds, hence the mass flux, differ from the
    x flux of the rotated problem."""
    O = oracle()
    L = np.array([1.0, 5.0, 10.0, 0.0, 300.0])
    R = np.array([1.2, 6.0, 12.0, 0.0, 372.0])
    fy = O.riemann(1, L, R)
    Lx = L[[0, 2, 1, 3, 4]]
    Rx = R[[0, 2, 1, 3, 4]]
    fx = O.riemann(0, Lx, Rx)
    assert abs(fy[0] - fx[0]) > 1e-6


# ---------------------------------------------------------------------------
# whole sweeps and steps
# ---------------------------------------------------------------------------
def bubble_spec(**kw):
    return pyoracle.make_spec([dict(shape=1, x0=10000.0, y0=10000.0, z0=2000.0, xr=2000.0, yr=2000.0, zr=2000.0,
                                    amp=2.0, pwr=1.0)], **kw)


@pytest.mark.parametrize("o,n", ORDERS)
def test_hydrostatic_rest_state_stays_at_rest(o, n):
    O = oracle(o, n, nx=6, ny=5, nz=12, bc=(0, 0, 1))
    st = O.new_state()
    dt = O.compute_time_step(st, 0.8)
    for _ in range(4):
        O.time_step(st, dt)
    # round-off only: (r*t)^gamma with t = rt/r differs from the background
    # pressure by an ulp at some GLL points
    assert np.abs(O.interior(st)).max() <= 1e-11


@pytest.mark.parametrize("bc", [(0, 0, 1), (0, 0, 0)])
def test_mass_conservation(bc):
    O = oracle(nx=20, ny=16, nz=12, bc=bc, xlen=20000.0, ylen=20000.0)
    st = O.init_state(bubble_spec(noise_amp=1e-3, noise_seed=4, u0=6.0, v0=-2.0))
    dt = O.compute_time_step(st, 0.8)
    m0 = O.interior(st)[0].sum()
    for _ in range(6):
        O.time_step(st, dt)
    I = O.interior(st)
    assert abs(I[0].sum() - m0) <= 1e-14 * np.abs(I[0]).sum()
    assert np.abs(I).max() > 0


def test_x_wall_is_not_mass_conserving():
    """Reference property: an x/y wall zeroes the halo normal momentum
    (S:645-661) but not the reconstructed edge value (only z does, S:1413-1416,
    2051-2054), so wind into an x wall carries mass through it."""
    O = oracle(nx=20, ny=16, nz=12, bc=(1, 0, 1))
    # (with a uniform wind the edge states at the two walls are equal and the
    # wall fluxes cancel; the noise makes them differ)
    st = O.init_state(bubble_spec(noise_amp=1e-3, noise_seed=4, u0=6.0))
    dt = O.compute_time_step(st, 0.8)
    m0 = O.interior(st)[0].sum()
    for _ in range(3):
        O.time_step(st, dt)
    I = O.interior(st)
    assert abs(I[0].sum() - m0) > 1e-8 * np.abs(I[0]).sum()


def test_dimension_switch_alternates():
    O = oracle(nx=8, ny=8, nz=8)
    assert O.dim_switch == 1
    st = O.init_state(bubble_spec())
    dt = O.compute_time_step(st, 0.8)
    O.time_step(st, dt)
    assert O.dim_switch == 0
    O.time_step(st, dt)
    assert O.dim_switch == 1


@pytest.mark.parametrize("direction", [0, 1])
def test_entropy_wave_translates_with_third_order_error(direction):
    """Uniform velocity and pressure with a sinusoidal density: the exact
    solution of one split sweep is a translation.  Error at a fixed time
    must fall at >= 2.5 order as the grid and dt are refined together."""
    L, u0, A = 20000.0, 20.0, 0.01
    T = 250.0
    errs = []
    for n in (16, 32, 64):
        if direction == 0:
            O = oracle(5, 3, nx=n, ny=1, nz=2, bc=(0, 0, 1), xlen=L, ylen=L)
        else:
            O = oracle(5, 3, nx=1, ny=n, nz=2, bc=(0, 0, 1), xlen=L, ylen=L)
        hs = O.hs
        h = L / n
        rbar = O.background(0)[hs:hs + 2]
        st = O.new_state()
        edges = np.arange(n + 1) * h

        def avg_sin(shift):
            ph = 2 * np.pi * (edges - shift) / L
            return -(np.cos(ph[1:]) - np.cos(ph[:-1])) * L / (2 * np.pi * h)

        I = O.interior(st)
        for k in range(2):
            dens = A * rbar[k] * avg_sin(0.0)
            if direction == 0:
                I[0, k, 0, :] = dens
                I[1, k, 0, :] = u0 * (rbar[k] + dens)
            else:
                I[0, k, :, 0] = dens
                I[2, k, :, 0] = u0 * (rbar[k] + dens)
        dt0 = O.compute_time_step(st, 0.5)
        nsteps = int(math.ceil(T / dt0))
        dt = T / nsteps
        for _ in range(nsteps):
            tend = O.tendencies_dir(st, dt, direction)
            O.interior(st)[...] += dt * tend
        I = O.interior(st)
        e = 0.0
        for k in range(2):
            exact = A * rbar[k] * avg_sin(u0 * T)
            got = I[0, k, 0, :] if direction == 0 else I[0, k, :, 0]
            e = max(e, np.abs(got - exact).max() / (A * rbar[k]))
        errs.append(e)
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert rates[-1] >= 2.5, (errs, rates)


def test_time_step_closed_form_at_rest():
    O = oracle(nx=10, ny=8, nz=12, zlen=10000.0)
    st = O.new_state()
    dt = O.compute_time_step(st, 0.8)
    hs = O.hs
    r = O.background(0)[hs:hs + 12]
    rt = O.background(3)[hs:hs + 12]
    p = C0 * (rt) ** GAMMA
    cs = np.sqrt(GAMMA * p / r)
    expect = (0.8 * min(O.dx, O.dy, O.dz) / cs).min()
    assert abs(dt - expect) <= 1e-13 * expect
    # cached: a very different state does not change it (S:166)
    st2 = O.init_state(bubble_spec(u0=100.0))
    assert O.compute_time_step(st2, 0.8) == dt


def test_initial_data_uniform_theta_matches_background():
    """theta' = a everywhere: (rho theta)' = a * hyDensCells(k) exactly in
    exact arithmetic (the z quadrature of rho_h is the background average)."""
    O = oracle(nx=6, ny=5, nz=7)
    a = 1.5
    spec = pyoracle.make_spec([dict(shape=0, x0=0.0, y0=0.0, z0=0.0, xr=1e300, yr=1e300, zr=1e300, amp=a)])
    st = O.init_state(spec)
    hs = O.hs
    rbar = O.background(0)[hs:hs + 7]
    rt = O.interior(st)[4]
    assert np.allclose(rt, a * rbar[:, None, None], rtol=1e-13, atol=0)
    assert np.all(O.interior(st)[:4] == 0)


def test_noise_generator_is_uniform_and_seeded():
    v = np.array([pyoracle.uniform_pm1(11, i) for i in range(20000)])
    assert v.min() >= -1 and v.max() < 1
    assert abs(v.mean()) < 0.02 and abs(v.var() - 1 / 3) < 0.01
    w = np.array([pyoracle.uniform_pm1(12, i) for i in range(100)])
    assert not np.array_equal(v[:100], w)


def test_weno9_ader5_stability_depends_on_cfl():
    """A property of the reference scheme, recorded so that the WENO9/ADER5
    workload is run where it is stable: with ngll = 5 the thermal bubble
    blows up within a few steps at the driver's CFL 0.8 and stays bounded at
    0.5 (48x40x32, 40 steps)."""
    from helpers import bubble_dicts

    def run(cfl):
        O = pyoracle.Oracle(48, 40, 32, ord=9, ngll=5, nthreads=8)
        st = O.init_state(pyoracle.make_spec(bubble_dicts()))
        dt = O.compute_time_step(st, cfl)
        O.dim_switch = 1
        for s in range(40):
            O.time_step(st, dt)
            I = O.interior(st)
            if not np.isfinite(I).all() or np.abs(I[3]).max() > 50:
                return s + 1
        return None

    assert run(0.8) is not None
    assert run(0.5) is None
