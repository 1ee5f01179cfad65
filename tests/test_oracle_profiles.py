"""The oracle's hydrostatic backgrounds (Profiles.h:8-47 as initState uses
them, S:370-425), pinned by what the mathematics fixes:

  * constant N: theta = theta0 exp(N^2 z / g) (the definition of a constant
    Brunt-Vaisala frequency, N^2 = g dln(theta)/dz);
  * both families: hydrostatic balance dp/dz = -rho g, checked with
    eighth-order central differences of the oracle's own p(z) against its
    rho(z) (a wrong exponent, sign or constant in the Exner function breaks it
    by orders of magnitude more than the difference error);
  * rho = P0 pi^(cp/Rd - 1) / (Rd theta) with the Exner function
    pi = 1 - g^2 (theta - theta0) / (cp N^2 theta theta0) (the integral of
    dpi/dz = -g / (cp theta)), evaluated here independently of the oracle's
    path through C0 and the equation of state; at z = 0, rho = P0 / (Rd theta0);
  * the background arrays of a data_spec 4 (igw) context equal these point
    values at the GLL points and their GLL quadratures;
  * the igw background at rest (no bubble, no wind) stays at rest."""
import math

import numpy as np
import pytest

import pyoracle

GRAV, CP, RD, P0 = 9.81, 1004.0, 287.0, 1.0e5
T0, BVF = 300.0, 0.01
ZS = np.linspace(-600.0, 10600.0, 29)


def theta_exact(z):
    return T0 * math.exp(BVF * BVF * z / GRAV)


def rho_exact(z):
    th = theta_exact(z)
    exner = 1.0 - GRAV * GRAV * (th - T0) / (CP * BVF * BVF * th * T0)
    return P0 * exner ** (CP / RD - 1.0) / (RD * th)


def test_bvf_theta_is_exponential():
    for z in ZS:
        th, _, _ = pyoracle.profile(1, z)
        assert abs(th - theta_exact(z)) <= 4e-16 * th, z
    # the definition of the constant frequency, by differences of the oracle
    h = 1.0
    for z in (0.0, 2500.0, 9000.0):
        d = (pyoracle.profile(1, z + h)[0] - pyoracle.profile(1, z - h)[0]) / (2 * h)
        n2 = GRAV * d / pyoracle.profile(1, z)[0]
        assert abs(n2 - BVF * BVF) <= 1e-9 * BVF * BVF


@pytest.mark.parametrize("family", [0, 1])
def test_hydrostatic_balance(family):
    c = [4.0 / 5, -1.0 / 5, 4.0 / 105, -1.0 / 280]      # eighth-order first derivative
    h = 4.0
    worst = 0.0
    for z in ZS:
        dp = sum(ck * (pyoracle.profile(family, z + (k + 1) * h)[2] - pyoracle.profile(family, z - (k + 1) * h)[2])
                 for k, ck in enumerate(c)) / h
        rho = pyoracle.profile(family, z)[1]
        worst = max(worst, abs(dp + rho * GRAV) / (rho * GRAV))
    assert worst <= 1e-9, worst


def test_bvf_density_closed_form():
    for z in ZS:
        _, rh, ph = pyoracle.profile(1, z)
        assert abs(rh - rho_exact(z)) <= 5e-15 * rh, z
    _, rh0, ph0 = pyoracle.profile(1, 0.0)
    assert abs(rh0 - P0 / (RD * T0)) <= 2e-15 * rh0
    assert abs(ph0 - P0) <= 2e-15 * P0
    # constant theta is the N -> 0 limit: same surface state
    _, rc0, pc0 = pyoracle.profile(0, 0.0)
    assert abs(rc0 - rh0) <= 2e-15 * rh0 and abs(pc0 - ph0) <= 2e-15 * P0


def test_igw_background_arrays():
    nz = 16
    O = pyoracle.Oracle(4, 4, nz, zlen=10000.0, data_spec=4)
    hs, ngll, dz = O.hs, O.ngll, 10000.0 / nz
    pts_o, wts_o = O.table(5), O.table(6)
    pts_g = O.table(7)
    rg, pg, tg, rtg = (O.background(w).reshape(nz, ngll) for w in (4, 5, 6, 7))
    rc, tc, rtc = O.background(0), O.background(2), O.background(3)
    for k in range(nz):
        for kk in range(ngll):
            z = (k + 0.5) * dz + pts_g[kk] * dz
            assert abs(tg[k, kk] - theta_exact(z)) <= 4e-16 * tg[k, kk]
            assert abs(rg[k, kk] - rho_exact(z)) <= 5e-15 * rg[k, kk]
            assert abs(rtg[k, kk] - rg[k, kk] * tg[k, kk]) <= 2e-16 * rtg[k, kk]
            assert abs(pg[k, kk] - pyoracle.profile(1, z)[2]) <= 2e-16 * pg[k, kk]
    for k in range(nz + 2 * hs):
        zs = [(k - hs + 0.5) * dz + p * dz for p in pts_o]
        assert abs(tc[k] - sum(w * theta_exact(z) for w, z in zip(wts_o, zs))) <= 1e-15 * tc[k]
        assert abs(rc[k] - sum(w * rho_exact(z) for w, z in zip(wts_o, zs))) <= 1e-13 * rc[k]
        assert abs(rtc[k] - sum(w * rho_exact(z) * theta_exact(z) for w, z in zip(wts_o, zs))) <= 1e-13 * rtc[k]


@pytest.mark.parametrize("o,n", [(3, 2), (5, 3)])
def test_igw_background_at_rest_stays_at_rest(o, n):
    O = pyoracle.Oracle(6, 5, 12, ord=o, ngll=n, bc=(0, 0, 1), data_spec=4)
    st = O.new_state()
    dt = O.compute_time_step(st, 0.8)
    for _ in range(4):
        O.time_step(st, dt)
    I = O.interior(st)
    rtbar = O.background(3)[O.hs:O.hs + 12]
    # round-off of the background's own pressure at the GLL points
    assert np.abs(I[1:4]).max() <= 1e-11
    assert np.abs(I[4]).max() <= 1e-13 * rtbar.max()
