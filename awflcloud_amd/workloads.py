"""Synthetic initial data of the benchmark workloads, as awfl_init_spec values.

All on the constant-theta (300 K) hydrostatic background of a 20 x 20 x 10 km
box; the bubble shape is the cosine ellipsoid of Profiles.h:60-73.

  thermal_spec       warm bubble 2 K, r = 2 km at (10, 10, 2) km, plus seeded
                     uniform +-1e-3 K noise (64x64x32, 256x256x128 runs)
  two_bubbles_spec   +2 K (r = 1.5 km) and -1.5 K (r = 2.5 km) bubbles at
                     seeded random centres in the lower half of the box
                     (WENO9/ADER5 run)
  band_limited_spec  thermal_spec plus a seeded band-limited theta' field,
                     wavenumbers 4..16, 0.2 K rms (FP32 comparison run)

and the reference's own four data specs (reference_case / reference_spec):
thermal, collision, straka and igw, as Spatial::initState builds them.
"""
import numpy as np

from .dycore import band_limited_modes, bubble_spec


def bubble(amp=2.0, r=2000.0, x0=10000.0, y0=10000.0, z0=2000.0):
    return dict(shape=1, x0=x0, y0=y0, z0=z0, xr=r, yr=r, zr=r, amp=amp, pwr=1.0)


def thermal_spec(noise_amp=1e-3, noise_seed=1, **kw):
    return bubble_spec([bubble()], noise_amp=noise_amp, noise_seed=noise_seed, **kw)


def two_bubbles(seed=3, xlen=20000.0, ylen=20000.0, zlen=10000.0):
    """(+2 K, r 1.5 km) and (-1.5 K, r 2.5 km), centres uniform over x, y and
    the lower half in z"""
    rng = np.random.default_rng(seed)
    out = []
    for amp, r in ((2.0, 1500.0), (-1.5, 2500.0)):
        out.append(bubble(amp=amp, r=r, x0=float(rng.uniform(0.0, xlen)), y0=float(rng.uniform(0.0, ylen)),
                          z0=float(rng.uniform(0.0, 0.5 * zlen))))
    return out


def two_bubbles_spec(seed=3):
    return bubble_spec(two_bubbles(seed))


def band_limited_spec(seed=4, nmodes=64, rms=0.2):
    return thermal_spec(modes=band_limited_modes(seed=seed, n=nmodes, kmin=4, kmax=16, rms=rms))


REFERENCE_DATA_SPECS = {"thermal": 1, "collision": 2, "straka": 3, "igw": 4}


def reference_case(name, xlen=20000.0, ylen=20000.0, zlen=10000.0, sim2d=False):
    """Parameters of one of the reference's initial-data specs (S:428-485):

      thermal    linear cone, +2 K, radius 2 km, at (xlen/2, ylen/2, 2 km)
                 (S:444-450; Profiles.h:49-58)
      collision  linear cones +20 K at 2 km and -20 K at 8 km height, radius
                 2 km, at (xlen/2, ylen/2) (S:451-459)
      straka     cosine ellipsoid -15 K at (xlen/2, ylen/2, 3 km), radii
                 4 / 2 / 2 km (S:460-466; Profiles.h:60-73)
      igw        constant-N background (data_spec 4), theta' = 0.01 sin(pi
                 z / zlen) / sqrt(1 + xn^2 + yn^2) about (xlen/3, ylen/3) with
                 a 5 km horizontal radius, and 20 m/s winds in x and (3-D) y
                 (S:467-481; Profiles.h:76-82); in 2-D every quadrature
                 point sits at y = ylen/3 (S:469)

    Returns dict(data_spec, bubbles, u0, v0, y2d); the bubble dicts use the
    shape codes of awfl_init_spec (0 linear cone, 1 cosine, 2 igw)."""
    def ell(shape, x0, y0, z0, xr, yr, zr, amp):
        return dict(shape=shape, x0=x0, y0=y0, z0=z0, xr=xr, yr=yr, zr=zr, amp=amp, pwr=1.0)
    out = dict(data_spec=REFERENCE_DATA_SPECS[name], u0=0.0, v0=0.0, y2d=0.0)
    if name == "thermal":
        out["bubbles"] = [ell(0, xlen / 2, ylen / 2, 2000.0, 2000.0, 2000.0, 2000.0, 2.0)]
    elif name == "collision":
        out["bubbles"] = [ell(0, xlen / 2, ylen / 2, 2000.0, 2000.0, 2000.0, 2000.0, 20.0),
                          ell(0, xlen / 2, ylen / 2, 8000.0, 2000.0, 2000.0, 2000.0, -20.0)]
    elif name == "straka":
        out["bubbles"] = [ell(1, xlen / 2, ylen / 2, 3000.0, 4000.0, 2000.0, 2000.0, -15.0)]
    else:
        out["bubbles"] = [ell(2, xlen / 3, ylen / 3, 0.0, 5000.0, 5000.0, 1.0, 0.01)]
        out["u0"] = 20.0
        out["v0"] = 0.0 if sim2d else 20.0
        out["y2d"] = ylen / 3 if sim2d else 0.0
    return out


def reference_spec(name, xlen=20000.0, ylen=20000.0, zlen=10000.0, sim2d=False):
    """(data_spec, awfl_init_spec) of reference_case(name, ...); data_spec goes
    into the context's configuration (it selects the background)."""
    c = reference_case(name, xlen, ylen, zlen, sim2d)
    return c["data_spec"], bubble_spec(c["bubbles"], u0=c["u0"], v0=c["v0"], y2d=c["y2d"])
