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
