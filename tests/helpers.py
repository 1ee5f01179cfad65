"""Shared inputs and error norms for the tests."""
import numpy as np

XLEN, YLEN, ZLEN = 20000.0, 20000.0, 10000.0


def bubble_dicts(amp=2.0, r=2000.0, x0=10000.0, y0=10000.0, z0=2000.0, shape=1):
    return [dict(shape=shape, x0=x0, y0=y0, z0=z0, xr=r, yr=r, zr=r, amp=amp, pwr=1.0)]


def rel_err(a, b):
    """Per-variable max-norm error relative to the max norm of the expected
    field, maximised over the five variables."""
    a = np.asarray(a)
    b = np.asarray(b)
    worst = 0.0
    for l in range(a.shape[0]):
        scale = np.abs(b[l]).max()
        if scale == 0:
            scale = 1.0
        worst = max(worst, np.abs(a[l] - b[l]).max() / scale)
    return worst


def per_var_rel_err(a, b):
    out = []
    for l in range(a.shape[0]):
        scale = np.abs(b[l]).max()
        out.append(np.abs(a[l] - b[l]).max() / (scale if scale else 1.0))
    return out


def full_rel_err(a, b, rbar, rtbar):
    """Per-variable max-norm error relative to the max norm of the full
    conserved field: rho = rho_bar + rho', rho*theta = (rho*theta)_bar +
    (rho*theta)', momenta as stored.  rbar/rtbar broadcast over (k, j, i)."""
    out = []
    for l in range(5):
        ref = b[l] + (rbar if l == 0 else rtbar if l == 4 else 0.0)
        out.append(np.abs(a[l] - b[l]).max() / max(np.abs(ref).max(), 1e-300))
    return out
