"""MI355X-native ADER-WENO dynamical-core step (HIP, gfx950).

The compute path is the C-ABI library ``libawfl_dycore.so`` built from
``csrc/``; this package only binds it (ctypes) for Python callers and tests.
"""
from .dycore import Dycore, InitSpec, lib, bubble_spec, band_limited_modes  # noqa: F401
