"""The C-ABI library loads and exports every entry point declared in
include/awfl_dycore.h; argument errors are rejected before any device call.
(No compute call is made: this runs without a GPU.)"""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "awfl_dycore.h")
LIB = os.path.join(ROOT, "awflcloud_amd", "libawfl_dycore.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(awfl_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("awfl_create", "awfl_destroy", "awfl_compute_time_step", "awfl_compute_tendencies",
                 "awfl_apply_tendencies", "awfl_time_step", "awfl_sweep", "awfl_num_split", "awfl_layout"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    L = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_invalid_configurations_are_rejected():
    import sys
    sys.path.insert(0, ROOT)
    from awflcloud_amd.dycore import AwflConfig, lib
    L = lib()
    h = C.c_void_p()
    base = dict(nx=8, ny=8, nz=8, xlen=1.0, ylen=1.0, zlen=1.0, bc_x=0, bc_y=0, bc_z=1, weno_scalars=1,
                weno_winds=1, ord=5, ngll=3, data_spec=1, precision=0)
    for change, code in [(dict(nx=0), -1), (dict(ord=4), -4), (dict(ngll=7), -4), (dict(bc_x=3), -1),
                         (dict(xlen=0.0), -1), (dict(data_spec=9), -1), (dict(precision=5), -1),
                         (dict(ord=9, ngll=5, precision=1), -4)]:
        cfg = AwflConfig(**{**base, **change})
        rc = L.awfl_create(C.byref(cfg), C.byref(h))
        assert rc == code, (change, rc)
        assert L.awfl_last_error(None)
    assert L.awfl_create(None, C.byref(h)) == -1
    assert L.awfl_num_split(None) == -1
