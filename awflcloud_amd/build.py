"""Build the HIP library (gfx950) and the CPU oracle in-tree.

    python -m awflcloud_amd.build            # tables + libawfl_dycore.so + oracle
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libawfl_dycore.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 arithmetic stays off in every HIP build of this package.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_tables():
    run([sys.executable, os.path.join(CSRC, "gen_tables.py"), os.path.join(CSRC, "awfl_tables.h")])


def build_hip():
    src = os.path.join(CSRC, "awfl_dycore.hip")
    tmp = LIB + ".tmp"
    run([HIPCC] + HIP_FLAGS + [src, "-o", tmp])
    os.replace(tmp, LIB)


def build_oracle():
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_tables()
    build_hip()
    build_oracle()


if __name__ == "__main__":
    build_all()
