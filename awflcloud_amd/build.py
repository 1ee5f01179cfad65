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
# -fno-signed-zeros lets the compiler drop the addition of a zero-initialised
# accumulator's first term (x + 0.0 is not x only for x = -0.0); no result
# changes except the sign of an exact zero, which nothing here depends on.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fno-signed-zeros",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# translation units of libawfl_dycore.so: FP64 sweeps + ABI, FP32 perturbation form, adjoint,
# alternative time integrators, tracers
HIP_UNITS = ["awfl_dycore.hip", "awfl_pert.hip", "awfl_adjoint.hip", "awfl_splitting.hip", "awfl_tracers.hip"]


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_tables():
    run([sys.executable, os.path.join(CSRC, "gen_tables.py"), os.path.join(CSRC, "awfl_tables.h")])


def build_hip():
    """Compile the units in parallel, then link one shared library."""
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for u in HIP_UNITS:
        obj = os.path.join(objdir, os.path.splitext(u)[0] + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, u), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((u, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [u for u, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    tmp = LIB + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, LIB)


# Parity builds of the FP64 (5, 3) sweeps (test libraries: the C ABI of
# awfl_dycore.hip only), each with some of the production arithmetic
# switched back toward the reference's, to attribute the difference between
# the HIP step and the CPU oracle (tests/test_gpu_parity_builds.py).  Masks:
# AWFL_PAR_DIV 1, _POW 2, _LIBPOW 4, _BRIDGE 8, _ORDER 16 (awfl_fastmath.hpp).
PARITY_BUILDS = {
    "div": (1, False),          # IEEE division and sqrt for rcp() and the sound speed
    "pow": (2, False),          # general pow_gamma in place of the near-background series
    "contract": (0, False),     # no FMA contraction
    "bridge": (8, False),       # WENO bridge polynomial by its three statements
    "fast_all": (1 | 2 | 8, True),
    "ref_libm": (16 | 1 | 4, True),   # reference operation order, IEEE /, sqrt, device pow()
    "ref_pg": (16 | 1, True),         # reference operation order, IEEE /, sqrt, pow_gamma
}


def parity_lib(name):
    return os.path.join(PKG, "libawfl_parity_%s.so" % name)


def build_parity():
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    for name, (mask, nocontract) in PARITY_BUILDS.items():
        obj = os.path.join(objdir, "parity_%s.o" % name)
        cmd = [HIPCC] + HIP_FLAGS + ["-DAWFL_LEAN", "-DAWFL_PARITY=%d" % mask] + \
            (["-ffp-contract=off"] if nocontract or name == "contract" else []) + \
            ["-c", os.path.join(CSRC, "awfl_dycore.hip"), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((name, obj, subprocess.Popen(cmd)))
    failed = [n for n, _, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("hipcc failed for parity builds " + ", ".join(failed))
    for name, obj, _ in procs:
        tmp = parity_lib(name) + ".tmp"
        run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", obj, "-o", tmp])
        os.replace(tmp, parity_lib(name))


def build_driver():
    """C++ driver over the header-only Spatial/Temporal adapter (include/awfl)."""
    src = os.path.join(PKG, "driver", "driver.cpp")
    exe = os.path.join(PKG, "awfl_driver")
    run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
         "-I" + os.path.join(ROOT, "include"), src, "-o", exe, "-L" + PKG, "-lawfl_dycore",
         "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])


def build_adapter_test():
    """Native test of the C++ adapter (run by tests/test_gpu_adapter.py)."""
    src = os.path.join(ROOT, "tests", "native", "adapter_test.cpp")
    exe = os.path.join(ROOT, "tests", "native", "adapter_test")
    run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
         "-I" + os.path.join(ROOT, "include"), src, "-o", exe, "-L" + PKG, "-lawfl_dycore",
         "-Wl,-rpath," + PKG, "-Wl,-rpath,$ORIGIN/../../awflcloud_amd",
         "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    # the same adapter with tracer transport compiled in
    src = os.path.join(ROOT, "tests", "native", "adapter_tracers_test.cpp")
    exe = os.path.join(ROOT, "tests", "native", "adapter_tracers_test")
    run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-DAWFL_TRACERS", "-I/opt/rocm/include",
         "-I" + os.path.join(ROOT, "include"), src, "-o", exe, "-L" + PKG, "-lawfl_dycore",
         "-Wl,-rpath," + PKG, "-Wl,-rpath,$ORIGIN/../../awflcloud_amd",
         "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])


def build_probe():
    """VALU issue-rate probe (FMA chains in FP32/FP64, rcp_f32), run on the GPU by hand."""
    run([HIPCC] + [f for f in HIP_FLAGS if f != "-fPIC"] + [os.path.join(CSRC, "issue_probe.hip"), "-o",
                                                            os.path.join(PKG, "issue_probe")])


def build_oracle():
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_tables()
    build_hip()
    build_parity()
    build_driver()
    build_adapter_test()
    build_probe()
    build_oracle()


if __name__ == "__main__":
    build_all()
