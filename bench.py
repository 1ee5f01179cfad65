#!/usr/bin/env python3
"""Throughput of the fused FP64 ADER-WENO step on one MI355X.

Workload: rising thermal bubble, 256 x 256 x 128 cells over 20 x 20 x 10 km,
periodic x/y, solid walls in z, WENO order 5 with ADER time order 3, cosine
bubble (2 K, r = 2 km at 10, 10, 2 km) plus seeded +-1e-3 K noise, dt from the
CFL condition (0.8) computed once and cached.  One step = three split sweeps
(x,y,z then z,y,x alternately), each one fused kernel.

Prints ONE JSON line.  `value` is cell-updates/s = nx*ny*nz*steps / wall of
awfl_time_step calls with the state resident in HBM.  `roofline` is for the
slowest (dominant) sweep kernel, timed by HIP events that the library records
on its stream around every sweep launch inside the timed region: its FP64
vector rate by the reference's operations per cell against 78.6 TFLOP/s, or
its algorithmic HBM bytes (read + write 5 FP64 variables per cell = 80 B/cell)
against 8 TB/s, whichever fraction is larger (the sweeps are VALU-bound); the
other figure and the executed-instruction issue fraction sit beside it.
`cpu_baseline` is the FP64 CPU oracle (a C restatement of the reference step)
timed on this host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP64_VALU_PEAK_TFS = 78.6    # MI355X FP64 vector
BYTES_PER_CELL_SWEEP = 80.0  # read + write 5 doubles
# FP64 flops per cell of each WENO5/ADER3 sweep (x, y, z) and of a whole step
# (three sweeps and three tendency applications) at 256x256x128, counted by the
# op-counting build of the oracle, which follows the reference's statements
# (oracle/opcount.cpp; add/sub, mul, div, sqrt, pow = 1 each):
# `oracle/opcount 256 256 128 5 3`.
FLOPS_PER_CELL_SWEEP = {0: 2488.3, 1: 2480.3, 2: 2553.6}
FLOPS_PER_CELL_STEP = 7552.2

# VALU wave-instructions x 64 per cell of each FP64 sweep kernel (x, y, z) at
# 256x256x128, from rocprofv3 SQ_INSTS_VALU (profiles/sweeps_valu_counters.md);
# MI355X: 256 CUs x 4 SIMDs at 2.4 GHz, one FP64 wave64 VALU instruction per
# 4 cycles per SIMD
VALU_INSTS_PER_CELL = {0: 1505.0, 1: 1363.0, 2: 1436.0}
N_SIMD, CLOCK_HZ = 1024, 2.4e9

NX, NY, NZ = 256, 256, 128
# HBM bytes (read + written) per launch of each FP64 sweep kernel at
# 256x256x128, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes calibrated by a
# copy kernel of known size (profiles/sweeps_fp64_256x256x128_hbm_traffic.md)
TRAFFIC_BYTES = {0: 682.0e6, 1: 751.1e6, 2: 732.8e6}
# FP32 perturbation-form comparison: 512^2 x 256 bubble + noise + band-limited
# theta' field, 200 steps FP64 vs FP32 from the same initial state
NX4, NY4, NZ4, STEPS4 = 512, 512, 256, 200


def cpu_baseline(nthreads, target_s=12.0, shape=(64, 64, 32), max_steps=100, sized=True):
    """Oracle step rate on the bubble (64x64x32 unless given) with `nthreads`
    threads: one step to size the sample, then up to `max_steps` steps for
    about `target_s` seconds (sized=False: just `max_steps` timed steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    nx, ny, nz = shape
    O = pyoracle.Oracle(nx, ny, nz, bc=(0, 0, 1), nthreads=nthreads)
    spec = pyoracle.make_spec([dict(shape=1, x0=10000.0, y0=10000.0, z0=2000.0, xr=2000.0, yr=2000.0,
                                    zr=2000.0, amp=2.0, pwr=1.0)], noise_amp=1e-3, noise_seed=1)
    st = O.init_state(spec)
    dt = O.compute_time_step(st, 0.8)
    if sized:
        t1 = O.run_steps(st, dt, 1)
        steps = max(1, min(max_steps, int(target_s / max(t1, 1e-3))))
    else:
        steps = max_steps
    t = O.run_steps(st, dt, steps)
    return nx * ny * nz * steps / t, steps, t


def fp32_leg(steps, warmup, chunk=20):
    """FP64 and FP32 runs of the 512x512x256 workload, each through its fused
    C-ABI time step (awfl_time_step / awfl_time_step_f32); throughput of each
    and the FP32 errors of theta', w and rho' against the FP64 run.  The two
    runs advance in alternating chunks of `chunk` steps, each chunk timed on
    its own, so that both see the same clocks and temperature."""
    import torch
    from awflcloud_amd import Dycore, bubble_spec, band_limited_modes
    from awflcloud_amd.dycore import AWFL_FP32_PERT
    D = Dycore(NX4, NY4, NZ4, bc=(0, 0, 1), ord=5, ngll=3, precision=AWFL_FP32_PERT)
    spec = bubble_spec([dict(shape=1, x0=10000.0, y0=10000.0, z0=2000.0, xr=2000.0, yr=2000.0, zr=2000.0,
                             amp=2.0, pwr=1.0)], noise_amp=1e-3, noise_seed=1,
                       modes=band_limited_modes(seed=4, n=64, kmin=4, kmax=16, rms=0.2))
    st0 = D.init_state(spec)
    dt = D.compute_time_step(st0, 0.8)

    # the two runs keep their own dimSwitch sequence (both start at 1)
    ds = {"64": 1, "32": 1}

    def run64(st, n):
        D.dim_switch = ds["64"]
        for _ in range(n):
            D.time_step(st, dt)
        ds["64"] = D.dim_switch

    def run32(s, c, n):
        D.dim_switch = ds["32"]
        for _ in range(n):
            D.time_step_f32(s, c, dt)
        ds["32"] = D.dim_switch

    def timed(fn, *a):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(*a)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    scratch = st0.clone()
    run64(scratch, warmup)
    s, c = D.to_f32(st0)
    run32(s, c, warmup)
    del scratch
    ds["64"] = ds["32"] = 1
    st64 = st0.clone()
    s, c = D.to_f32(st0)
    w64 = w32 = 0.0
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        w64 += timed(run64, st64, n)
        w32 += timed(run32, s, c, n)
        done += n
    st32 = D.from_f32(s, c)
    del s, c

    hs, nz = D.hs, D.nz
    r = torch.tensor(D.background(0)[hs:hs + nz], dtype=torch.float64, device="cuda")[:, None, None]
    rt = torch.tensor(D.background(3)[hs:hs + nz], dtype=torch.float64, device="cuda")[:, None, None]
    a, b = D.interior(st32), D.interior(st64)
    errs = {}
    pairs = (("theta_K", (rt + a[4]) / (r + a[0]), (rt + b[4]) / (r + b[0])),
             ("w_m_s", a[3] / (r + a[0]), b[3] / (r + b[0])),
             ("rho_kg_m3", a[0], b[0]))
    for name, x, y in pairs:
        d = x - y
        errs[name] = {"l2": float(d.pow(2).mean().sqrt()), "linf": float(d.abs().max())}
    ncell = NX4 * NY4 * NZ4
    v64, v32 = ncell * steps / w64, ncell * steps / w32
    return {"workload": "thermal bubble + 1e-3 K noise + band-limited theta' (|k| 4-16, 0.2 K rms, 64 modes) "
                        "%dx%dx%d, WENO5/ADER3, %d steps" % (NX4, NY4, NZ4, steps),
            "fp64_cell_updates_per_s": v64, "fp32_cell_updates_per_s": v32, "fp32_over_fp64": v32 / v64,
            "fp64_ms_per_step": w64 / steps * 1e3, "fp32_ms_per_step": w32 / steps * 1e3,
            "errors_vs_fp64": errs, "error_norms": "L2 = rms over cells, Linf = max; FP32 minus FP64",
            "finite": bool(torch.isfinite(st32).all())}


def weno9_leg(steps=50, warmup=2, cfl=0.5):
    """WENO9/ADER5 FP64, 192x192x96, two opposite seeded bubbles: fused steps
    through awfl_time_step, with the per-sweep mean kernel time.  CFL 0.5:
    the scheme with ngll = 5 is unstable at the driver's 0.8 (the oracle and
    this path alike blow up within a few steps there)."""
    import torch
    from awflcloud_amd import Dycore
    from awflcloud_amd.workloads import two_bubbles_spec
    nx, ny, nz = 192, 192, 96
    D = Dycore(nx, ny, nz, bc=(0, 0, 1), ord=9, ngll=5)
    st = D.init_state(two_bubbles_spec(seed=3))
    dt = D.compute_time_step(st, cfl)
    D.dim_switch = 1
    for _ in range(warmup):
        D.time_step(st, dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        D.time_step(st, dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = torch.empty_like(st)
    kms = {}
    for d in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            D.sweep(st, out, dt, d)
        e1.record()
        torch.cuda.synchronize()
        kms["xyz"[d]] = e0.elapsed_time(e1) / 5
    return {"workload": "two opposite bubbles (+2 K r 1.5 km, -1.5 K r 2.5 km, seed 3) %dx%dx%d, WENO9/ADER5, "
                        "FP64, %d steps, CFL %.1f" % (nx, ny, nz, steps, cfl),
            "cell_updates_per_s": nx * ny * nz * steps / wall, "ms_per_step": wall / steps * 1e3,
            "sweep_ms": kms, "finite": bool(torch.isfinite(st).all())}


def tracer_leg(steps=20, warmup=2):
    """The driver's three tracers (driver.cpp:27-29: uniform, theta replica,
    block; initTracers S:505-560, all positive) transported with the state on
    the 256x256x128 thermal: fused steps with tracers (awfl_time_step_tracers)
    against the same steps without, plus the uniform tracer's mixing-ratio
    error and the tracer mass drift (periodic x/y, walls in z: conserved)."""
    import torch
    from awflcloud_amd import Dycore
    from awflcloud_amd.workloads import thermal_spec
    nx, ny, nz = 256, 256, 128
    D = Dycore(nx, ny, nz, bc=(0, 0, 1), ord=5, ngll=3)
    st = D.init_state(thermal_spec())
    dt = D.compute_time_step(st, 0.8)
    D.set_tracers([1, 1, 1])
    tr = D.init_tracers()
    ref = st.clone()

    def run(fn, n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    D.dim_switch = 1
    run(lambda: D.time_step(ref, dt), warmup)
    w0 = run(lambda: D.time_step(ref, dt), steps)
    m0 = D.tracer_interior(tr).sum(dim=(1, 2, 3))
    D.dim_switch = 1
    run(lambda: D.time_step_tracers(st, tr, dt), warmup)
    w1 = run(lambda: D.time_step_tracers(st, tr, dt), steps)
    m1 = D.tracer_interior(tr).sum(dim=(1, 2, 3))
    hs = D.hs
    rb = torch.tensor(D.background(0)[hs:hs + nz], dtype=torch.float64, device="cuda")
    q = D.tracer_interior(tr)[0] / (D.interior(st)[0] + rb[:, None, None])
    ncell = nx * ny * nz
    return {"workload": "thermal bubble %dx%dx%d, WENO5/ADER3, FP64, 3 positive tracers (uniform, theta replica, "
                        "block), %d steps" % (nx, ny, nz, steps),
            "cell_updates_per_s": ncell * steps / w1, "ms_per_step": w1 / steps * 1e3,
            "ms_per_step_without_tracers": w0 / steps * 1e3, "with_over_without": w1 / w0,
            "uniform_ratio_max_err": float((q - 1.0).abs().max()),
            "tracer_mass_drift_rel": float(((m1 - m0).abs() / m0.abs()).max()),
            "tracer_min": [float(v) for v in D.tracer_interior(tr).amin(dim=(1, 2, 3))]}


def adjoint_leg(reps=3):
    """Discrete adjoint on the 128x128x64 bubble: 60 FP64 steps with a
    checkpoint every 10 (awfl_forward_ckpt) against the adjoint run
    (awfl_adjoint_ckpt: recompute each segment, then 180 split-sweep
    adjoints); and one split-sweep adjoint against one forward sweep."""
    import torch
    from awflcloud_amd import Dycore, bubble_spec
    nx, ny, nz, nsteps, every = 128, 128, 64, 60, 10
    D = Dycore(nx, ny, nz, bc=(0, 0, 1), ord=5, ngll=3)
    base = D.init_state(bubble_spec([dict(shape=1, x0=10000.0, y0=10000.0, z0=2000.0, xr=2000.0, yr=2000.0,
                                          zr=2000.0, amp=2.0, pwr=1.0)], noise_amp=1e-3, noise_seed=1))
    dt = D.compute_time_step(base, 0.8)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    fwd, bwd = [], []
    for r in range(reps + 1):
        st = base.clone()
        D.dim_switch = 1
        box = {}
        fwd.append(timed(lambda: box.update(ckpt=D.forward_ckpt(st, dt, nsteps, every))))
        adj = torch.randn_like(st)
        bwd.append(timed(lambda: D.adjoint_ckpt(box["ckpt"], dt, nsteps, every, 1, adj)))
    fwd, bwd = min(fwd[1:]), min(bwd[1:])
    sweep = {}
    out = torch.empty_like(base)
    y = torch.randn_like(base)
    yin = torch.empty_like(base)
    for d in range(3):
        D.sweep(base, out, dt, d)
        D.sweep_adjoint(base, y, dt, d, out=yin)
        tf = timed(lambda: [D.sweep(base, out, dt, d) for _ in range(10)]) / 10
        ta = timed(lambda: [D.sweep_adjoint(base, y, dt, d, out=yin) for _ in range(10)]) / 10
        sweep["xyz"[d]] = {"forward_ms": tf * 1e3, "adjoint_ms": ta * 1e3, "adjoint_over_forward": ta / tf}
    # the adjoint of one whole step's three split sweeps against their forward
    sf = sum(v["forward_ms"] for v in sweep.values())
    sa = sum(v["adjoint_ms"] for v in sweep.values())
    return {"workload": "thermal bubble 128x128x64, WENO5/ADER3, FP64, %d steps, checkpoint every %d"
                        % (nsteps, every),
            "forward_ms": fwd * 1e3, "backward_ms": bwd * 1e3, "backward_over_forward": bwd / fwd,
            "note": "backward = segment recomputation + split-sweep adjoints",
            "sweep": sweep,
            "step_sweeps": {"forward_ms": sf, "adjoint_ms": sa, "adjoint_over_forward": sa / sf}}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-fp32", action="store_true", help="skip the 512x512x256 FP32-vs-FP64 leg")
    ap.add_argument("--fp32-steps", type=int, default=STEPS4)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing leg")
    ap.add_argument("--no-weno9", action="store_true", help="skip the WENO9/ADER5 leg")
    ap.add_argument("--no-tracers", action="store_true", help="skip the tracer-transport leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU runs are not part of it "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from awflcloud_amd import Dycore, bubble_spec

    D = Dycore(NX, NY, NZ, bc=(0, 0, 1), ord=5, ngll=3)
    spec = bubble_spec([dict(shape=1, x0=10000.0, y0=10000.0, z0=2000.0, xr=2000.0, yr=2000.0, zr=2000.0,
                             amp=2.0, pwr=1.0)], noise_amp=1e-3, noise_seed=1)
    state = D.init_state(spec)
    dt = D.compute_time_step(state, 0.8)
    mass0 = D.state_sums(state)[0]

    # Temporal_ader::timeStep through the C ABI's fused step (awfl_time_step);
    # the library brackets every sweep launch with HIP events on its stream
    for _ in range(args.warmup):
        D.time_step(state, dt)
    torch.cuda.synchronize()
    D.sweep_timing(3 * args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        D.time_step(state, dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sweep_ms, sweep_n = D.sweep_times()
    D.sweep_timing(0)
    assert sweep_n == [args.steps] * 3, sweep_n

    ncell = NX * NY * NZ
    value = ncell * args.steps / wall
    kms = {d: sweep_ms[d] for d in range(3)}
    dom = max(kms, key=kms.get)
    achieved = BYTES_PER_CELL_SWEEP * ncell / (kms[dom] * 1e-3) / 1e9
    flops = FLOPS_PER_CELL_SWEEP[dom] * ncell / (kms[dom] * 1e-3) / 1e12
    # executed VALU wave-instructions per SIMD-cycle against the FP64 limit
    # of one per 4 cycles, from the counted instructions per cell
    issue = VALU_INSTS_PER_CELL[dom] * ncell / 64 / (kms[dom] * 1e-3) / (N_SIMD * CLOCK_HZ / 4)
    hbm_frac, valu_frac = achieved / HBM_PEAK_GBS, flops / FP64_VALU_PEAK_TFS
    step_flops = FLOPS_PER_CELL_STEP * value / 1e12
    mass1 = D.state_sums(state)[0]
    finite = bool(torch.isfinite(state).all())

    out = {
        "metric": "cell-updates/s per ADER step, FP64 vs FP32 at stated error; adjoint:forward time",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "thermal bubble 256x256x128, WENO5/ADER3, FP64, periodic x/y, wall z",
                   "nx": NX, "ny": NY, "nz": NZ, "ord": 5, "ngll": 3, "dt": dt, "precision": "fp64"},
        # the binding roof is the larger fraction: the FP64 vector ALU (by the
        # reference's operations per cell) rather than HBM
        "roofline": ({"bound": "valu", "achieved": flops, "peak": FP64_VALU_PEAK_TFS, "unit": "TFLOP/s",
                      "frac": valu_frac} if valu_frac >= hbm_frac else
                     {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": hbm_frac}) | {
            "traffic": TRAFFIC_BYTES[dom],
            "hbm": {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_frac,
                    "bytes_per_cell": BYTES_PER_CELL_SWEEP},
            "valu_issue_frac": issue,
            "kernel": ["sweep_x", "sweep_y", "sweep_z"][dom],
            "kernel_ms": {["x", "y", "z"][d]: kms[d] for d in kms},
            "timing": "HIP events recorded by the library on its stream around every sweep launch of "
                      "awfl_time_step in the timed region"},
        "valu": {"achieved_tflops": flops, "peak_tflops": FP64_VALU_PEAK_TFS, "frac": flops / FP64_VALU_PEAK_TFS,
                 "flops_per_cell_sweep": FLOPS_PER_CELL_SWEEP[dom],
                 "step_tflops": step_flops, "step_frac": step_flops / FP64_VALU_PEAK_TFS,
                 "flops_per_cell_step": FLOPS_PER_CELL_STEP,
                 "flop_count": "reference operations per cell (oracle/opcount.cpp), FMA counted as 2"},
        "checks": {"finite": finite, "mass_drift_rel": abs(mass1 - mass0) / max(abs(mass0), 1.0)},
    }
    if not args.no_fp32:
        torch.cuda.empty_cache()
        out["fp32_pert"] = fp32_leg(args.fp32_steps, 2)
    if not args.no_adjoint:
        torch.cuda.empty_cache()
        out["adjoint"] = adjoint_leg()
    if not args.no_weno9:
        torch.cuda.empty_cache()
        out["weno9"] = weno9_leg()
    if not args.no_tracers:
        torch.cuda.empty_cache()
        out["tracers"] = tracer_leg()
    if not args.no_cpu:
        # the headline workload itself, one step on one core (~16 s)
        v1, n1, t1 = cpu_baseline(1, shape=(NX, NY, NZ), max_steps=1, sized=False)
        out["cpu_baseline"] = {"value": v1, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                               "sample": "CPU oracle (C restatement, -O3, no FMA), the 256x256x128 bubble "
                                         "above, WENO5/ADER3, its first step (%.1f s), 1 thread" % t1}
        ncores = len(os.sched_getaffinity(0))
        nthr = min(ncores, 16)
        vs, ns, ts = cpu_baseline(1, args.cpu_seconds / 2)
        out["cpu_baseline"]["small_1_core"] = {"value": vs, "threads": 1,
                                               "sample": "64x64x32 bubble, %d steps in %.1f s" % (ns, ts)}
        vN, nN, tN = cpu_baseline(nthr, args.cpu_seconds / 2)
        out["cpu_baseline"]["all_cores"] = {"value": vN, "threads": nthr, "host_cpus": ncores,
                                            "sample": "64x64x32 bubble, %d steps in %.1f s" % (nN, tN)}
        # the headline workload on all cores: two steps of 256x256x128
        v2, n2, t2 = cpu_baseline(nthr, 1e9, shape=(NX, NY, NZ), max_steps=2)
        out["cpu_baseline"]["all_cores_256x256x128"] = {
            "value": v2, "threads": nthr, "ms_per_step": t2 / n2 * 1e3,
            "sample": "the headline bubble at 256x256x128, %d step(s) after one untimed step" % n2}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
