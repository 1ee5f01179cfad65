"""Launch each fused sweep kernel a fixed number of times on the thermal
bubble (default 256x256x128; FP64 sweeps, or with --precision the FP32
perturbation-form sweeps, the adjoint sweeps or the sweeps with the driver's
three tracers), plus one elementwise kernel of known size
(read 8N, write 8N bytes) that calibrates HBM byte counters.  Meant to run
under rocprofv3:

    rocprofv3 --kernel-trace --stats -f csv -d <dir> -- python3 -m awflcloud_amd.profile_sweeps
    rocprofv3 --pmc FETCH_SIZE -f csv -d <dir> -- python3 -m awflcloud_amd.profile_sweeps
"""
import argparse
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--ord", type=int, default=5)
    ap.add_argument("--ngll", type=int, default=3)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=512)
    ap.add_argument("--precision", choices=("fp64", "fp32", "adjoint", "tracers"), default="fp64",
                    help="fp64 sweeps, FP32 perturbation-form sweeps, FP64 adjoint sweeps, or FP64 sweeps "
                         "with the driver's three tracers")
    a = ap.parse_args()
    import torch
    from awflcloud_amd import Dycore, bubble_spec
    from awflcloud_amd.dycore import AWFL_FP32_PERT, AWFL_FP64
    prec = AWFL_FP32_PERT if a.precision == "fp32" else AWFL_FP64
    D = Dycore(a.nx, a.ny, a.nz, ord=a.ord, ngll=a.ngll, precision=prec)
    st = D.init_state(bubble_spec([dict(shape=1, x0=10000.0, y0=10000.0, z0=2000.0, xr=2000.0, yr=2000.0,
                                        zr=2000.0, amp=2.0, pwr=1.0)], noise_amp=1e-3, noise_seed=1))
    dt = D.compute_time_step(st, 0.8)
    if a.precision == "fp64":
        out = torch.empty_like(st)
        for d in range(3):
            for _ in range(a.reps):
                D.sweep(st, out, dt, d)
    elif a.precision == "tracers":
        D.set_tracers([1, 1, 1])
        tr = D.init_tracers()
        out, tro = torch.empty_like(st), torch.empty_like(tr)
        for d in range(3):
            for _ in range(a.reps):
                D.sweep_tracers(st, tr, out, tro, dt, d)
    elif a.precision == "adjoint":
        y = torch.randn_like(st)
        for d in range(3):
            for _ in range(a.reps):
                D.sweep_adjoint(st, y, dt, d)
    else:
        s, c = D.to_f32(st)
        so, co = torch.empty_like(s), torch.empty_like(c)
        for d in range(3):
            for _ in range(a.reps):
                D.sweep_f32(s, c, so, co, dt, d)
    n = a.copy_mib * 1024 * 1024 // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(a.reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print("profile_sweeps: %s %dx%dx%d ord %d ngll %d, %d launches per sweep; copy probe %d MiB read + %d MiB "
          "write" % (a.precision, a.nx, a.ny, a.nz, a.ord, a.ngll, a.reps, a.copy_mib, a.copy_mib))
    sys.stdout.flush()


if __name__ == "__main__":
    main()
