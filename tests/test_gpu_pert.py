"""FP32 perturbation form (awfl_*_f32) against the FP64 path.

The FP32 state is a pair (s, c) of float32 arrays holding s - c, advanced by
FP32 sweeps on deviations from the FP64 hydrostatic background with
Kahan-compensated updates.  The FP64 side is the FP64 HIP path, which the
parity tests hold to the CPU oracle at <= 1e-10 after 100 steps; one case
also compares with the oracle directly.

Errors stated for every case (printed; FP32 minus FP64, interior cells):
  theta' [K]       theta = ((rho theta)bar + (rho theta)') / (rhobar + rho')
  w [m/s]          rho*w / rho
  rho' [kg/m^3]
  L2 = root mean square over cells, Linf = max.
Gate: theta' L2 <= 1e-4 K (after 100 steps, and for the 512x512x256 case after
200 steps).  w and rho' are gated at 1e-4 of the FP64 field's own RMS.
"""
import numpy as np
import pytest

from helpers import bubble_dicts

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

THETA_L2_MAX = 1e-4


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _dycore(nx, ny, nz, **kw):
    from awflcloud_amd import Dycore
    from awflcloud_amd.dycore import AWFL_FP32_PERT
    return Dycore(nx, ny, nz, precision=AWFL_FP32_PERT, **kw)


def _bg(D):
    hs, nz = D.hs, D.nz
    r = torch.tensor(D.background(0)[hs:hs + nz], dtype=torch.float64, device="cuda")[:, None, None]
    rt = torch.tensor(D.background(3)[hs:hs + nz], dtype=torch.float64, device="cuda")[:, None, None]
    return r, rt


def _errors(D, st32, st64):
    """theta', w, rho' errors of the FP32 run (unpacked to FP64) against FP64"""
    r, rt = _bg(D)
    a = D.interior(st32)
    b = D.interior(st64)
    out = {}
    th_a = (rt + a[4]) / (r + a[0])
    th_b = (rt + b[4]) / (r + b[0])
    w_a = a[3] / (r + a[0])
    w_b = b[3] / (r + b[0])
    for name, x, y in (("theta'", th_a, th_b), ("w", w_a, w_b), ("rho'", a[0], b[0])):
        d = (x - y).double()
        out[name] = dict(l2=float(d.pow(2).mean().sqrt()), linf=float(d.abs().max()),
                         rms=float((y - (rt / r if name == "theta'" else 0.0)).pow(2).mean().sqrt()))
    return out


def _report(tag, err):
    for k, v in err.items():
        print("%s %-6s L2 %.3e  Linf %.3e  (FP64 field rms %.3e)" % (tag, k, v["l2"], v["linf"], v["rms"]),
              flush=True)


def _gate(err):
    assert err["theta'"]["l2"] <= THETA_L2_MAX, err
    for k in ("w", "rho'"):
        assert err[k]["l2"] <= 1e-4 * max(err[k]["rms"], 1e-300), (k, err)


def _run_pair(D, st0, nsteps):
    """FP64 and FP32 runs of nsteps from the same FP64 initial state"""
    st64 = st0.clone()
    s, c = D.to_f32(st0)
    dt = D.compute_time_step(st0, 0.8)
    D.dim_switch = 1
    for _ in range(nsteps):
        D.time_step(st64, dt)
    D.dim_switch = 1
    for _ in range(nsteps):
        D.time_step_f32(s, c, dt)
    st32 = D.from_f32(s, c)
    torch.cuda.synchronize()
    return st32, st64


def test_f32_round_trip():
    _need_gpu()
    from awflcloud_amd import bubble_spec
    D = _dycore(40, 36, 20)
    st = D.init_state(bubble_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=3, u0=7.0, v0=-3.0))
    s, c = D.to_f32(st)
    back = D.from_f32(s, c)
    err = (back - st).abs()
    assert bool((err <= 2.0 ** -47 * st.abs()).all()), float(err.max())
    # s is the float32 rounding of the value
    assert torch.equal(s, st.float())


def test_f32_rest_state_is_exact():
    """Zero deviations: every flux and source of the perturbation form is
    exactly zero, so the state stays exactly zero (the FP64 form, built on
    full pressures, drifts at rounding level)."""
    _need_gpu()
    from awflcloud_amd import bubble_spec
    D = _dycore(32, 24, 16)
    st = D.init_state(bubble_spec())
    dt = D.compute_time_step(st, 0.8)
    s, c = D.to_f32(st)
    s.zero_()
    c.zero_()
    for _ in range(20):
        D.time_step_f32(s, c, dt)
    torch.cuda.synchronize()
    assert float(s.abs().max()) == 0.0 and float(c.abs().max()) == 0.0


def test_f32_sweeps_compose_time_step():
    _need_gpu()
    from awflcloud_amd import bubble_spec
    D = _dycore(48, 40, 24, bc=(0, 1, 1))
    st = D.init_state(bubble_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=5, u0=5.0))
    dt = D.compute_time_step(st, 0.8)
    s0, c0 = D.to_f32(st)
    s1, c1 = s0.clone(), c0.clone()
    D.dim_switch = 1
    D.time_step_f32(s1, c1, dt)          # x, y, z
    D.time_step_f32(s1, c1, dt)          # z, y, x
    bufs = [torch.empty_like(s0) for _ in range(4)]
    cur = (s0, c0)
    for k, d in enumerate((0, 1, 2, 2, 1, 0)):
        nxt = (bufs[2 * (k % 2)], bufs[2 * (k % 2) + 1])
        D.sweep_f32(cur[0], cur[1], nxt[0], nxt[1], dt, d)
        cur = nxt
    torch.cuda.synchronize()
    # halo cells of an output are never written: compare the interiors
    I = D.interior
    assert torch.equal(I(cur[0]), I(s1)) and torch.equal(I(cur[1]), I(c1))


@pytest.mark.parametrize("case", [
    dict(name="thermal 64x64x32", n=(64, 64, 32), kw={}, bub=dict()),
    dict(name="walls xyz", n=(48, 40, 24), kw=dict(bc=(1, 1, 1)), bub=dict(x0=7000.0)),
    dict(name="ord3 ngll2", n=(64, 48, 32), kw=dict(ord=3, ngll=2), bub=dict()),
    dict(name="wind + collision", n=(48, 48, 32), kw=dict(data_spec=2), bub=dict(amp=-2.0, z0=6000.0)),
    dict(name="2-D (ny=1)", n=(96, 1, 48), kw={}, bub=dict(y0=0.0)),
])
def test_f32_matches_fp64_after_100_steps(case):
    _need_gpu()
    from awflcloud_amd import bubble_spec
    nx, ny, nz = case["n"]
    print("case:", case["name"], flush=True)
    D = _dycore(nx, ny, nz, **case["kw"])
    u0 = 10.0 if case["name"].startswith("wind") else 0.0
    st0 = D.init_state(bubble_spec(bubble_dicts(**case["bub"]), noise_amp=1e-3, noise_seed=11, u0=u0,
                                   y2d=1.0 if ny == 1 else 0.0))
    st32, st64 = _run_pair(D, st0, 100)
    err = _errors(D, st32, st64)
    _report(case["name"], err)
    _gate(err)


def test_f32_matches_oracle():
    """The FP32 form against the CPU FP64 oracle directly (20 steps)."""
    _need_gpu()
    import pyoracle
    from awflcloud_amd import bubble_spec
    nx, ny, nz = 40, 32, 24
    D = _dycore(nx, ny, nz)
    O = pyoracle.Oracle(nx, ny, nz, nthreads=8)
    st0 = D.init_state(bubble_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=2))
    host = st0.cpu().numpy().copy()
    dt = D.compute_time_step(st0, 0.8)
    O.dim_switch = 1
    O.run_steps(host, dt, 20)
    s, c = D.to_f32(st0)
    D.dim_switch = 1
    for _ in range(20):
        D.time_step_f32(s, c, dt)
    st32 = D.from_f32(s, c)
    err = _errors(D, st32, torch.tensor(host, device="cuda"))
    _report("vs oracle", err)
    _gate(err)


def test_f32_mass_is_conserved_periodic():
    """Periodic x/y and walls in z: the flux-difference form conserves the
    total of rho' up to FP32 rounding of the updates."""
    _need_gpu()
    from awflcloud_amd import bubble_spec
    D = _dycore(64, 48, 32)
    st0 = D.init_state(bubble_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=4, u0=6.0, v0=2.0))
    dt = D.compute_time_step(st0, 0.8)
    s, c = D.to_f32(st0)
    m0 = D.state_sums(st0)[0]
    for _ in range(100):
        D.time_step_f32(s, c, dt)
    m1 = D.state_sums(D.from_f32(s, c))[0]
    r, _ = _bg(D)
    total = float(r.sum()) * D.nx * D.ny
    print("mass change %.3e of total %.3e" % (m1 - m0, total), flush=True)
    assert abs(m1 - m0) <= 1e-9 * total


def test_f32_large_band_limited():
    """512x512x256, the thermal bubble plus noise plus a seeded band-limited
    theta' field (|k| 4..16, 0.2 K rms), 200 steps FP32 vs FP64."""
    _need_gpu()
    from awflcloud_amd import bubble_spec, band_limited_modes
    D = _dycore(512, 512, 256)
    modes = band_limited_modes(seed=4, n=64, kmin=4, kmax=16, rms=0.2)
    st0 = D.init_state(bubble_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=1, modes=modes))
    st32, st64 = _run_pair(D, st0, 200)
    err = _errors(D, st32, st64)
    _report("512x512x256 200 steps", err)
    _gate(err)
