"""Tangent-linear and discrete-adjoint sweeps (awfl_sweep_tangent /
awfl_sweep_adjoint / awfl_time_step_adjoint / awfl_*_ckpt) on the GPU.

Checks, all through the C ABI:
  * tangent sweep value == awfl_sweep (<= 1e-13 relative);
  * dot-product identity <y, J x> == <J^T y, x> for each sweep direction and
    boundary configuration (<= 1e-11 relative);
  * tangent == fourth-order central differences of awfl_sweep (<= 1e-7);
  * time-step adjoint == reverse composition of the sweep adjoints;
  * the checkpointed run's gradient of J = integral of theta'^2 over a
    4 x 4 x 2 km box centred at (10, 10, 6) km after 60 steps, with respect
    to the initial theta' field, against central finite differences
    (eps = 1e-6) along 8 seeded random directions (<= 1e-3 relative: see
    that test), and the dot-product identity of the whole 60-step run.
"""
import numpy as np
import pytest

from helpers import bubble_dicts

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _interior_dot(D, a, b):
    return float((D.interior(a) * D.interior(b)).sum())


def _rand_like_state(D, gen, scale=(1e-3, 1.0, 1.0, 1.0, 0.5)):
    x = D.new_state()
    hs = D.hs
    for l, s in enumerate(scale):
        x[l, hs:hs + D.nz, hs:hs + D.ny, hs:hs + D.nx] = s * torch.randn(
            (D.nz, D.ny, D.nx), generator=gen, dtype=torch.float64, device="cuda")
    return x


def _setup(nx, ny, nz, bc=(0, 0, 1), ord=5, ngll=3, u0=4.0, v0=-2.0, seed=7):
    from awflcloud_amd import Dycore, bubble_spec
    D = Dycore(nx, ny, nz, bc=bc, ord=ord, ngll=ngll)
    st = D.init_state(bubble_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=seed, u0=u0, v0=v0,
                                  y2d=1.0 if ny == 1 else 0.0))
    dt = D.compute_time_step(st, 0.8)
    # two steps so that the state carries winds, sound waves and a deformed bubble
    D.dim_switch = 1
    D.time_step(st, dt)
    D.time_step(st, dt)
    return D, st, dt


CASES = [
    dict(n=(24, 20, 16), bc=(0, 0, 1), ord=5, ngll=3),
    dict(n=(24, 20, 16), bc=(1, 1, 1), ord=5, ngll=3),
    dict(n=(24, 20, 16), bc=(0, 0, 0), ord=5, ngll=3),
    dict(n=(20, 16, 12), bc=(0, 1, 1), ord=3, ngll=2),
    dict(n=(32, 1, 16), bc=(0, 0, 1), ord=5, ngll=3),
    # x lines longer than one work-group (threads take two cells each)
    dict(n=(300, 6, 5), bc=(1, 0, 1), ord=5, ngll=3),
    # lines too long for LDS: the multi-kernel adjoint (x, then z)
    dict(n=(704, 4, 3), bc=(0, 1, 1), ord=5, ngll=3),
    dict(n=(4, 3, 704), bc=(0, 0, 1), ord=5, ngll=3),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-bc%s-o%d" % ("x".join(map(str, c["n"])),
                                                                        "".join(map(str, c["bc"])), c["ord"]))
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_sweep_dot_product(case, direction):
    _need_gpu()
    nx, ny, nz = case["n"]
    print("case", case, "dir", direction, flush=True)
    D, st, dt = _setup(nx, ny, nz, bc=case["bc"], ord=case["ord"], ngll=case["ngll"])
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1234 + direction)
    x = _rand_like_state(D, gen)
    y = _rand_like_state(D, gen)
    out, dout = D.sweep_tangent(st, x, dt, direction)
    ref = D.new_state()
    D.sweep(st, ref, dt, direction)
    vi, ri = D.interior(out), D.interior(ref)
    verr = float((vi - ri).abs().max() / ri.abs().max())
    xb = D.sweep_adjoint(st, y, dt, direction)
    lhs = _interior_dot(D, y, dout)
    rhs = _interior_dot(D, xb, x)
    scale = float((D.interior(y) * D.interior(dout)).abs().sum())
    print("value err %.2e  <y,Jx> %.15e  <J^T y,x> %.15e  rel %.2e" % (verr, lhs, rhs, abs(lhs - rhs) / scale),
          flush=True)
    assert verr <= 1e-13
    assert abs(lhs - rhs) <= 1e-11 * scale
    # adjoint halo is zero
    hs = D.hs
    mask = torch.ones_like(xb, dtype=torch.bool)
    mask[:, hs:hs + nz, hs:hs + ny, hs:hs + nx] = False
    assert float(xb[mask].abs().max()) == 0.0


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_sweep_tangent_matches_differences(direction):
    """On a state whose every stencil has total variation far above the WENO
    epsilon (sqrt(1e-20)): there the weights are smooth functions of the
    stencil over the difference step.  (On smooth small-amplitude data the
    weights change regime within a step of 1e-4 and differences do not
    approximate the derivative.)"""
    _need_gpu()
    D, st, dt = _setup(24, 20, 16, bc=(0, 1, 1))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(99 + direction)
    st = st + _rand_like_state(D, gen, scale=(2e-3, 3.0, 3.0, 3.0, 1.0))
    x = _rand_like_state(D, gen)
    _, dout = D.sweep_tangent(st, x, dt, direction)
    h = 1e-5
    outs = []
    for m in (1.0, -1.0, 2.0, -2.0):
        o = D.new_state()
        D.sweep(st + m * h * x, o, dt, direction)
        outs.append(D.interior(o))
    fd = (8.0 * (outs[0] - outs[1]) - (outs[2] - outs[3])) / (12.0 * h)
    err = float((fd - D.interior(dout)).abs().max() / D.interior(dout).abs().max())
    print("dir %d tangent vs fd rel err %.2e" % (direction, err), flush=True)
    assert err <= 1e-7


def test_time_step_adjoint_is_reverse_of_sweeps():
    _need_gpu()
    D, st, dt = _setup(24, 20, 16)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    y = _rand_like_state(D, gen)
    for ds, dirs in ((1, (0, 1, 2)), (0, (2, 1, 0))):
        # split inputs of one step
        ins = [st.clone()]
        for d in dirs[:-1]:
            o = D.new_state()
            D.sweep(ins[-1], o, dt, d)
            ins.append(o)
        a = y.clone()
        for d, s_in in zip(reversed(dirs), reversed(ins)):
            a = D.sweep_adjoint(s_in, a, dt, d)
        b = y.clone()
        D.time_step_adjoint(st, b, dt, ds)
        err = float((D.interior(a) - D.interior(b)).abs().max() / D.interior(a).abs().max())
        print("dimSwitch %d: step adjoint vs sweeps %.2e" % (ds, err), flush=True)
        assert err <= 1e-14


def test_adjoint_output_buffer_and_halo():
    """An adjoint sweep into a caller's buffer full of NaN equals the one into
    a fresh buffer, halo included (zero); the input adjoint's halo is not
    read; the time-step adjoint, which zeroes the halo once after its chain
    of sweeps, leaves it zero too."""
    _need_gpu()
    D, st, dt = _setup(24, 20, 16)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(6)
    y = _rand_like_state(D, gen)
    hs = D.hs
    halo = torch.ones(D.state_shape, dtype=torch.bool, device="cuda")
    halo[:, hs:hs + D.nz, hs:hs + D.ny, hs:hs + D.nx] = False
    y7 = y.clone()
    y7[halo] = 7.0
    for d in range(3):
        ref = D.sweep_adjoint(st, y, dt, d)
        buf = torch.full_like(st, float("nan"))
        got = D.sweep_adjoint(st, y7, dt, d, out=buf)
        assert got is buf
        assert torch.equal(got, ref)
        assert float(ref[halo].abs().max()) == 0.0
    for ds in (1, 0):
        a, b = y.clone(), y7.clone()
        D.time_step_adjoint(st, a, dt, ds)
        D.time_step_adjoint(st, b, dt, ds)
        assert torch.equal(a, b)
        assert float(a[halo].abs().max()) == 0.0


def test_checkpointed_adjoint_equals_stepwise():
    """awfl_adjoint_ckpt (segments of 4) == the step-by-step adjoint"""
    _need_gpu()
    D, st, dt = _setup(24, 20, 16)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(8)
    y = _rand_like_state(D, gen)
    nsteps, every = 10, 4
    D.dim_switch = 1
    states = []
    s = st.clone()
    for n in range(nsteps):
        states.append((s.clone(), D.dim_switch))
        D.time_step(s, dt)
    a = y.clone()
    for s_before, ds in reversed(states):
        D.time_step_adjoint(s_before, a, dt, ds)
    ds_after = D.dim_switch
    D.dim_switch = 1
    s2 = st.clone()
    ckpt = D.forward_ckpt(s2, dt, nsteps, every)
    # the last segment runs as split sweeps that keep their inputs: the same
    # state bit for bit, and the dimSwitch the fused steps leave
    assert torch.equal(s2, s)
    assert D.dim_switch == ds_after
    b = y.clone()
    D.adjoint_ckpt(ckpt, dt, nsteps, every, 1, b)
    err = float((D.interior(a) - D.interior(b)).abs().max() / D.interior(a).abs().max())
    print("checkpointed vs stepwise %.2e" % err, flush=True)
    assert err <= 1e-14
    # a second adjoint run on the same checkpoints recomputes every segment
    # (the first one overwrote the kept inputs): bit for bit the same
    c = y.clone()
    D.adjoint_ckpt(ckpt, dt, nsteps, every, 1, c)
    assert torch.equal(b, c)
    # one segment (nsteps < every): the kept inputs serve the first adjoint,
    # the second recomputes them, a step adjoint in between rewrites them:
    # all three bit for bit the same
    D.dim_switch = 1
    s3 = st.clone()
    ckpt1 = D.forward_ckpt(s3, dt, 3, every)
    d1, d2, d3 = y.clone(), y.clone(), y.clone()
    D.adjoint_ckpt(ckpt1, dt, 3, every, 1, d1)
    D.adjoint_ckpt(ckpt1, dt, 3, every, 1, d2)
    D.time_step_adjoint(states[0][0], y.clone(), dt, 1)
    D.adjoint_ckpt(ckpt1, dt, 3, every, 1, d3)
    ref = y.clone()
    for s_before, ds in reversed(states[:3]):
        D.time_step_adjoint(s_before, ref, dt, ds)
    assert torch.equal(d1, d2) and torch.equal(d1, d3)
    err1 = float((D.interior(ref) - D.interior(d1)).abs().max() / D.interior(ref).abs().max())
    assert err1 <= 1e-14


def _c5_problem():
    from awflcloud_amd import Dycore, bubble_spec
    from awflcloud_amd import adjoint as A
    D = Dycore(128, 128, 64)
    base = D.init_state(bubble_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=1))
    dt = D.compute_time_step(base, 0.8)
    mask = A.box_mask(D, (8000.0, 12000.0, 8000.0, 12000.0, 5000.0, 7000.0))
    return D, base, dt, mask


def _branch_records(D, base, dt, nsteps, theta=None):
    """the upwind branch of every face of every split of the run from base
    (+ rho_bar theta), recorded step by step; returns (records, final state)"""
    from awflcloud_amd import adjoint as A
    recs = [D.new_branch_record() for _ in range(nsteps)]
    s = (A.with_theta(D, base, theta) if theta is not None else base.clone()).contiguous()
    D.dim_switch = 1
    for n in range(nsteps):
        D.time_step_branches(s, dt, recs[n], 1)
    return recs, s


def _frozen_run(D, base, dt, recs, theta):
    """the run from base + rho_bar theta with every face's upwind branch
    imposed from recs (a smooth map near the recorded run)"""
    from awflcloud_amd import adjoint as A
    s = A.with_theta(D, base, theta).contiguous()
    D.dim_switch = 1
    for r in recs:
        D.time_step_branches(s, dt, r, 2)
    return s


def _flips(a, b):
    return sum(int((x != y).sum()) for x, y in zip(a, b))


def test_gradient_of_box_objective_against_differences(figures):
    """128x128x64 bubble, 60 steps, checkpoints every 10: dJ/dtheta'_0 by the
    adjoint against central differences (eps = 1e-6 K) along 8 seeded random
    directions.

    The forward is only piecewise smooth: the Riemann flux takes waves 1-3
    from the left or right state on the sign of the averaged normal velocity
    (strict >, S:863, 1221, 1602), and with zero initial winds many faces sit
    near u_n = 0, where a perturbation of eps flips the choice.  The adjoint
    differentiates the branch the forward took.  So the differences are taken
    of the forward with every face's branch frozen as the base run took it
    (awfl_time_step_branches): that map is smooth and its derivative at the
    base is what the adjoint computes; gate 1e-6.  The plain forward's
    differences and the number of faces whose branch differs between its +eps
    and -eps runs are reported beside them."""
    _need_gpu()
    from awflcloud_amd import adjoint as A
    D, base, dt, mask = _c5_problem()
    nsteps, every = 60, 10
    theta = torch.zeros((D.nz, D.ny, D.nx), dtype=torch.float64, device="cuda", requires_grad=True)
    final = A.run(A.with_theta(D, base, theta), D, dt, nsteps, every)
    J = A.theta_box_objective(D, final, mask)
    J.backward()
    grad = theta.grad.detach()
    recs, s_rec = _branch_records(D, base, dt, nsteps)
    # recording changes nothing, and replaying the recorded branches from the
    # base state reproduces the run (the same statements in other kernels)
    s_rep = _frozen_run(D, base, dt, recs, torch.zeros_like(grad))
    e_rec = float((s_rec - final.detach()).abs().max() / final.detach().abs().max())
    e_rep = float((s_rep - s_rec).abs().max() / s_rec.abs().max())
    faces = sum(int(r.numel()) for r in recs)
    figures("J = %.10e, |grad| = %.3e, %d face records; recording run vs forward %.1e, replay vs recording %.1e"
            % (float(J), float(grad.norm()), faces, e_rec, e_rep))
    assert e_rec <= 1e-14 and e_rep <= 1e-14

    def J_of(st):
        return float(A.theta_box_objective(D, st, mask))

    gen = torch.Generator(device="cuda")
    gen.manual_seed(2024)
    eps = 1e-6
    worst = 0.0
    for k in range(8):
        d = torch.randn((D.nz, D.ny, D.nx), generator=gen, dtype=torch.float64, device="cuda")
        ad = float((grad * d).sum())
        with torch.no_grad():
            fd_frozen = (J_of(_frozen_run(D, base, dt, recs, eps * d))
                         - J_of(_frozen_run(D, base, dt, recs, -eps * d))) / (2 * eps)
            rp, sp = _branch_records(D, base, dt, nsteps, eps * d)
            rm, sm = _branch_records(D, base, dt, nsteps, -eps * d)
            fd_plain = (J_of(sp) - J_of(sm)) / (2 * eps)
        rel = abs(fd_frozen - ad) / abs(ad)
        rel_plain = abs(fd_plain - ad) / abs(ad)
        worst = max(worst, rel)
        figures("direction %d: adjoint %.12e, frozen-branch fd %.12e (rel %.2e), plain fd rel %.2e, "
                "faces flipped +eps vs -eps %d, +eps vs base %d, -eps vs base %d"
                % (k, ad, fd_frozen, rel, rel_plain, _flips(rp, rm), _flips(rp, recs), _flips(rm, recs)))
        del rp, rm
    assert worst <= 1e-6


def test_whole_run_dot_product():
    """<y, L x> == <L^T y, x> for the 60-step run, L its tangent-linear map
    (composition of tangent sweeps) and L^T the checkpointed adjoint."""
    _need_gpu()
    D, base, dt, _ = _c5_problem()
    nsteps, every = 60, 10
    gen = torch.Generator(device="cuda")
    gen.manual_seed(77)
    x = _rand_like_state(D, gen)
    y = _rand_like_state(D, gen)
    s, dx = base.clone(), x.clone()
    for n in range(nsteps):
        dirs = (0, 1, 2) if n % 2 == 0 else (2, 1, 0)
        for d in dirs:
            s, dx = D.sweep_tangent(s, dx, dt, d)
    D.dim_switch = 1
    s2 = base.clone()
    ckpt = D.forward_ckpt(s2, dt, nsteps, every)
    ferr = float((D.interior(s) - D.interior(s2)).abs().max() / D.interior(s2).abs().max())
    a = y.clone()
    D.adjoint_ckpt(ckpt, dt, nsteps, every, 1, a)
    lhs = _interior_dot(D, y, dx)
    rhs = _interior_dot(D, a, x)
    scale = float((D.interior(y) * D.interior(dx)).abs().sum())
    print("tangent-run state vs forward %.2e; <y,Lx> %.15e <L^T y,x> %.15e rel %.2e"
          % (ferr, lhs, rhs, abs(lhs - rhs) / scale), flush=True)
    assert ferr <= 1e-10
    assert abs(lhs - rhs) <= 1e-9 * scale


def test_difference_quotients_plateau(figures):
    """The c5 gradient along one per-cell random and one smooth direction,
    against central differences for eps = 1e-3 ... 1e-8: of the plain forward,
    which stay at 1e-7 - 1e-4 without converging (the upwind switch at u_n = 0
    makes it piecewise smooth), and of the forward with the base run's upwind
    branches frozen, which converge like eps^2 down to the rounding floor."""
    _need_gpu()
    from awflcloud_amd import adjoint as A
    import math
    D, base, dt, mask = _c5_problem()
    nsteps, every = 60, 10
    theta = torch.zeros((D.nz, D.ny, D.nx), dtype=torch.float64, device="cuda", requires_grad=True)
    A.theta_box_objective(D, A.run(A.with_theta(D, base, theta), D, dt, nsteps, every), mask).backward()
    g = theta.grad.detach()
    recs, _ = _branch_records(D, base, dt, nsteps)

    def J_plain(th):
        with torch.no_grad():
            return float(A.theta_box_objective(D, A.run(A.with_theta(D, base, th), D, dt, nsteps, every), mask))

    def J_frozen(th):
        with torch.no_grad():
            return float(A.theta_box_objective(D, _frozen_run(D, base, dt, recs, th), mask))

    gen = torch.Generator(device="cuda")
    gen.manual_seed(2024)
    d_rand = torch.randn((D.nz, D.ny, D.nx), generator=gen, dtype=torch.float64, device="cuda")
    z = (torch.arange(D.nz, device="cuda", dtype=torch.float64)[:, None, None] + 0.5) / D.nz
    y = (torch.arange(D.ny, device="cuda", dtype=torch.float64)[None, :, None] + 0.5) / D.ny
    x = (torch.arange(D.nx, device="cuda", dtype=torch.float64)[None, None, :] + 0.5) / D.nx
    d_smooth = (torch.cos(2 * math.pi * (3 * x + 2 * y) + 0.3) * torch.sin(math.pi * z)
                + 0.5 * torch.cos(2 * math.pi * (x - 4 * y) + 1.1))
    for name, d in (("random", d_rand), ("smooth", d_smooth)):
        ad = float((g * d).sum())
        frozen = {}
        for eps in (1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8):
            fd = (J_plain(eps * d) - J_plain(-eps * d)) / (2 * eps)
            ff = (J_frozen(eps * d) - J_frozen(-eps * d)) / (2 * eps)
            rel, frozen[eps] = abs(fd - ad) / abs(ad), abs(ff - ad) / abs(ad)
            figures("%s direction, eps %.0e: adjoint %.12e, plain fd rel %.2e, frozen-branch fd rel %.2e"
                    % (name, eps, ad, rel, frozen[eps]))
            assert rel <= 1e-2
        assert frozen[1e-5] <= 1e-6 and frozen[1e-6] <= 1e-6
