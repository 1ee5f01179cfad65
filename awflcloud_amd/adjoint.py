"""Gradients of final-time objectives of the FP64 dycore run with respect to
its initial state, through torch.autograd.

``run(state0, dycore, dt, nsteps, every)`` is differentiable: its forward is
``awfl_forward_ckpt`` (nsteps fused steps, a checkpoint every ``every``
steps) and its backward ``awfl_adjoint_ckpt`` (the discrete adjoint,
recomputing each checkpoint segment).  dt is held fixed.  Objectives and the
map from a control (e.g. an initial theta' field) to the initial state are
ordinary torch code around it.
"""
import torch


class DycoreRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state0, dycore, dt, nsteps, every):
        ds0 = dycore.dim_switch
        st = state0.detach().clone().contiguous()
        ckpt = dycore.forward_ckpt(st, dt, nsteps, every)
        ctx.dycore, ctx.dt, ctx.nsteps, ctx.every, ctx.ds0, ctx.ckpt = dycore, dt, nsteps, every, ds0, ckpt
        return st

    @staticmethod
    def backward(ctx, grad):
        adj = grad.detach().clone().contiguous()
        ctx.dycore.adjoint_ckpt(ctx.ckpt, ctx.dt, ctx.nsteps, ctx.every, ctx.ds0, adj)
        return adj, None, None, None, None


def run(state0, dycore, dt, nsteps, every=10):
    """Final state after nsteps steps from state0 (differentiable in state0).
    The dycore's dimSwitch is reset to 1 first, as for a fresh run."""
    dycore.dim_switch = 1
    return DycoreRun.apply(state0, dycore, dt, nsteps, every)


def background(dycore):
    """(rho_bar, (rho theta)_bar) cell averages of the interior levels, shaped
    (nz, 1, 1) for broadcasting over (nz, ny, nx)"""
    hs, nz = dycore.hs, dycore.nz
    r = torch.tensor(dycore.background(0)[hs:hs + nz], dtype=torch.float64, device="cuda")
    rt = torch.tensor(dycore.background(3)[hs:hs + nz], dtype=torch.float64, device="cuda")
    return r[:, None, None], rt[:, None, None]


def with_theta(dycore, base, theta):
    """Initial state whose rho*theta' is base's plus rho_bar * theta (theta: a
    per-cell potential-temperature perturbation, shape (nz, ny, nx))."""
    r, _ = background(dycore)
    hs = dycore.hs
    pad = torch.zeros_like(base)
    pad[4, hs:hs + dycore.nz, hs:hs + dycore.ny, hs:hs + dycore.nx] = r * theta
    return base + pad


def theta_prime(dycore, state):
    """theta - theta_bar on the interior: ((rt_bar + rt') / (r_bar + r')) - rt_bar / r_bar"""
    r, rt = background(dycore)
    s = dycore.interior(state)
    return (rt + s[4]) / (r + s[0]) - rt / r


def box_mask(dycore, box, xlen=20000.0, ylen=20000.0, zlen=10000.0):
    """cells whose centres lie in box = (x0, x1, y0, y1, z0, z1) [m]"""
    x = (torch.arange(dycore.nx, dtype=torch.float64, device="cuda") + 0.5) * (xlen / dycore.nx)
    y = (torch.arange(dycore.ny, dtype=torch.float64, device="cuda") + 0.5) * (ylen / dycore.ny)
    z = (torch.arange(dycore.nz, dtype=torch.float64, device="cuda") + 0.5) * (zlen / dycore.nz)
    mx = (x >= box[0]) & (x <= box[1])
    my = (y >= box[2]) & (y <= box[3])
    mz = (z >= box[4]) & (z <= box[5])
    return (mz[:, None, None] & my[None, :, None] & mx[None, None, :]).to(torch.float64)


def theta_box_objective(dycore, state, mask):
    """J = integral of theta'^2 over the masked cells [K^2 m^3]"""
    vol = dycore.dx * dycore.dy * dycore.dz
    return (theta_prime(dycore, state) ** 2 * mask).sum() * vol
