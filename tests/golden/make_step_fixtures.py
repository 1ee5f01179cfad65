#!/usr/bin/env python3
"""Write the oracle's step fixtures (inputs and outputs) into tests/golden/:

  oracle_steps_12x10x8.npz   two 12 x 10 x 8 boxes (periodic x/y with wall z,
                             and walls everywhere), WENO5/ADER3, a bubble
                             with winds and seeded noise: the initial state,
                             dt, the tendency of every direction, the states
                             after steps 1, 2 and 3 (dimSwitch starting at 1);
                             three tracers (uniform, smooth, step) with their
                             tendencies and the tracers after 2 steps.
  oracle_c1_norms.json       the 64 x 64 x 32 thermal bubble (WENO5/ADER3,
                             noise 1e-3 K seed 1): per-variable L2 and Linf
                             norms and sums of the interior at steps 1, 10, 100.

The oracle is the CPU restatement of the reference's step (oracle/); these
fixtures pin it against later changes and give the GPU tests outputs to meet
without running the oracle.  Run from the repository root after
`make -C oracle`:
    python tests/golden/make_step_fixtures.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as P  # noqa: E402

BUBBLE = dict(shape=1, x0=10000.0, y0=10000.0, z0=2000.0, xr=2000.0, yr=2000.0, zr=2000.0, amp=2.0, pwr=1.0)


def tracer_fields(O, st):
    rho = O.density(st)
    x = (np.arange(O.nx) + 0.5) * O.dx
    y = (np.arange(O.ny) + 0.5) * O.dy
    z = (np.arange(O.nz) + 0.5) * O.dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    blob = np.exp(-(((X - 8000) / 2500) ** 2 + ((Y - 9000) / 2500) ** 2 + ((Z - 4000) / 1500) ** 2))
    blob[blob < 0.05] = 0.0
    tr = np.zeros(O.tracer_shape())
    O.tracer_interior(tr)[...] = np.stack([0.25 * rho, blob * rho, (blob > 0.3) * rho])
    return tr


def small_case(bc):
    n = (12, 10, 8)
    O = P.Oracle(*n, bc=bc)
    st0 = O.init_state(P.make_spec([BUBBLE], noise_amp=1e-3, noise_seed=5, u0=5.0, v0=-3.0))
    dt = O.compute_time_step(st0, 0.8)
    out = {"state0": st0.copy(), "dt": np.array(dt)}
    for d in range(3):
        out["tend%d" % d] = O.tendencies_dir(st0.copy(), dt, d)
    st = st0.copy()
    O.dim_switch = 1
    for s in range(3):
        O.time_step(st, dt)
        out["state%d" % (s + 1)] = st.copy()
    # tracers
    O.set_tracers([0, 1, 1])
    tr0 = tracer_fields(O, st0)
    out["tracers0"] = tr0.copy()
    for d in range(3):
        _, tt = O.tendencies_dir_tr(st0.copy(), tr0.copy(), dt, d)
        out["ttend%d" % d] = tt
    st, tr = st0.copy(), tr0.copy()
    O.dim_switch = 1
    for s in range(2):
        O.time_step_tr(st, tr, dt)
    out["tracers2"] = tr
    out["tracers2_state"] = st
    return out


def c1_norms():
    O = P.Oracle(64, 64, 32, bc=(0, 0, 1))
    st = O.init_state(P.make_spec([BUBBLE], noise_amp=1e-3, noise_seed=1))
    dt = O.compute_time_step(st, 0.8)
    O.dim_switch = 1
    rec = {"dt": dt, "steps": {}}
    done = 0
    for target in (1, 10, 100):
        for _ in range(target - done):
            O.time_step(st, dt)
        done = target
        I = O.interior(st)
        rec["steps"][str(target)] = {
            "l2": [float(np.sqrt(np.mean(I[l] ** 2))) for l in range(5)],
            "linf": [float(np.abs(I[l]).max()) for l in range(5)],
            "sum": [float(I[l].sum()) for l in range(5)]}
    return rec


def main():
    out = {}
    for tag, bc in (("per", (0, 0, 1)), ("wall", (1, 1, 1))):
        for k, v in small_case(bc).items():
            out["%s_%s" % (tag, k)] = v
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "oracle_steps_12x10x8.npz"), **out)
    with open(os.path.join(ROOT, "tests", "golden", "oracle_c1_norms.json"), "w") as f:
        json.dump({"workload": "thermal bubble 64x64x32 (cosine, 2 K, r 2 km at (10, 10, 2) km), noise 1e-3 K "
                               "seed 1, WENO5/ADER3, periodic x/y, wall z, CFL 0.8, dimSwitch starting at 1",
                   **c1_norms()}, f, indent=1)
    print("wrote oracle_steps_12x10x8.npz, oracle_c1_norms.json")


if __name__ == "__main__":
    main()
