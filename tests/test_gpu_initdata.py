"""The reference's four initial-data specs (S:428-485) and the constant-N
background (S:384-391, Profiles.h:28-47) on the GPU against the oracle.

  * initState of thermal (linear cone), collision, straka and igw, 3-D and
    2-D: the device-built state against the oracle's, per variable;
  * the igw background arrays, bit for bit (both are built on the host in
    FP64 from the same formulas);
  * ten igw steps with the 20 m/s winds, fused, against the oracle with the
    gates of the other step tests."""
import numpy as np
import pytest

from helpers import full_rel_err, per_var_rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _pair(name, nx, ny, nz, bc=(0, 0, 1), nthreads=0):
    import pyoracle
    from awflcloud_amd import Dycore
    from awflcloud_amd.workloads import reference_case, reference_spec
    sim2d = ny == 1
    ds, spec = reference_spec(name, sim2d=sim2d)
    case = reference_case(name, sim2d=sim2d)
    D = Dycore(nx, ny, nz, bc=bc, data_spec=ds)
    O = pyoracle.Oracle(nx, ny, nz, bc=bc, data_spec=ds, nthreads=nthreads)
    ospec = pyoracle.make_spec(case["bubbles"], u0=case["u0"], v0=case["v0"], y2d=case["y2d"])
    return D, O, spec, ospec


CASES = [("thermal", (40, 36, 24)), ("collision", (40, 36, 24)), ("straka", (48, 40, 24)), ("igw", (40, 36, 24)),
         ("thermal", (64, 1, 32)), ("straka", (64, 1, 32)), ("igw", (64, 1, 32))]


@pytest.mark.parametrize("name,n", CASES, ids=["%s-%dx%dx%d" % (c, *n) for c, n in CASES])
def test_initial_data_matches_oracle(name, n, figures):
    _need_gpu()
    D, O, spec, ospec = _pair(name, *n)
    st = D.init_state(spec)
    got = D.interior(st).cpu().numpy()
    ref = O.interior(O.init_state(ospec))
    # the fields the spec sets are non-trivial
    assert np.abs(ref[4]).max() > 0
    if name == "igw":
        assert np.abs(ref[1]).max() > 0 and (n[1] == 1 or np.abs(ref[2]).max() > 0)
    err = per_var_rel_err(got, ref)
    figures("per-variable error %s" % ["%.1e" % e for e in err])
    # device sqrt/pow/cos/sin against glibc: a few ulp of each cell's integrand
    assert max(err) <= 1e-13
    assert np.all(got[0] == 0) and np.all(got[3] == 0)


def test_igw_background_bitwise():
    _need_gpu()
    D, O, _, _ = _pair("igw", 8, 8, 32)
    for w in range(8):
        assert np.array_equal(D.background(w), O.background(w)), w


@pytest.mark.parametrize("n,bc", [((40, 36, 24), (0, 0, 1)), ((64, 1, 32), (0, 0, 1)), ((32, 28, 20), (1, 1, 1))])
def test_igw_steps_match_oracle(n, bc, figures):
    """Ten fused steps of the igw case (constant-N background, 20 m/s winds)."""
    _need_gpu()
    D, O, spec, ospec = _pair("igw", *n, bc=bc)
    st = D.init_state(spec)
    host = st.cpu().numpy().copy()
    dt = D.compute_time_step(st, 0.8)
    assert abs(dt - O.compute_time_step(host.copy(), 0.8)) <= 1e-14 * dt
    O.set_dt_cache(dt)
    hs = O.hs
    rbar = O.background(0)[hs:hs + n[2]][:, None, None]
    rtbar = O.background(3)[hs:hs + n[2]][:, None, None]
    for s in range(10):
        D.time_step(st, dt)
        O.time_step(host, dt)
        g = D.interior(st).cpu().numpy()
        ref = O.interior(host)
        f = full_rel_err(g, ref, rbar, rtbar)
        assert max(f[0], f[4]) <= 1e-12 * (s + 1), (s, f)
        assert max(f[1:4]) <= 1e-11 * (s + 1), (s, f)
    e = per_var_rel_err(g, ref)
    figures("after 10 steps: full %s pert %s" % (["%.1e" % v for v in f], ["%.1e" % v for v in e]))
    assert max(e) <= 1e-10
    assert bool(torch.isfinite(st).all())
