"""awfl_time_step_noalloc (the step that allocates nothing) against
awfl_time_step through the C ABI: bit for bit, over two calls in a row of
which the second has a smaller grid, in both sweep orders (x,y,z and z,y,x),
in 3-D and 2-D, for every (ord, ngll) the library builds; states beyond its
scratch are refused; the call leaves free device memory unchanged (no
state-sized buffer of the kind awfl_time_step allocates, and no scratch
memory the runtime would allocate for kernels that spill)."""
import pytest

from helpers import bubble_dicts

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _state(D, seed):
    from awflcloud_amd import bubble_spec
    return D.init_state(bubble_spec(bubble_dicts(), noise_amp=1e-3, noise_seed=seed, u0=4.0, v0=-3.0))


def _compare(nx, ny, nz, bc, ord, ngll, steps, seed):
    from awflcloud_amd import Dycore
    A = Dycore(nx, ny, nz, bc=bc, ord=ord, ngll=ngll)
    B = Dycore(nx, ny, nz, bc=bc, ord=ord, ngll=ngll)
    sa = _state(A, seed)
    sb = sa.clone()
    dt = A.compute_time_step(sa, 0.5)
    B.set_time_step(dt)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(steps):
        A.time_step_noalloc(sa, dt)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    for _ in range(steps):
        B.time_step(sb, dt)
    torch.cuda.synchronize()
    same = bool(torch.equal(A.interior(sa), B.interior(sb)))
    # halos are never written by either
    halo_same = bool(torch.equal(sa, sb))
    drop = free0 - free1
    A.close()
    B.close()
    return same, halo_same, drop, sa.numel() * 8


CASES = [((48, 40, 32), (0, 0, 1), 5, 3), ((40, 1, 32), (0, 0, 1), 5, 3), ((32, 28, 24), (1, 1, 1), 5, 3),
         ((36, 32, 20), (0, 0, 1), 3, 2), ((40, 36, 24), (0, 0, 1), 9, 5), ((36, 32, 28), (1, 1, 1), 9, 5)]


@pytest.mark.parametrize("n,bc,ord,ngll", CASES, ids=["%dx%dx%d-bc%d%d%d-o%d" % (*n, *bc, o) for n, bc, o, _ in CASES])
def test_noalloc_matches_time_step(n, bc, ord, ngll, figures):
    """Three steps (x,y,z / z,y,x / x,y,z), then a second call sequence on a
    smaller grid that reuses the same scratch."""
    _need_gpu()
    same, halo_same, drop, nbytes = _compare(*n, bc, ord, ngll, 3, 1)
    small = (max(n[0] // 2, 8), 1 if n[1] == 1 else max(n[1] // 2, 8), max(n[2] // 2, 8))
    same2, halo_same2, drop2, nbytes2 = _compare(*small, bc, ord, ngll, 2, 2)
    figures("bitwise equal %s / %s (smaller %s); free device memory change %d / %d bytes (state %d / %d)"
            % (same, same2, small, drop, drop2, nbytes, nbytes2))
    assert same and halo_same and same2 and halo_same2
    assert drop < nbytes and drop2 < nbytes2


def test_noalloc_rejects_large_state():
    _need_gpu()
    from awflcloud_amd import Dycore
    from awflcloud_amd.dycore import NOALLOC_MAX_BYTES, AwflError
    D = Dycore(256, 256, 40)
    assert 5 * 44 * 260 * 260 * 8 > NOALLOC_MAX_BYTES
    st = torch.zeros(D.state_shape, dtype=torch.float64, device="cuda")
    with pytest.raises(AwflError, match="exceeds its scratch"):
        D.time_step_noalloc(st, 0.1)
    assert bool((st == 0).all())
    D.close()
