"""Device build of pow_gamma (through the equation-of-state entry point)
agrees with the host build to 1 ulp and with the reference EOS to 2 ulp."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_device_pow_gamma_matches_host():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from awflcloud_amd import Dycore
    from test_fastmath import C0, GAMMA, host_pow_gamma
    D = Dycore(8, 8, 8)
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(20, 600, 50000), np.exp(rng.uniform(np.log(1e-3), np.log(1e8), 50000))])
    p = D.eos_pressure(torch.from_numpy(x).cuda()).cpu().numpy()
    host = C0 * host_pow_gamma(x)
    dh = np.abs(p - host) / np.spacing(np.abs(host))
    ref = C0 * x ** GAMMA
    dr = np.abs(p - ref) / np.spacing(np.abs(ref))
    i, j = int(np.argmax(dh)), int(np.argmax(dr))
    print("device vs host: max %.2f ulp at x=%r (%r vs %r); vs libm: max %.2f ulp at x=%r"
          % (dh[i], x[i], p[i], host[i], dr[j], x[j]))
    assert dh.max() <= 1.0
    assert dr.max() <= 3.0
