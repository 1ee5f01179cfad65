"""ctypes wrapper of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OracleConfig(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xlen", C.c_double), ("ylen", C.c_double), ("zlen", C.c_double),
                ("bc_x", C.c_int), ("bc_y", C.c_int), ("bc_z", C.c_int),
                ("weno_scalars", C.c_int), ("weno_winds", C.c_int),
                ("ord", C.c_int), ("ngll", C.c_int), ("data_spec", C.c_int), ("nthreads", C.c_int)]


class OracleInitSpec(C.Structure):
    _fields_ = [("nbub", C.c_int), ("shape", C.c_int * 4),
                ("x0", C.c_double * 4), ("y0", C.c_double * 4), ("z0", C.c_double * 4),
                ("xr", C.c_double * 4), ("yr", C.c_double * 4), ("zr", C.c_double * 4),
                ("amp", C.c_double * 4), ("pwr", C.c_double * 4),
                ("noise_amp", C.c_double), ("noise_seed", C.c_uint64),
                ("nmodes", C.c_int), ("modes", C.POINTER(C.c_double)),
                ("u0", C.c_double), ("v0", C.c_double),
                ("y2d", C.c_double), ("perturb_amp", C.c_double), ("perturb_seed", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: build it with `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        D = C.POINTER(C.c_double)
        L.oracle_create.restype = P
        L.oracle_create.argtypes = [C.POINTER(OracleConfig)]
        L.oracle_destroy.argtypes = [P]
        L.oracle_hs.argtypes = [P]
        L.oracle_get_table.argtypes = [P, C.c_int, D]
        L.oracle_get_background.argtypes = [P, C.c_int, D]
        L.oracle_compute_time_step.restype = C.c_double
        L.oracle_compute_time_step.argtypes = [P, D, C.c_double]
        L.oracle_set_dt_cache.argtypes = [P, C.c_double]
        L.oracle_compute_tendencies.argtypes = [P, D, D, C.c_double, C.c_int]
        L.oracle_tendencies_dir.argtypes = [P, D, D, C.c_double, C.c_int]
        L.oracle_time_step.argtypes = [P, D, C.c_double]
        L.oracle_run_steps.restype = C.c_double
        L.oracle_run_steps.argtypes = [P, D, C.c_double, C.c_int]
        L.oracle_get_dim_switch.argtypes = [P]
        L.oracle_set_dim_switch.argtypes = [P, C.c_int]
        L.oracle_weno_coefs.argtypes = [P, D, D]
        L.oracle_reconstruct.argtypes = [P, D, C.c_int, D]
        L.oracle_diff_transform.argtypes = [P, C.c_int, D, C.c_double, C.c_int]
        L.oracle_riemann.argtypes = [P, C.c_int, D, D, C.c_int, D]
        L.oracle_init_state.argtypes = [P, C.POINTER(OracleInitSpec), D]
        L.oracle_set_tracers.argtypes = [P, C.c_int, C.POINTER(C.c_int)]
        L.oracle_num_tracers.argtypes = [P]
        L.oracle_compute_tendencies_tr.argtypes = [P, D, D, D, D, C.c_double, C.c_int]
        L.oracle_tendencies_dir_tr.argtypes = [P, D, D, D, D, C.c_double, C.c_int]
        L.oracle_time_step_tr.argtypes = [P, D, D, C.c_double]
        L.oracle_init_tracers.argtypes = [P, D]
        L.oracle_profile.argtypes = [C.c_int, C.c_double, D, D, D]
        L.oracle_set_pow_jitter.argtypes = [C.c_uint64]
        L.oracle_set_pow_mode.argtypes = [C.c_int]
        L.oracle_eos_pow.restype = C.c_double
        L.oracle_eos_pow.argtypes = [C.c_double]
        L.oracle_uniform_pm1.restype = C.c_double
        L.oracle_uniform_pm1.argtypes = [C.c_uint64, C.c_uint64]
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def profile(family, z):
    """(theta, rho, p) of the background at height z: family 0 constant theta,
    1 constant Brunt-Vaisala frequency (Profiles.h:8-47)."""
    th, rh, ph = C.c_double(), C.c_double(), C.c_double()
    lib().oracle_profile(family, z, C.byref(th), C.byref(rh), C.byref(ph))
    return th.value, rh.value, ph.value


def make_spec(bubbles=(), noise_amp=0.0, noise_seed=0, modes=None, u0=0.0, v0=0.0, y2d=0.0, perturb_amp=0.0,
              perturb_seed=0):
    """bubbles: iterable of dicts(shape, x0, y0, z0, xr, yr, zr, amp, pwr)."""
    s = OracleInitSpec()
    bubbles = list(bubbles)
    assert len(bubbles) <= 4
    s.nbub = len(bubbles)
    for b, d in enumerate(bubbles):
        s.shape[b] = d.get("shape", 1)
        for key in ("x0", "y0", "z0", "xr", "yr", "zr", "amp"):
            getattr(s, key)[b] = d[key]
        s.pwr[b] = d.get("pwr", 1.0)
    s.noise_amp = noise_amp
    s.noise_seed = noise_seed
    s._modes_keep = None
    if modes is not None and len(modes):
        m = np.ascontiguousarray(np.asarray(modes, dtype=np.float64).reshape(-1, 5))
        s._modes_keep = m
        s.nmodes = m.shape[0]
        s.modes = _dp(m)
    s.u0, s.v0 = u0, v0
    s.y2d, s.perturb_amp, s.perturb_seed = y2d, perturb_amp, perturb_seed
    return s


class Oracle:
    """One reference Spatial + Temporal instance on the CPU."""

    def __init__(self, nx, ny, nz, xlen=20000.0, ylen=20000.0, zlen=10000.0, bc=(0, 0, 1),
                 ord=5, ngll=3, weno_scalars=True, weno_winds=True, data_spec=1, nthreads=0):
        cfg = OracleConfig(nx, ny, nz, xlen, ylen, zlen, bc[0], bc[1], bc[2], int(weno_scalars),
                           int(weno_winds), ord, ngll, data_spec, nthreads)
        self.L = lib()
        self.h = self.L.oracle_create(C.byref(cfg))
        if not self.h:
            raise ValueError("oracle_create rejected the configuration")
        self.nx, self.ny, self.nz, self.ord, self.ngll = nx, ny, nz, ord, ngll
        self.hs = (ord - 1) // 2
        self.dx, self.dy, self.dz = xlen / nx, ylen / ny, zlen / nz

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oracle_destroy(self.h)
            self.h = None

    # --- layout helpers ---
    def state_shape(self):
        hs = self.hs
        return (5, self.nz + 2 * hs, self.ny + 2 * hs, self.nx + 2 * hs)

    def tend_shape(self):
        return (5, self.nz, self.ny, self.nx)

    def new_state(self):
        return np.zeros(self.state_shape())

    def interior(self, st):
        hs = self.hs
        return st[:, hs:hs + self.nz, hs:hs + self.ny, hs:hs + self.nx]

    # --- wrapped calls ---
    def table(self, which):
        out = np.zeros(1024)
        n = self.L.oracle_get_table(self.h, which, _dp(out))
        return out[:n].copy()

    def background(self, which):
        out = np.zeros(self.nz * max(self.ngll, 1) + 2 * self.hs + self.nz)
        n = self.L.oracle_get_background(self.h, which, _dp(out))
        return out[:n].copy()

    def init_state(self, spec):
        st = self.new_state()
        self.L.oracle_init_state(self.h, C.byref(spec), _dp(st))
        return st

    def compute_time_step(self, state, cfl=0.8):
        return self.L.oracle_compute_time_step(self.h, _dp(state), cfl)

    def set_dt_cache(self, dt):
        self.L.oracle_set_dt_cache(self.h, dt)

    def compute_tendencies(self, state, dt, split):
        tend = np.zeros(self.tend_shape())
        self.L.oracle_compute_tendencies(self.h, _dp(state), _dp(tend), dt, split)
        return tend

    def tendencies_dir(self, state, dt, direction):
        tend = np.zeros(self.tend_shape())
        self.L.oracle_tendencies_dir(self.h, _dp(state), _dp(tend), dt, direction)
        return tend

    def time_step(self, state, dt):
        self.L.oracle_time_step(self.h, _dp(state), dt)

    # --- tracers ---
    def set_tracers(self, positive):
        """positive: one flag per tracer (tracerPos)."""
        flags = (C.c_int * max(1, len(positive)))(*[int(bool(p)) for p in positive])
        if self.L.oracle_set_tracers(self.h, len(positive), flags) != 0:
            raise ValueError("oracle_set_tracers failed")
        self.ntr = len(positive)

    def tracer_shape(self):
        hs = self.hs
        return (self.ntr, self.nz + 2 * hs, self.ny + 2 * hs, self.nx + 2 * hs)

    def tracer_tend_shape(self):
        return (self.ntr, self.nz, self.ny, self.nx)

    def tendencies_dir_tr(self, state, tracers, dt, direction):
        """(state tendency, tracer tendency) of one direction; tracers is
        divided by and multiplied back by the density as the reference does."""
        tend = np.zeros(self.tend_shape())
        ttend = np.zeros(self.tracer_tend_shape())
        self.L.oracle_tendencies_dir_tr(self.h, _dp(state), _dp(tend), _dp(tracers), _dp(ttend), dt, direction)
        return tend, ttend

    def compute_tendencies_tr(self, state, tracers, dt, split):
        tend = np.zeros(self.tend_shape())
        ttend = np.zeros(self.tracer_tend_shape())
        self.L.oracle_compute_tendencies_tr(self.h, _dp(state), _dp(tend), _dp(tracers), _dp(ttend), dt, split)
        return tend, ttend

    def init_tracers(self):
        tr = np.zeros(self.tracer_shape())
        self.L.oracle_init_tracers(self.h, _dp(tr))
        return tr

    def time_step_tr(self, state, tracers, dt):
        self.L.oracle_time_step_tr(self.h, _dp(state), _dp(tracers), dt)

    def tracer_interior(self, tr):
        hs = self.hs
        return tr[:, hs:hs + self.nz, hs:hs + self.ny, hs:hs + self.nx]

    def density(self, state):
        """Full density of every interior cell (rho' + hyDensCells)."""
        hs = self.hs
        rb = self.background(0)[hs:hs + self.nz]
        return self.interior(state)[0] + rb[:, None, None]

    # --- alternative time integrators over the same spatial operator ---
    def sweep(self, state, dt, direction):
        """One split sweep in place: interior += dt * tendency of `direction`
        (the update of Temporal_ader.h:63-65 for one split)."""
        tend = self.tendencies_dir(state, dt, direction)
        self.interior(state)[...] += dt * tend

    def strang_step(self, state, dt, ds):
        """qjrmsCode/TimeIntegrator.cpp:32-65: x, y and then z as two half
        steps when ds is set, the reverse order otherwise; y is skipped when
        ny == 1 (run2d)."""
        seq = [(0, dt), (1, dt), (2, dt / 2), (2, dt / 2)]
        if not ds:
            seq = seq[::-1]
        for d, h in seq:
            if d == 1 and self.ny == 1:
                continue
            self.sweep(state, h, d)

    def semi_discrete_tendency(self, state):
        """Sum of the three directional tendencies without the ADER time
        average (dt = 0 in the transform: only the reconstructed values reach
        the faces), as the semi-discrete operator of an explicit RK method."""
        tend = self.tendencies_dir(state, 0.0, 0)
        if self.ny > 1:
            tend += self.tendencies_dir(state, 0.0, 1)
        tend += self.tendencies_dir(state, 0.0, 2)
        return tend

    def ssprk3_step(self, state, dt):
        """qjrmsCode/TimeIntegrator.cpp:68-96 with applyTendencies of :99-110
        (state2 = c0 state0 + c1 state1 + ct dt tend, in that order)."""
        I = self.interior
        tmp = state.copy()
        t = self.semi_discrete_tendency(state)
        I(tmp)[...] = 1.0 * I(state) + 0.0 * I(tmp) + 1.0 * dt * t
        t = self.semi_discrete_tendency(tmp)
        I(tmp)[...] = 0.75 * I(state) + 0.25 * I(tmp) + 0.25 * dt * t
        t = self.semi_discrete_tendency(tmp)
        I(state)[...] = (1.0 / 3.0) * I(state) + (2.0 / 3.0) * I(tmp) + (2.0 / 3.0) * dt * t

    def run_steps(self, state, dt, nsteps):
        return self.L.oracle_run_steps(self.h, _dp(state), dt, nsteps)

    @property
    def dim_switch(self):
        return self.L.oracle_get_dim_switch(self.h)

    @dim_switch.setter
    def dim_switch(self, v):
        self.L.oracle_set_dim_switch(self.h, int(v))

    def weno_coefs(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        aw = np.zeros(self.ord)
        self.L.oracle_weno_coefs(self.h, _dp(u), _dp(aw))
        return aw

    def reconstruct(self, u, doweno=True):
        u = np.ascontiguousarray(u, dtype=np.float64)
        g = np.zeros(self.ngll)
        self.L.oracle_reconstruct(self.h, _dp(u), int(doweno), _dp(g))
        return g

    def diff_transform(self, direction, arrs, dx, k_cell=0):
        a = np.ascontiguousarray(arrs, dtype=np.float64).copy()
        self.L.oracle_diff_transform(self.h, direction, _dp(a), dx, k_cell)
        return a

    def riemann(self, direction, qL, qR, kface=0):
        qL = np.ascontiguousarray(qL, dtype=np.float64)
        qR = np.ascontiguousarray(qR, dtype=np.float64)
        f = np.zeros(5)
        self.L.oracle_riemann(self.h, direction, _dp(qL), _dp(qR), kface, _dp(f))
        return f


def set_pow_jitter(seed):
    """Sensitivity probe: move every pow() of the step by one ulp (0 = off)."""
    lib().oracle_set_pow_jitter(seed)


def set_pow_mode(mode):
    """0: the C library's pow (default); 1: (rho theta)^gamma by the
    double-double method of the HIP parity builds (attribution runs only)."""
    lib().oracle_set_pow_mode(mode)


def eos_pow(x):
    """x^gamma as the oracle's step evaluates it (current pow mode)."""
    return lib().oracle_eos_pow(x)


def uniform_pm1(seed, idx):
    return lib().oracle_uniform_pm1(seed, idx)
