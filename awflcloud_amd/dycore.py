"""ctypes binding of the C ABI in include/awfl_dycore.h.

The Python-side mirror of the reference's Spatial operator
(Spatial_euler3d_cons_expl_cart_fv_Agrid.h) and of Temporal_ader::timeStep,
operating on torch device tensors as plain device pointers.  Every numerical
operation runs in the HIP library; if the library is missing the import of a
``Dycore`` fails loudly -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libawfl_dycore.so")
# per-device scratch of awfl_time_step_noalloc (AWFL_NOALLOC_MAX_BYTES)
NOALLOC_MAX_BYTES = 64 * 1024 * 1024

AWFL_BC_PERIODIC, AWFL_BC_WALL = 0, 1
AWFL_FP64, AWFL_FP32_PERT = 0, 1


class AwflConfig(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xlen", C.c_double), ("ylen", C.c_double), ("zlen", C.c_double),
                ("bc_x", C.c_int), ("bc_y", C.c_int), ("bc_z", C.c_int),
                ("weno_scalars", C.c_int), ("weno_winds", C.c_int),
                ("ord", C.c_int), ("ngll", C.c_int), ("data_spec", C.c_int), ("precision", C.c_int)]


class InitSpec(C.Structure):
    _fields_ = [("nbub", C.c_int), ("shape", C.c_int * 4),
                ("x0", C.c_double * 4), ("y0", C.c_double * 4), ("z0", C.c_double * 4),
                ("xr", C.c_double * 4), ("yr", C.c_double * 4), ("zr", C.c_double * 4),
                ("amp", C.c_double * 4), ("pwr", C.c_double * 4),
                ("noise_amp", C.c_double), ("noise_seed", C.c_uint64),
                ("nmodes", C.c_int), ("modes", C.POINTER(C.c_double)),
                ("u0", C.c_double), ("v0", C.c_double),
                ("y2d", C.c_double), ("perturb_amp", C.c_double), ("perturb_seed", C.c_uint64)]


def bubble_spec(bubbles=(), noise_amp=0.0, noise_seed=0, modes=None, u0=0.0, v0=0.0, y2d=0.0, perturb_amp=0.0,
                perturb_seed=0):
    """bubbles: iterable of dicts(shape, x0, y0, z0, xr, yr, zr, amp, pwr)."""
    s = InitSpec()
    bubbles = list(bubbles)
    if len(bubbles) > 4:
        raise ValueError("at most 4 ellipsoids")
    s.nbub = len(bubbles)
    for b, d in enumerate(bubbles):
        s.shape[b] = d.get("shape", 1)
        for key in ("x0", "y0", "z0", "xr", "yr", "zr", "amp"):
            getattr(s, key)[b] = d[key]
        s.pwr[b] = d.get("pwr", 1.0)
    s.noise_amp = noise_amp
    s.noise_seed = noise_seed
    s._keep = None
    if modes is not None and len(modes):
        m = np.ascontiguousarray(np.asarray(modes, dtype=np.float64).reshape(-1, 5))
        s._keep = m
        s.nmodes = m.shape[0]
        s.modes = m.ctypes.data_as(C.POINTER(C.c_double))
    s.u0, s.v0 = u0, v0
    s.y2d, s.perturb_amp, s.perturb_seed = y2d, perturb_amp, perturb_seed
    return s


def band_limited_modes(seed, n=64, kmin=4.0, kmax=16.0, rms=0.2):
    """Seeded band-limited theta' field as n Fourier modes (kx, ky, kz, amp, phase).

    Integer wavevectors with kmin <= |k| <= kmax (distinct, kx > 0 so no two
    modes alias), random phases and one common amplitude chosen so that the
    field's mean square over the periodic box is rms**2 (each cosine
    contributes amp**2 / 2)."""
    rng = np.random.default_rng(seed)
    lim = int(np.floor(kmax))
    seen, rows = set(), []
    while len(rows) < n:
        k = rng.integers(-lim, lim + 1, size=3)
        k[0] = abs(k[0])
        if k[0] == 0 or not (kmin <= np.sqrt((k * k).sum()) <= kmax) or tuple(k) in seen:
            continue
        seen.add(tuple(k))
        rows.append([float(k[0]), float(k[1]), float(k[2]), 0.0, rng.uniform(0.0, 2.0 * np.pi)])
    m = np.array(rows)
    m[:, 3] = rms / np.sqrt(n / 2.0)
    return m


_P, _D, _I64 = C.c_void_p, C.c_double, C.c_int64
# C ABI of include/awfl_dycore.h: name -> (argument types, result type)
_SIGNATURES = {
    "awfl_create": ([C.POINTER(AwflConfig), C.POINTER(_P)], C.c_int),
    "awfl_destroy": ([_P], None),
    "awfl_last_error": ([_P], C.c_char_p),
    "awfl_layout": ([_P, C.POINTER(C.c_int), C.POINTER(_I64), C.POINTER(_I64)], C.c_int),
    "awfl_num_split": ([_P], C.c_int),
    "awfl_set_stream": ([_P, _P], C.c_int),
    "awfl_init_state": ([_P, _P, C.POINTER(InitSpec)], C.c_int),
    "awfl_compute_time_step": ([_P, _P, _D, C.POINTER(_D)], C.c_int),
    "awfl_set_time_step": ([_P, _D], C.c_int),
    "awfl_compute_tendencies": ([_P, _P, _P, _D, C.c_int], C.c_int),
    "awfl_apply_tendencies": ([_P, _P, _P, _D, C.c_int], C.c_int),
    "awfl_time_step": ([_P, _P, _D], C.c_int),
    "awfl_time_step_noalloc": ([_P, _P, _D], C.c_int),
    "awfl_sweep_timing": ([_P, C.c_int], C.c_int),
    "awfl_sweep_times": ([_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)], C.c_int),
    "awfl_time_step_branches": ([_P, _P, _D, _P, C.c_int], C.c_int),
    "awfl_sweep": ([_P, _P, _P, _D, C.c_int], C.c_int),
    "awfl_get_dim_switch": ([_P], C.c_int),
    "awfl_set_dim_switch": ([_P, C.c_int], C.c_int),
    "awfl_get_background": ([_P, C.c_int, C.POINTER(_D), _I64], C.c_int),
    "awfl_state_sums": ([_P, _P, C.POINTER(_D)], C.c_int),
    "awfl_eos_pressure": ([_P, _P, _P, _I64], C.c_int),
    "awfl_state_to_f32": ([_P, _P, _P, _P], C.c_int),
    "awfl_state_from_f32": ([_P, _P, _P, _P], C.c_int),
    "awfl_sweep_f32": ([_P, _P, _P, _P, _P, _D, C.c_int], C.c_int),
    "awfl_time_step_f32": ([_P, _P, _P, _D], C.c_int),
    "awfl_time_step_strang": ([_P, _P, _D], C.c_int),
    "awfl_time_step_ssprk3": ([_P, _P, _D], C.c_int),
    "awfl_sweep_tangent": ([_P, _P, _P, _P, _P, _D, C.c_int], C.c_int),
    "awfl_sweep_adjoint": ([_P, _P, _P, _P, _D, C.c_int], C.c_int),
    "awfl_time_step_adjoint": ([_P, _P, _P, _D, C.c_int], C.c_int),
    "awfl_forward_ckpt": ([_P, _P, _D, C.c_int, C.c_int, _P], C.c_int),
    "awfl_adjoint_ckpt": ([_P, _P, _D, C.c_int, C.c_int, C.c_int, _P], C.c_int),
    "awfl_set_tracers": ([_P, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "awfl_num_tracers": ([_P], C.c_int),
    "awfl_init_tracers": ([_P, _P], C.c_int),
    "awfl_compute_tendencies_tracers": ([_P, _P, _P, _P, _P, _D, C.c_int], C.c_int),
    "awfl_apply_tracer_tendencies": ([_P, _P, _P, _D], C.c_int),
    "awfl_sweep_tracers": ([_P, _P, _P, _P, _P, _D, C.c_int], C.c_int),
    "awfl_time_step_tracers": ([_P, _P, _P, _D], C.c_int),
}

_libs = {}


def lib(path=None):
    """Load libawfl_dycore.so, or with `path` one of the parity test builds of
    its FP64 sweeps (awfl_dycore.hip alone, so the FP32, adjoint, tracer and
    integrator entry points are absent there); raises if it is not built."""
    path = path or LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError("HIP library %s is missing; run `python -m awflcloud_amd.build`" % path)
        L = C.CDLL(path)
        for name, (argtypes, restype) in _SIGNATURES.items():
            if path != LIB_PATH and not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.argtypes = argtypes
            f.restype = restype
        _libs[path] = L
    return _libs[path]


class AwflError(RuntimeError):
    pass


class Dycore:
    """One Spatial/Temporal instance on the current HIP device."""

    def __init__(self, nx, ny, nz, xlen=20000.0, ylen=20000.0, zlen=10000.0, bc=(0, 0, 1), ord=5, ngll=3,
                 weno_scalars=True, weno_winds=True, data_spec=1, precision=AWFL_FP64, stream=None, library=None):
        import torch
        self.torch = torch
        self.L = lib(library)
        cfg = AwflConfig(nx, ny, nz, xlen, ylen, zlen, bc[0], bc[1], bc[2], int(weno_scalars), int(weno_winds),
                         ord, ngll, data_spec, precision)
        h = C.c_void_p()
        rc = self.L.awfl_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise AwflError("awfl_create failed (%d): %s" % (rc, self.L.awfl_last_error(None).decode()))
        self.h = h
        self.nx, self.ny, self.nz, self.ord, self.ngll = nx, ny, nz, ord, ngll
        self.hs = (ord - 1) // 2
        self.dx, self.dy, self.dz = xlen / nx, ylen / ny, zlen / nz
        self.ntr = 0
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._check(self.L.awfl_set_stream(self.h, C.c_void_p(stream)))

    def close(self):
        if getattr(self, "h", None):
            self.L.awfl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise AwflError("awfl call failed (%d): %s" % (rc, self.L.awfl_last_error(self.h).decode()))
        return rc

    @staticmethod
    def _ptr(t):
        return C.c_void_p(t.data_ptr())

    def _chk_state(self, t):
        if t.dtype != self.torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("state must be a contiguous float64 device tensor")
        if tuple(t.shape) != self.state_shape:
            raise ValueError("state shape %s != %s" % (tuple(t.shape), self.state_shape))

    def _chk_tend(self, t):
        if t.dtype != self.torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("tendency must be a contiguous float64 device tensor")
        if tuple(t.shape) != self.tend_shape:
            raise ValueError("tendency shape %s != %s" % (tuple(t.shape), self.tend_shape))

    # --- layout ---
    @property
    def state_shape(self):
        hs = self.hs
        return (5, self.nz + 2 * hs, self.ny + 2 * hs, self.nx + 2 * hs)

    @property
    def tend_shape(self):
        return (5, self.nz, self.ny, self.nx)

    def num_split(self):
        return self._check(self.L.awfl_num_split(self.h))

    def new_state(self):
        return self.torch.zeros(self.state_shape, dtype=self.torch.float64, device="cuda")

    def new_tend(self):
        return self.torch.zeros(self.tend_shape, dtype=self.torch.float64, device="cuda")

    def interior(self, st):
        hs = self.hs
        return st[:, hs:hs + self.nz, hs:hs + self.ny, hs:hs + self.nx]

    # --- Spatial ---
    def init_state(self, spec, state=None):
        if state is None:
            state = self.new_state()
        self._chk_state(state)
        self._check(self.L.awfl_init_state(self.h, self._ptr(state), C.byref(spec)))
        return state

    def compute_time_step(self, state, cfl=0.8):
        self._chk_state(state)
        dt = C.c_double()
        self._check(self.L.awfl_compute_time_step(self.h, self._ptr(state), cfl, C.byref(dt)))
        return dt.value

    def set_time_step(self, dt):
        self._check(self.L.awfl_set_time_step(self.h, dt))

    def compute_tendencies(self, state, tend, dt, split):
        self._chk_state(state)
        self._chk_tend(tend)
        self._check(self.L.awfl_compute_tendencies(self.h, self._ptr(state), self._ptr(tend), dt, split))

    def apply_tendencies(self, state, tend, dt, split):
        self._chk_state(state)
        self._chk_tend(tend)
        self._check(self.L.awfl_apply_tendencies(self.h, self._ptr(state), self._ptr(tend), dt, split))

    # --- Temporal ---
    def time_step(self, state, dt):
        self._chk_state(state)
        self._check(self.L.awfl_time_step(self.h, self._ptr(state), dt))

    def sweep_timing(self, max_sweeps):
        """time the next max_sweeps sweeps of time_step with HIP events on
        the library's stream (0 stops)"""
        self._check(self.L.awfl_sweep_timing(self.h, int(max_sweeps)))

    def sweep_times(self):
        """(mean ms per direction x, y, z; counts) of the sweeps timed since
        the last call"""
        ms = (C.c_double * 3)()
        n = (C.c_int64 * 3)()
        self._check(self.L.awfl_sweep_times(self.h, ms, n))
        return list(ms), list(n)

    def time_step_noalloc(self, state, dt):
        """time_step without any allocation (awfl_time_step_noalloc): its
        intermediate state lives in the library's per-device scratch of
        NOALLOC_MAX_BYTES; larger states raise."""
        self._chk_state(state)
        self._check(self.L.awfl_time_step_noalloc(self.h, self._ptr(state), dt))

    def new_branch_record(self):
        """uint8 device buffer for the upwind branches of one step (three
        splits of (nx+1)(ny+1)(nz+1) faces each)."""
        n = 3 * (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        return self.torch.zeros(n, dtype=self.torch.uint8, device="cuda")

    def time_step_branches(self, state, dt, branches, mode):
        """time_step recording (mode 1) or imposing (mode 2) the upwind branch
        of every face (awfl_time_step_branches)."""
        self._chk_state(state)
        if branches.dtype != self.torch.uint8 or not branches.is_cuda or branches.numel() < \
                3 * (self.nx + 1) * (self.ny + 1) * (self.nz + 1):
            raise ValueError("branch record must be a uint8 device tensor from new_branch_record()")
        self._check(self.L.awfl_time_step_branches(self.h, self._ptr(state), dt, self._ptr(branches), mode))

    def time_step_unfused(self, state, tend, dt):
        """Temporal_ader::timeStep through computeTendencies + applyTendencies."""
        for spl in range(self.num_split()):
            self.compute_tendencies(state, tend, dt, spl)
            self.apply_tendencies(state, tend, dt, spl)

    # --- tracers (masses rho*q, laid out like the state, one slab per tracer) ---
    def set_tracers(self, positive):
        """positive: one tracerPos flag per tracer"""
        flags = (C.c_int * max(1, len(positive)))(*[int(bool(p)) for p in positive])
        self._check(self.L.awfl_set_tracers(self.h, len(positive), flags))
        self.ntr = len(positive)

    @property
    def tracer_shape(self):
        hs = self.hs
        return (self.ntr, self.nz + 2 * hs, self.ny + 2 * hs, self.nx + 2 * hs)

    @property
    def tracer_tend_shape(self):
        return (self.ntr, self.nz, self.ny, self.nx)

    def new_tracers(self):
        return self.torch.zeros(self.tracer_shape, dtype=self.torch.float64, device="cuda")

    def init_tracers(self):
        """the reference's initTracers fields (uniform, theta replica, block)"""
        tr = self.new_tracers()
        self._check(self.L.awfl_init_tracers(self.h, self._ptr(tr)))
        return tr

    def _chk_tr(self, t, shape):
        if t.dtype != self.torch.float64 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != shape:
            raise ValueError("tracer array must be a contiguous float64 device tensor of shape %s" % (shape,))

    def sweep_tracers(self, state, tracers, out, tr_out, dt, direction):
        self._chk_state(state)
        self._chk_state(out)
        self._chk_tr(tracers, self.tracer_shape)
        self._chk_tr(tr_out, self.tracer_shape)
        self._check(self.L.awfl_sweep_tracers(self.h, self._ptr(state), self._ptr(tracers), self._ptr(out),
                                              self._ptr(tr_out), dt, direction))

    def compute_tendencies_tracers(self, state, tracers, tend, ttend, dt, split):
        self._chk_state(state)
        self._chk_tend(tend)
        self._chk_tr(tracers, self.tracer_shape)
        self._chk_tr(ttend, self.tracer_tend_shape)
        self._check(self.L.awfl_compute_tendencies_tracers(self.h, self._ptr(state), self._ptr(tracers),
                                                           self._ptr(tend), self._ptr(ttend), dt, split))

    def apply_tracer_tendencies(self, tracers, ttend, dt):
        self._chk_tr(tracers, self.tracer_shape)
        self._chk_tr(ttend, self.tracer_tend_shape)
        self._check(self.L.awfl_apply_tracer_tendencies(self.h, self._ptr(tracers), self._ptr(ttend), dt))

    def time_step_tracers(self, state, tracers, dt):
        self._chk_state(state)
        self._chk_tr(tracers, self.tracer_shape)
        self._check(self.L.awfl_time_step_tracers(self.h, self._ptr(state), self._ptr(tracers), dt))

    def tracer_interior(self, tr):
        hs = self.hs
        return tr[:, hs:hs + self.nz, hs:hs + self.ny, hs:hs + self.nx]

    def time_step_strang(self, state, dt):
        """x, y with dt and z as two dt/2 sweeps (qjrmsCode/TimeIntegrator.cpp:32-65)"""
        self._chk_state(state)
        self._check(self.L.awfl_time_step_strang(self.h, self._ptr(state), dt))

    def time_step_ssprk3(self, state, dt):
        """SSPRK3 on the semi-discrete operator (qjrmsCode/TimeIntegrator.cpp:68-110)"""
        self._chk_state(state)
        self._check(self.L.awfl_time_step_ssprk3(self.h, self._ptr(state), dt))

    def sweep(self, src, dst, dt, direction):
        self._chk_state(src)
        self._chk_state(dst)
        self._check(self.L.awfl_sweep(self.h, self._ptr(src), self._ptr(dst), dt, direction))

    @property
    def dim_switch(self):
        return self._check(self.L.awfl_get_dim_switch(self.h))

    @dim_switch.setter
    def dim_switch(self, v):
        self._check(self.L.awfl_set_dim_switch(self.h, int(v)))

    def background(self, which):
        cap = self.nz * self.ngll + self.nz + 2 * self.hs
        out = np.zeros(cap)
        n = self._check(self.L.awfl_get_background(self.h, which, out.ctypes.data_as(C.POINTER(C.c_double)), cap))
        return out[:n].copy()

    def eos_pressure(self, rt):
        """p = C0 (rho theta)^GAMMA on the device for a float64 device tensor."""
        if rt.dtype != self.torch.float64 or not rt.is_cuda or not rt.is_contiguous():
            raise ValueError("rt must be a contiguous float64 device tensor")
        p = self.torch.empty_like(rt)
        self._check(self.L.awfl_eos_pressure(self.h, self._ptr(rt), self._ptr(p), rt.numel()))
        return p

    # --- FP32 perturbation form: a state is a pair (s, c) holding s - c ---
    def _chk_f32(self, t):
        if t.dtype != self.torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("FP32 state must be a contiguous float32 device tensor")
        if tuple(t.shape) != self.state_shape:
            raise ValueError("FP32 state shape %s != %s" % (tuple(t.shape), self.state_shape))

    def to_f32(self, state):
        """FP64 state -> (s, c), float32 tensors with s - c = state to ~2^-48."""
        self._chk_state(state)
        s = self.torch.empty(self.state_shape, dtype=self.torch.float32, device="cuda")
        c = self.torch.empty_like(s)
        self._check(self.L.awfl_state_to_f32(self.h, self._ptr(state), self._ptr(s), self._ptr(c)))
        return s, c

    def from_f32(self, s, c, state=None):
        self._chk_f32(s)
        self._chk_f32(c)
        if state is None:
            state = self.new_state()
        self._chk_state(state)
        self._check(self.L.awfl_state_from_f32(self.h, self._ptr(s), self._ptr(c), self._ptr(state)))
        return state

    def time_step_f32(self, s, c, dt):
        self._chk_f32(s)
        self._chk_f32(c)
        self._check(self.L.awfl_time_step_f32(self.h, self._ptr(s), self._ptr(c), dt))

    def sweep_f32(self, s_in, c_in, s_out, c_out, dt, direction):
        for t in (s_in, c_in, s_out, c_out):
            self._chk_f32(t)
        self._check(self.L.awfl_sweep_f32(self.h, self._ptr(s_in), self._ptr(c_in), self._ptr(s_out),
                                          self._ptr(c_out), dt, direction))

    # --- tangent-linear and adjoint (FP64, dt held fixed) ---
    def sweep_tangent(self, state, dstate, dt, direction):
        """(sweep(state), J dstate) for one split sweep"""
        self._chk_state(state)
        self._chk_state(dstate)
        out, dout = self.new_state(), self.new_state()
        self._check(self.L.awfl_sweep_tangent(self.h, self._ptr(state), self._ptr(dstate), self._ptr(out),
                                              self._ptr(dout), dt, direction))
        return out, dout

    def sweep_adjoint(self, state, adj_out, dt, direction, out=None):
        """J(state)^T adj_out for one split sweep, into `out` (a new state
        when None; its halo comes back zero)"""
        self._chk_state(state)
        self._chk_state(adj_out)
        if out is None:
            adj_in = self.new_state()
        else:
            self._chk_state(out)
            adj_in = out
        self._check(self.L.awfl_sweep_adjoint(self.h, self._ptr(state), self._ptr(adj_out), self._ptr(adj_in),
                                              dt, direction))
        return adj_in

    def time_step_adjoint(self, state_before, adj, dt, dim_switch):
        """adj <- (d step / d state)^T adj in place, for the step taken from
        state_before with the given dimSwitch"""
        self._chk_state(state_before)
        self._chk_state(adj)
        self._check(self.L.awfl_time_step_adjoint(self.h, self._ptr(state_before), self._ptr(adj), dt,
                                                  int(dim_switch)))

    def forward_ckpt(self, state, dt, nsteps, every):
        """nsteps fused steps in place; returns the checkpoint tensor"""
        self._chk_state(state)
        nslot = max(1, (nsteps + every - 1) // every)
        ckpt = self.torch.empty((nslot,) + self.state_shape, dtype=self.torch.float64, device="cuda")
        self._check(self.L.awfl_forward_ckpt(self.h, self._ptr(state), dt, nsteps, every, self._ptr(ckpt)))
        return ckpt

    def adjoint_ckpt(self, ckpt, dt, nsteps, every, dim_switch0, adj):
        """adj: dJ/d(final state) in, dJ/d(initial state) out"""
        self._chk_state(adj)
        nslot = max(1, (nsteps + every - 1) // every)
        if ckpt.dtype != self.torch.float64 or not ckpt.is_cuda or not ckpt.is_contiguous() or \
                tuple(ckpt.shape) != (nslot,) + self.state_shape:
            raise ValueError("checkpoints must be the tensor forward_ckpt returned")
        self._check(self.L.awfl_adjoint_ckpt(self.h, self._ptr(ckpt), dt, nsteps, every, int(dim_switch0),
                                             self._ptr(adj)))

    def state_sums(self, state):
        self._chk_state(state)
        out = np.zeros(5)
        self._check(self.L.awfl_state_sums(self.h, self._ptr(state), out.ctypes.data_as(C.POINTER(C.c_double))))
        return out
