"""ctypes binding of the C ABI declared in include/impec_hip.h.

The product path: every call goes to libimpec_hip.so (hand-written HIP for
gfx950).  There is no CPU fallback; loading fails loudly when the library is
missing and context creation fails with IMPEC_ENODEV when no GPU is visible.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libimpec_hip.so")

IMPEC_OK = 0
IMPEC_EINVAL = -1
IMPEC_EHIP = -2
IMPEC_ENOCONV = -3
IMPEC_ENAN = -4
IMPEC_ENOMEM = -5
IMPEC_ENODEV = -6

SRC_PARALLEL, SRC_DIAGONAL = 0, 1
AVG_HARMONIC, AVG_MEANCONC = 0, 1
SPLIT_ALTERNATING, SPLIT_X_FIRST, SPLIT_Y_FIRST, SPLIT_STRANG, SPLIT_PARALLEL = range(5)
HYPER_MUSCL, HYPER_UPWIND = 0, 1
PARA_EXPLICIT, PARA_CN = 0, 1
PREC_FP64, PREC_MIXED, PREC_MIXED32 = 0, 1, 2
GUESS_ZERO, GUESS_PREVIOUS, GUESS_LEARNED, GUESS_EXTRAPOLATE = 0, 1, 2, 3

_INT_FIELDS = ["N1", "N2", "num_ghostcells", "source_type", "average_procedure", "dim_splitting",
               "hyper_method", "diffusion", "para_method", "is_test", "alpha_absolute",
               "max_iterations", "precision", "initial_guess", "mg_pre", "mg_post", "mg_variant", "mg_rows", "cn_guess",
               "transport32"]
_DBL_FIELDS = ["x_lb", "x_ub", "y_lb", "y_ub", "nu", "mu_res", "mu_inj", "porosity", "cfl",
               "inj_fluid_conc", "alpha_l", "alpha_t", "dt_diff_limiter", "d_m", "eps", "para_eps",
               "I", "tstart", "inner_tol", "cn_ahead", "cg_ahead"]


class Params(C.Structure):
    _fields_ = ([(n, C.c_int) for n in _INT_FIELDS]
                + [(n, C.c_double) for n in _DBL_FIELDS] + [("reserved_d", C.c_double * 1)])


class Stats(C.Structure):
    _fields_ = [("last_its", C.c_int), ("last_vcycles", C.c_int), ("last_parabolic_its", C.c_int),
                ("total_its", C.c_int), ("calls", C.c_int), ("last_restarts", C.c_int),
                ("last_elim_its", C.c_int), ("cn_ahead_misses", C.c_int),
                ("last_relres", C.c_double), ("last_parabolic_relres", C.c_double),
                ("total_mass_outflux", C.c_double), ("last_guess_alpha", C.c_double),
                ("last_guess_gamma", C.c_double), ("last_relres_floor", C.c_double),
                ("cg_ahead_misses", C.c_int), ("reserved_i2", C.c_int), ("reserved_d", C.c_double * 1)]


# every entry point declared in include/impec_hip.h: name -> (argtypes, restype)
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
SIGNATURES = {
    "impec_params_default": ([C.POINTER(Params)], C.c_int),
    "impec_create": ([C.POINTER(Params), C.POINTER(C.c_void_p)], C.c_int),
    "impec_destroy": ([_vp], C.c_int),
    "impec_last_error": ([_vp], C.c_char_p),
    "impec_strerror": ([C.c_int], C.c_char_p),
    "impec_synchronize": ([_vp], C.c_int),
    "impec_device_count": ([], C.c_int),
    "impec_set_concentration": ([_vp, _dp], C.c_int),
    "impec_get_concentration": ([_vp, _dp], C.c_int),
    "impec_set_pressure": ([_vp, _dp], C.c_int),
    "impec_get_pressure": ([_vp, _dp], C.c_int),
    "impec_set_velocity": ([_vp, _dp], C.c_int),
    "impec_get_velocity": ([_vp, _dp], C.c_int),
    "impec_set_source": ([_vp, _dp], C.c_int),
    "impec_set_permeability": ([_vp, _dp, _dp], C.c_int),
    "impec_set_pin_kappa": ([_vp, C.c_double], C.c_int),
    "impec_fill_boundaries": ([_vp], C.c_int),
    "impec_pressure_solve": ([_vp, _ip, _dp], C.c_int),
    "impec_velocities": ([_vp], C.c_int),
    "impec_max_stable_timestep": ([_vp, _dp], C.c_int),
    "impec_max_stable_timestep_noalloc": ([C.POINTER(Params), _vp, _vp, _vp], C.c_int),
    "impec_source_split": ([_vp, C.c_double], C.c_int),
    "impec_advect_split": ([_vp, C.c_double], C.c_int),
    "impec_advect_x": ([_vp, C.c_double], C.c_int),
    "impec_advect_y": ([_vp, C.c_double], C.c_int),
    "impec_intercell_flux": ([_vp, C.c_int, C.c_double, _dp], C.c_int),
    "impec_cnn_config_default": ([_vp], C.c_int),
    "impec_cnn_create": ([_vp, C.POINTER(C.c_void_p)], C.c_int),
    "impec_cnn_destroy": ([_vp], C.c_int),
    "impec_cnn_last_error": ([_vp], C.c_char_p),
    "impec_cnn_num_params": ([_vp, C.POINTER(C.c_longlong)], C.c_int),
    "impec_cnn_get_params": ([_vp, _fp], C.c_int),
    "impec_cnn_set_params": ([_vp, _fp], C.c_int),
    "impec_cnn_get_scale": ([_vp, C.POINTER(C.c_float)], C.c_int),
    "impec_cnn_set_scale": ([_vp, C.c_float], C.c_int),
    "impec_cnn_forward": ([_vp, _fp, C.c_int, C.c_int, C.c_int, _fp], C.c_int),
    "impec_cnn_gradients": ([_vp, _fp, _fp, C.c_int, C.c_int, C.c_int, _fp, C.POINTER(C.c_double)], C.c_int),
    "impec_cnn_collect": ([_vp, _vp, C.c_int], C.c_int),
    "impec_cnn_pairs": ([_vp, C.POINTER(C.c_int)], C.c_int),
    "impec_cnn_train": ([_vp, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_double)], C.c_int),
    "impec_set_warmstart": ([_vp, _vp], C.c_int),
    "impec_dispersion_cn": ([_vp, C.c_double, _ip], C.c_int),
    "impec_hyperbolic_solve": ([_vp, C.c_double], C.c_int),
    "impec_step": ([_vp, C.c_double, C.c_double, _dp], C.c_int),
    "impec_total_mass": ([_vp, _dp], C.c_int),
    "impec_mass_error": ([_vp, C.c_double, _dp], C.c_int),
    "impec_divergence_error": ([_vp, _dp], C.c_int),
    "impec_velocity_boundary_error": ([_vp, _dp], C.c_int),
    "impec_minmax_concentration": ([_vp, _dp, _dp], C.c_int),
    "impec_get_stats": ([_vp, C.POINTER(Stats)], C.c_int),
    "impec_pressure_operator": ([_vp, _dp], C.c_int),
    "impec_parabolic_operator": ([_vp, C.c_double, _dp, _dp], C.c_int),
    "impec_profile_enable": ([_vp, C.c_int], C.c_int),
    "impec_profile_read": ([_vp, C.c_char_p, C.POINTER(C.c_longlong), _dp], C.c_int),
    "impec_profile_reset": ([_vp], C.c_int),
    "impec_copy_calibration": ([C.c_longlong, C.c_int, _dp], C.c_int),
    "impec_ensemble_create": ([C.POINTER(Params), C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "impec_ensemble_destroy": ([_vp], C.c_int),
    "impec_ensemble_last_error": ([_vp], C.c_char_p),
    "impec_ensemble_member": ([_vp, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "impec_ensemble_set_graphs": ([_vp, C.c_int, C.c_int], C.c_int),
    "impec_ensemble_autotune": ([_vp, C.c_int, _ip], C.c_int),
    "impec_ensemble_step": ([_vp, _dp, _ip], C.c_int),
}

_lib = None


def load(path=None):
    """Loads the in-tree libimpec_hip.so (or the build at `path`, which must
    then be passed explicitly on the first call); raises if it is not built."""
    global _lib
    if _lib is None:
        path = path or LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError("%s is not built; run `python -m impec2d_amd.build`" % path)
        lib = C.CDLL(path)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


class CnnConfig(C.Structure):
    _fields_ = [("layers", C.c_int), ("channels", C.c_int), ("capacity", C.c_int), ("crop", C.c_int),
                ("seed", C.c_ulonglong), ("mse_loss", C.c_int), ("reserved", C.c_int * 3)]


class Network:
    """The learned-initial-guess CNN (impec_cnn_* entry points)."""

    def __init__(self, layers=8, capacity=4096, crop=64, seed=1234, mse_loss=False):
        self.lib = load()
        cfg = CnnConfig()
        self.lib.impec_cnn_config_default(C.byref(cfg))
        cfg.layers, cfg.capacity, cfg.crop, cfg.seed = layers, capacity, crop, seed
        cfg.mse_loss = 1 if mse_loss else 0
        h = C.c_void_p()
        st = self.lib.impec_cnn_create(C.byref(cfg), C.byref(h))
        if st != IMPEC_OK:
            raise ImpecError(st, "impec_cnn_create")
        self.h = h
        self.layers = layers
        n = C.c_longlong(0)
        self.lib.impec_cnn_num_params(self.h, C.byref(n))
        self.nparam = n.value

    def _check(self, st):
        if st != IMPEC_OK:
            raise ImpecError(st, self.lib.impec_cnn_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.impec_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        self._check(self.lib.impec_cnn_get_params(self.h, out.ctypes.data_as(_fp)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self._check(self.lib.impec_cnn_set_params(self.h, p.ctypes.data_as(_fp)))

    def scale(self):
        v = C.c_float(0)
        self._check(self.lib.impec_cnn_get_scale(self.h, C.byref(v)))
        return v.value

    def set_scale(self, v):
        self._check(self.lib.impec_cnn_set_scale(self.h, float(v)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        n, hh, ww, _ = x.shape
        y = np.zeros((n, hh, ww), dtype=np.float32)
        self._check(self.lib.impec_cnn_forward(self.h, x.ctypes.data_as(_fp), n, hh, ww, y.ctypes.data_as(_fp)))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        n, hh, ww, _ = x.shape
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_double(0)
        self._check(self.lib.impec_cnn_gradients(self.h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), n, hh, ww,
                                                 g.ctypes.data_as(_fp), C.byref(loss)))
        return g, loss.value

    def collect(self, ctx, crops):
        self._check(self.lib.impec_cnn_collect(self.h, ctx.h, int(crops)))

    def pairs(self):
        v = C.c_int(0)
        self._check(self.lib.impec_cnn_pairs(self.h, C.byref(v)))
        return v.value

    def train(self, iters, batch=16, lr=1e-3):
        loss = C.c_double(0)
        self._check(self.lib.impec_cnn_train(self.h, int(iters), int(batch), float(lr), C.byref(loss)))
        return loss.value


class ImpecError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (load().impec_strerror(code).decode(), code, msg))
        self.code = code


def default_params(**kw):
    p = Params()
    load().impec_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k) or k.startswith("reserved"):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Context:
    """One device context (impec_ctx) with the reference's host layouts."""

    def __init__(self, params):
        self.lib = load()
        self.params = params
        h = C.c_void_p()
        st = self.lib.impec_create(C.byref(params), C.byref(h))
        if st != IMPEC_OK:
            raise ImpecError(st, "impec_create failed")
        self.h = h
        self.N1, self.N2, self.b = params.N1, params.N2, params.num_ghostcells

    def close(self):
        if self.h:
            self.lib.impec_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != IMPEC_OK:
            raise ImpecError(st, self.lib.impec_last_error(self.h).decode())
        return st

    # ---- shapes
    def scalar_shape(self):
        return (self.N1 + 2 * self.b, self.N2 + 2 * self.b)

    def velocity_shape(self):
        return (self.N1 + 1 + 2 * self.b, self.N2 + 1 + 2 * self.b, 2)

    # ---- transfers
    def set_concentration(self, c):
        c = np.ascontiguousarray(c, dtype=np.float64)
        assert c.shape == self.scalar_shape()
        self._check(self.lib.impec_set_concentration(self.h, _ptr(c)))

    def get_concentration(self):
        c = np.empty(self.scalar_shape())
        self._check(self.lib.impec_get_concentration(self.h, _ptr(c)))
        return c

    def set_pressure(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert p.shape == self.scalar_shape()
        self._check(self.lib.impec_set_pressure(self.h, _ptr(p)))

    def get_pressure(self):
        p = np.empty(self.scalar_shape())
        self._check(self.lib.impec_get_pressure(self.h, _ptr(p)))
        return p

    def set_velocity(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.shape == self.velocity_shape()
        self._check(self.lib.impec_set_velocity(self.h, _ptr(v)))

    def get_velocity(self):
        v = np.empty(self.velocity_shape())
        self._check(self.lib.impec_get_velocity(self.h, _ptr(v)))
        return v

    def set_source(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        assert q.size == self.N1 * self.N2
        self._check(self.lib.impec_set_source(self.h, _ptr(q)))

    def set_permeability(self, K11=None, K22=None):
        a = None if K11 is None else np.ascontiguousarray(K11, dtype=np.float64)
        b = None if K22 is None else np.ascontiguousarray(K22, dtype=np.float64)
        self._check(self.lib.impec_set_permeability(self.h, None if a is None else _ptr(a),
                                                    None if b is None else _ptr(b)))

    def set_warmstart(self, net):
        """Attach the network used by initial_guess = GUESS_LEARNED (None detaches)."""
        self._check(self.lib.impec_set_warmstart(self.h, None if net is None else net.h))
        self._net = net

    def set_pin_kappa(self, kappa):
        self._check(self.lib.impec_set_pin_kappa(self.h, kappa))

    # ---- stages
    def fill_boundaries(self):
        self._check(self.lib.impec_fill_boundaries(self.h))

    def pressure_solve(self):
        its = C.c_int(0)
        rr = C.c_double(0)
        self._check(self.lib.impec_pressure_solve(self.h, C.byref(its), C.byref(rr)))
        return its.value, rr.value

    def velocities(self):
        self._check(self.lib.impec_velocities(self.h))

    def max_stable_timestep(self):
        dt = C.c_double(0)
        self._check(self.lib.impec_max_stable_timestep(self.h, C.byref(dt)))
        return dt.value

    def source_split(self, dt):
        self._check(self.lib.impec_source_split(self.h, dt))

    def advect_split(self, dt):
        self._check(self.lib.impec_advect_split(self.h, dt))

    def advect_x(self, dt):
        self._check(self.lib.impec_advect_x(self.h, dt))

    def advect_y(self, dt):
        self._check(self.lib.impec_advect_y(self.h, dt))

    def intercell_flux(self, direction, dt):
        """F (direction 0, shape (N1+1, N2+4)) or G (direction 1, shape (N1+4, N2+1))."""
        shape = (self.N1 + 1, self.N2 + 4) if direction == 0 else (self.N1 + 4, self.N2 + 1)
        out = np.zeros(shape)
        self._check(self.lib.impec_intercell_flux(self.h, int(direction), dt, out.ctypes.data_as(_dp)))
        return out

    def dispersion_cn(self, dt):
        its = C.c_int(0)
        self._check(self.lib.impec_dispersion_cn(self.h, dt, C.byref(its)))
        return its.value

    def hyperbolic_solve(self, dt):
        self._check(self.lib.impec_hyperbolic_solve(self.h, dt))

    def step(self, t=0.0, t_next=-1.0):
        dt = C.c_double(0)
        self._check(self.lib.impec_step(self.h, t, t_next, C.byref(dt)))
        return dt.value

    def synchronize(self):
        self._check(self.lib.impec_synchronize(self.h))

    # ---- diagnostics
    def _scalar(self, fn, *args):
        out = C.c_double(0)
        self._check(fn(self.h, *args, C.byref(out)))
        return out.value

    def total_mass(self):
        return self._scalar(self.lib.impec_total_mass)

    def mass_error(self, t):
        return self._scalar(self.lib.impec_mass_error, C.c_double(t))

    def divergence_error(self):
        return self._scalar(self.lib.impec_divergence_error)

    def velocity_boundary_error(self):
        return self._scalar(self.lib.impec_velocity_boundary_error)

    def minmax_concentration(self):
        a, b = C.c_double(0), C.c_double(0)
        self._check(self.lib.impec_minmax_concentration(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def stats(self):
        s = Stats()
        self._check(self.lib.impec_get_stats(self.h, C.byref(s)))
        return s

    def pressure_operator(self):
        A = np.empty(9 * self.N1 * self.N2)
        self._check(self.lib.impec_pressure_operator(self.h, _ptr(A)))
        return A

    def parabolic_operator(self, dt):
        A = np.empty(9 * self.N1 * self.N2)
        r = np.empty(self.N1 * self.N2)
        self._check(self.lib.impec_parabolic_operator(self.h, dt, _ptr(A), _ptr(r)))
        return A, r

    # ---- measurement
    def profile(self, on=True):
        self._check(self.lib.impec_profile_enable(self.h, int(on)))

    def profile_reset(self):
        self._check(self.lib.impec_profile_reset(self.h))

    def profile_read(self, name):
        n = C.c_longlong(0)
        ms = C.c_double(0)
        self.lib.impec_profile_read(self.h, name.encode(), C.byref(n), C.byref(ms))
        return n.value, ms.value


class _Member(Context):
    """A member context lent by an ensemble (not destroyed by close())."""

    def __init__(self, ens, handle, params):
        self.lib = ens.lib
        self.params = params
        self.h = handle
        self.N1, self.N2, self.b = params.N1, params.N2, params.num_ghostcells
        self._ens = ens

    def close(self):
        self.h = None


class Ensemble:
    """Independent problems of different sizes advanced together
    (impec_ensemble): one launch per pass over all members, graph replays per
    step.  ``members[k]`` is member k's Context (source, permeability,
    concentration, fields, diagnostics)."""

    def __init__(self, params_list):
        self.lib = load()
        n = len(params_list)
        arr = (Params * n)(*params_list)
        h = C.c_void_p()
        st = self.lib.impec_ensemble_create(arr, n, C.byref(h))
        if st != IMPEC_OK:
            raise ImpecError(st, "impec_ensemble_create failed")
        self.h = h
        self.n = n
        self.members = []
        for k in range(n):
            mh = C.c_void_p()
            self._check(self.lib.impec_ensemble_member(self.h, k, C.byref(mh)))
            self.members.append(_Member(self, mh, params_list[k]))

    def _check(self, st):
        if st != IMPEC_OK:
            raise ImpecError(st, self.lib.impec_ensemble_last_error(self.h).decode())
        return st

    def close(self):
        if self.h:
            for m in self.members:
                m.close()
            self.lib.impec_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graphs(self, on, slots=0):
        self._check(self.lib.impec_ensemble_set_graphs(self.h, int(bool(on)), int(slots)))

    def autotune(self, reps=3):
        rows = np.zeros(self.n, dtype=np.int32)
        self._check(self.lib.impec_ensemble_autotune(self.h, int(reps), rows.ctypes.data_as(_ip)))
        return rows

    def step(self):
        """One IMPEC step of every member; returns (dt, its) arrays."""
        dt = np.zeros(self.n)
        its = np.zeros(self.n, dtype=np.int32)
        self._check(self.lib.impec_ensemble_step(self.h, dt.ctypes.data_as(_dp), its.ctypes.data_as(_ip)))
        return dt, its
