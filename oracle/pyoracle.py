"""ctypes front end of the CPU oracle (oracle/impec_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the impec2d_amd package.

Arrays use the reference's host layouts (GenSimDefs.h:22-31,
Impec2DSolver.cpp:151-182): scalar fields (N1+2b, N2+2b) with base -b,
velocity (N1+1+2b, N2+1+2b, 2), q flat (N1*N2,) indexed i*N2+j.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_INT_FIELDS = ["N1", "N2", "b", "stype", "avgproc", "dim_split", "hyper", "diffusion",
               "para_cn", "is_test", "testno", "alpha_abs", "precond", "max_iter"]
_DBL_FIELDS = ["x_lb", "x_ub", "y_lb", "y_ub", "nu", "mu_res", "mu_inj", "porosity", "cfl",
               "inj_fluid_conc", "alpha_l", "alpha_t", "dt_diff_limit", "d_m", "eps", "para_eps",
               "I", "well_rc", "tstart", "tfinal"]


class OrParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in _INT_FIELDS] + [(n, C.c_double) for n in _DBL_FIELDS]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        L.or_ctx_size.restype = C.c_size_t
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_threads.restype = C.c_int
        L.or_set_mutation.argtypes = [C.c_int]
        L.or_set_mutation.restype = C.c_int
        L.or_params_size.restype = C.c_size_t
        L.or_init.argtypes = [vp, C.POINTER(OrParams), vp, vp]
        for fn in ["or_fill_neumann_scalar", "or_fill_periodic_scalar", "or_fill_boundaries",
                   "or_fill_neumann_velocity"]:
            getattr(L, fn).argtypes = [vp, dp]
        L.or_mu.argtypes = [vp, C.c_double]
        L.or_mu.restype = C.c_double
        L.or_assemble_elliptic.argtypes = [vp, dp, dp]
        L.or_matvec9.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.or_dense_solve9.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.or_pressure_solve.argtypes = [vp, dp, dp, dp, C.c_int]
        L.or_pressure_solve_dense.argtypes = [vp, dp, dp, dp]
        L.or_calc_velocities.argtypes = [vp, dp, dp, dp]
        for fn in ["or_max_stable_timestep", "or_max_stable_timestep_combined",
                   "or_max_stable_advection_timestep"]:
            getattr(L, fn).argtypes = [vp, dp]
            getattr(L, fn).restype = C.c_double
        L.or_tr_integrate.argtypes = [vp, dp, dp, C.c_double]
        L.or_tr_integrate.restype = C.c_double
        L.or_source_split.argtypes = [vp, dp, dp, C.c_double]
        L.or_flux_F.argtypes = [vp, dp, dp, dp, C.c_double]
        L.or_flux_G.argtypes = [vp, dp, dp, dp, C.c_double]
        L.or_advect_x.argtypes = [vp, dp, dp, C.c_double]
        L.or_advect_y.argtypes = [vp, dp, dp, C.c_double]
        L.or_advect_split.argtypes = [vp, dp, dp, C.c_double]
        L.or_assemble_parabolic.argtypes = [vp, dp, C.c_double, dp]
        L.or_rhs_parabolic.argtypes = [vp, dp, dp, C.c_double, dp]
        L.or_parabolic_solve.argtypes = [vp, dp, dp, C.c_double]
        L.or_hyperbolic_solve.argtypes = [vp, dp, dp, dp, C.c_double]
        L.or_total_mass.argtypes = [vp, dp]
        L.or_total_mass.restype = C.c_double
        L.or_mass_error.argtypes = [vp, dp, C.c_double]
        L.or_mass_error.restype = C.c_double
        L.or_divergence_error.argtypes = [vp, dp, dp]
        L.or_divergence_error.restype = C.c_double
        L.or_velocity_boundary_error.argtypes = [vp, dp]
        L.or_velocity_boundary_error.restype = C.c_double
        L.or_impec_step.argtypes = [vp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_int,
                                    C.POINTER(C.c_double)]
        L.or_init_one_cell.argtypes = [vp, dp]
        L.or_init_circle.argtypes = [vp, dp]
        L.or_exact_vol_frac_circle.argtypes = [C.c_double] * 5
        L.or_exact_vol_frac_circle.restype = C.c_double
        L.or_set_kappa.argtypes = [vp, C.c_double]
        L.or_total_mass_outflux.argtypes = [vp]
        L.or_total_mass_outflux.restype = C.c_double
        for fn in ["or_last_its", "or_last_para_its", "or_last_vcycles"]:
            getattr(L, fn).argtypes = [vp]
            getattr(L, fn).restype = C.c_int
        L.or_last_relres.argtypes = [vp]
        L.or_last_relres.restype = C.c_double
        _lib = L
    return _lib


def default_params(**kw):
    """Parameter defaults of the shipped example configs (examples/parallel.cfg)."""
    p = dict(N1=64, N2=64, b=2, stype=1, avgproc=0, dim_split=4, hyper=0, diffusion=0, para_cn=1,
             is_test=0, testno=1, alpha_abs=0, precond=0, max_iter=10000,
             x_lb=0.0, x_ub=1.0, y_lb=0.0, y_ub=1.0, nu=1.0, mu_res=1.0, mu_inj=0.1,
             porosity=1.0, cfl=0.5, inj_fluid_conc=1.0, alpha_l=1.0, alpha_t=1.0,
             dt_diff_limit=0.9, d_m=0.0, eps=1e-14, para_eps=1e-14, I=1.0, well_rc=0.05,
             tstart=0.0, tfinal=1e30)
    for k, v in kw.items():
        if k not in p:
            raise KeyError(k)
        p[k] = v
    return p


class Oracle:
    """One oracle context: parameters, optional K arrays and solver state."""

    def __init__(self, params, K11=None, K22=None):
        L = lib()
        self.p = dict(params)
        self._P = OrParams(**self.p)
        self._buf = C.create_string_buffer(L.or_ctx_size())
        self.ctx = C.cast(self._buf, C.c_void_p)
        self.N1, self.N2, self.b = self.p["N1"], self.p["N2"], self.p["b"]
        self._K11 = None if K11 is None else np.ascontiguousarray(K11, dtype=np.float64)
        self._K22 = None if K22 is None else np.ascontiguousarray(K22, dtype=np.float64)
        k11 = None if self._K11 is None else self._K11.ctypes.data
        k22 = None if self._K22 is None else self._K22.ctypes.data
        if L.or_init(self.ctx, C.byref(self._P), k11, k22) != 0:
            raise ValueError("or_init rejected the parameters")
        self.L = L
        self.dx = (self.p["x_ub"] - self.p["x_lb"]) / self.N1
        self.dy = (self.p["y_ub"] - self.p["y_lb"]) / self.N2

    # --- array factories -------------------------------------------------
    def scalar(self):
        return np.zeros((self.N1 + 2 * self.b, self.N2 + 2 * self.b))

    def velocity(self):
        return np.zeros((self.N1 + 1 + 2 * self.b, self.N2 + 1 + 2 * self.b, 2))

    def interior(self, a):
        b = self.b
        return a[b:b + self.N1, b:b + self.N2]

    # --- set-up ----------------------------------------------------------
    def source(self, profile="one_cell"):
        q = np.zeros(self.N1 * self.N2)
        if profile == "one_cell":
            self.L.or_init_one_cell(self.ctx, q)
        elif profile == "circle":
            self.L.or_init_circle(self.ctx, q)
        else:
            raise ValueError(profile)
        return q

    # --- pieces of the step -----------------------------------------------
    def fill(self, a):
        self.L.or_fill_boundaries(self.ctx, a)

    def fill_neumann(self, a):
        self.L.or_fill_neumann_scalar(self.ctx, a)

    def fill_velocity(self, v):
        self.L.or_fill_neumann_velocity(self.ctx, v)

    def assemble(self, c):
        A = np.zeros(9 * self.N1 * self.N2)
        self.L.or_assemble_elliptic(self.ctx, c, A)
        return A

    def pressure_solve(self, c, p, q, warm=False):
        return self.L.or_pressure_solve(self.ctx, c, p, q, int(warm))

    def pressure_solve_dense(self, c, p, q):
        return self.L.or_pressure_solve_dense(self.ctx, c, p, q)

    def velocities(self, p, v, c):
        self.L.or_calc_velocities(self.ctx, p, v, c)

    def timestep(self, v):
        return self.L.or_max_stable_timestep(self.ctx, v)

    def tr(self, c, q, dt):
        return self.L.or_tr_integrate(self.ctx, c, q, dt)

    def flux(self, direction, c, v, dt):
        """F (direction 0, (N1+1) x (N2+2b)) or G (direction 1, (N1+2b) x (N2+1))."""
        b = self.b
        shape = (self.N1 + 1, self.N2 + 2 * b) if direction == 0 else (self.N1 + 2 * b, self.N2 + 1)
        out = np.zeros(shape)
        (self.L.or_flux_F if direction == 0 else self.L.or_flux_G)(self.ctx, out, c, v, dt)
        return out

    def advect_x(self, c, v, dt):
        self.L.or_advect_x(self.ctx, c, v, dt)

    def advect_y(self, c, v, dt):
        self.L.or_advect_y(self.ctx, c, v, dt)

    def advect_split(self, c, v, dt):
        self.L.or_advect_split(self.ctx, c, v, dt)

    def parabolic_matrix(self, v, dt):
        A = np.zeros(9 * self.N1 * self.N2)
        self.L.or_assemble_parabolic(self.ctx, v, dt, A)
        return A

    def parabolic_rhs(self, c, v, dt):
        r = np.zeros(self.N1 * self.N2)
        self.L.or_rhs_parabolic(self.ctx, c, v, dt, r)
        return r

    def parabolic_solve(self, c, v, dt):
        return self.L.or_parabolic_solve(self.ctx, c, v, dt)

    def hyperbolic_solve(self, c, v, q, dt):
        return self.L.or_hyperbolic_solve(self.ctx, c, v, q, dt)

    def step(self, c, p, v, q, t=0.0, tnext=-1.0, warm=False):
        dt = C.c_double(0)
        st = self.L.or_impec_step(self.ctx, c, p, v, q, t, tnext, int(warm), C.byref(dt))
        return st, dt.value

    # --- diagnostics -------------------------------------------------------
    def mass(self, c):
        return self.L.or_total_mass(self.ctx, c)

    def mass_error(self, c, t):
        return self.L.or_mass_error(self.ctx, c, t)

    def div_error(self, v, q):
        return self.L.or_divergence_error(self.ctx, v, q)

    def noflow_error(self, v):
        return self.L.or_velocity_boundary_error(self.ctx, v)

    @property
    def total_mass_outflux(self):
        return self.L.or_total_mass_outflux(self.ctx)

    @property
    def last_its(self):
        return self.L.or_last_its(self.ctx)

    @property
    def last_para_its(self):
        return self.L.or_last_para_its(self.ctx)

    @property
    def last_relres(self):
        return self.L.or_last_relres(self.ctx)

    @property
    def last_vcycles(self):
        return self.L.or_last_vcycles(self.ctx)

    def set_kappa(self, kappa):
        self.L.or_set_kappa(self.ctx, kappa)


def set_threads(n):
    """OpenMP threads of the oracle's per-cell loops (results do not depend on
    it); returns the count in effect."""
    return lib().or_set_threads(int(n))


# planted mistakes of the oracle (impec_oracle.c, or_set_mutation): tests use
# them to show that their pins fail on a wrong formula
MUTATIONS = {"cn_swap_alpha": 1, "cn_drop_phi": 2, "cn_avg_f": 3, "cn_avg_g": 4, "cn_dm_cross": 5,
             "cn_rhs_cross": 6, "split_xfirst": 7, "split_yfirst": 8, "split_strang": 9, "split_alt": 10,
             "split_parallel": 11}


class mutated:
    """Context manager: the oracle with one planted mistake (TEST ONLY)."""

    def __init__(self, name):
        self.m = MUTATIONS[name]

    def __enter__(self):
        assert lib().or_set_mutation(self.m) == 0
        return self

    def __exit__(self, *exc):
        lib().or_set_mutation(0)
        return False


def matvec9(N1, N2, A, x):
    y = np.zeros(N1 * N2)
    lib().or_matvec9(N1, N2, np.ascontiguousarray(A), np.ascontiguousarray(x), y)
    return y


def dense_solve9(N1, N2, A, b):
    x = np.zeros(N1 * N2)
    st = lib().or_dense_solve9(N1, N2, np.ascontiguousarray(A), np.ascontiguousarray(b), x)
    if st != 0:
        raise RuntimeError("dense solve failed: %d" % st)
    return x


def exact_vol_frac_circle(x1, y1, r, dx, dy):
    return lib().or_exact_vol_frac_circle(x1, y1, r, dx, dy)
