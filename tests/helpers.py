"""Shared set-up for the parity tests: one parameter set drives both the
oracle (oracle/pyoracle.py) and the HIP context (impec2d_amd/capi.py)."""
import numpy as np

from oracle import pyoracle as po

# oracle name -> C-ABI name
_MAP = {"N1": "N1", "N2": "N2", "b": "num_ghostcells", "stype": "source_type",
        "avgproc": "average_procedure", "dim_split": "dim_splitting", "hyper": "hyper_method",
        "diffusion": "diffusion", "para_cn": "para_method", "is_test": "is_test",
        "alpha_abs": "alpha_absolute", "max_iter": "max_iterations",
        "x_lb": "x_lb", "x_ub": "x_ub", "y_lb": "y_lb", "y_ub": "y_ub", "nu": "nu",
        "mu_res": "mu_res", "mu_inj": "mu_inj", "porosity": "porosity", "cfl": "cfl",
        "inj_fluid_conc": "inj_fluid_conc", "alpha_l": "alpha_l", "alpha_t": "alpha_t",
        "dt_diff_limit": "dt_diff_limiter", "d_m": "d_m", "eps": "eps", "para_eps": "para_eps",
        "I": "I", "tstart": "tstart"}


def params(**kw):
    """Oracle parameter dict; defaults are the C1 quarter-five-spot case."""
    return po.default_params(**kw)


def gpu_params(p, **extra):
    from impec2d_amd import capi
    g = capi.default_params()
    for k, v in p.items():
        if k in _MAP:
            setattr(g, _MAP[k], v)
    for k, v in extra.items():
        setattr(g, k, v)
    return g


def pair(p, K11=None, K22=None, **extra):
    from impec2d_amd import capi
    o = po.Oracle(p, K11, K22)
    g = capi.Context(gpu_params(p, **extra))
    if K11 is not None:
        g.set_permeability(K11, K22)
    return o, g


def random_c(o, seed=0, lo=0.0, hi=1.0):
    rng = np.random.default_rng(seed)
    c = o.scalar()
    o.interior(c)[:] = rng.uniform(lo, hi, (o.N1, o.N2))
    o.fill(c)
    return c


def smooth_c(o, seed=0):
    """A front-like concentration field in [0, 1]."""
    rng = np.random.default_rng(seed)
    c = o.scalar()
    x = (np.arange(o.N1) + 0.5) / o.N1
    y = (np.arange(o.N2) + 0.5) / o.N2
    X, Y = np.meshgrid(x, y, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2)
    f = 0.5 * (1 - np.tanh((r - 0.4 - 0.05 * np.sin(7 * np.arctan2(Y, X + 1e-9))) / 0.03))
    o.interior(c)[:] = np.clip(f + 0.01 * rng.standard_normal(f.shape), 0, 1)
    o.fill(c)
    return c


def rel_err(a, b):
    a = np.asarray(a); b = np.asarray(b)
    s = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / s
