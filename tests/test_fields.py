"""The seeded lognormal-permeability generator (impec2d_amd/fields.py): the
statistics it promises, checked on samples (mean 0, variance 1, squared-
exponential correlation exp(-(h/l)^2))."""
import numpy as np

from impec2d_amd import fields


def test_grf_moments_and_correlation():
    n, l = 256, 0.05
    dx = 1.0 / n
    vals, corr = [], {1: [], 6: [], 13: []}
    for seed in range(6):
        g = fields.gaussian_random_field(n, n, dx, dx, l, seed)
        vals.append(g)
        for lag in corr:
            corr[lag].append(np.mean(g[:, :-lag] * g[:, lag:]))
    allv = np.concatenate([v.ravel() for v in vals])
    assert abs(allv.mean()) < 0.1
    assert abs(allv.var() - 1.0) < 0.15
    for lag, cs in corr.items():
        expect = np.exp(-((lag * dx) / l) ** 2)
        assert abs(np.mean(cs) / allv.var() - expect) < 0.08, (lag, np.mean(cs), expect)


def test_lognormal_permeability_layout_and_seed():
    K = fields.lognormal_permeability(40, 24, b=2, sigma=1.0, seed=3)
    assert K.shape == (44, 28) and np.all(K > 0)
    assert np.array_equal(K, fields.lognormal_permeability(40, 24, b=2, sigma=1.0, seed=3))
    assert not np.array_equal(K, fields.lognormal_permeability(40, 24, b=2, sigma=1.0, seed=4))
    # sigma scales ln K
    K2 = fields.lognormal_permeability(40, 24, b=2, sigma=2.0, seed=3)
    assert np.allclose(np.log(K2), 2.0 * np.log(K))
