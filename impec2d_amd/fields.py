"""Synthetic input fields for the lognormal-permeability workloads.

``lognormal_permeability`` draws ln K as a stationary Gaussian random field
with mean 0, standard deviation ``sigma`` and squared-exponential covariance
C(h) = exp(-(|h| / corr_len)^2) on the cell centres of the scalar layout
(N1 + 2b) x (N2 + 2b) (ghost cells included, as impec_set_permeability
expects), by circulant embedding: the covariance is laid out on a periodic
grid of twice the extent, its FFT gives the embedding's eigenvalues (small
negative ones, from the truncated tail, are set to zero), and one complex
Gaussian draw scaled by their square roots yields two independent fields, of
which the real part is used.  Host-side set-up only; numpy's PCG64 generator
with the given seed makes the field reproducible.
"""
import functools

import numpy as np

try:
    import scipy.fft as _fft
    _FFT_KW = {"workers": -1}
except ImportError:   # numpy's FFT gives the same field to rounding
    _fft = np.fft
    _FFT_KW = {}


@functools.lru_cache(maxsize=16)
def _embedding_scale(m1, m2, dx, dy, corr_len):
    """sqrt(eigenvalues / (m1 m2)) of the circulant embedding of the covariance
    on the periodic m1 x m2 grid (negative eigenvalues clipped to zero)."""
    i = np.arange(m1)
    j = np.arange(m2)
    hx = np.minimum(i, m1 - i) * dx
    hy = np.minimum(j, m2 - j) * dy
    cov = np.exp(-((hx[:, None] ** 2 + hy[None, :] ** 2) / corr_len ** 2))
    lam = _fft.fft2(cov, **_FFT_KW).real
    lam = np.maximum(lam, 0.0)
    out = np.sqrt(lam / (m1 * m2))
    out.setflags(write=False)
    return out


def gaussian_random_field(n1, n2, dx, dy, corr_len, seed):
    """Zero-mean, unit-variance GRF with covariance exp(-(h/corr_len)^2) on an
    n1 x n2 grid of spacing (dx, dy)."""
    m1, m2 = 2 * n1, 2 * n2
    scale = _embedding_scale(m1, m2, float(dx), float(dy), float(corr_len))
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal((m1, m2)) + 1j * rng.standard_normal((m1, m2))
    xi *= scale
    f = _fft.fft2(xi, overwrite_x=True, **_FFT_KW) if _FFT_KW else _fft.fft2(xi)
    return np.ascontiguousarray(f.real[:n1, :n2])


def lognormal_permeability(N1, N2, b=2, sigma=1.0, corr_len=0.05, seed=0, x_extent=1.0, y_extent=1.0):
    """K = exp(sigma * g) on the ghosted scalar layout (N1 + 2b) x (N2 + 2b)."""
    dx, dy = x_extent / N1, y_extent / N2
    g = gaussian_random_field(N1 + 2 * b, N2 + 2 * b, dx, dy, corr_len, seed)
    return np.exp(sigma * g)
