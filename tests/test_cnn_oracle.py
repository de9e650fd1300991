"""The CPU oracle of the learned-guess network (oracle/cnn_oracle.py), pinned
without the GPU: BF16 rounding on known values, the convolution against a
direct loop, and the FP64 back-propagation against central finite
differences of the loss."""
import numpy as np

from oracle import cnn_oracle as co


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7, -2.5, 0.0, 65504.0], dtype=np.float32)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    # 1 + 3 * 2^-9 is above the midpoint -> 1 + 2^-7
    expect = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -7, -2.5, 0.0, 65536.0], dtype=np.float32)
    assert np.array_equal(co.bf16(x), expect)


def test_conv_against_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 7, 3))
    W = rng.standard_normal((9, 3, 4))
    y = co.conv(x, W)
    for n, i, j in [(0, 0, 0), (1, 4, 6), (0, 2, 3), (1, 0, 6)]:
        s = np.zeros(4)
        for tap in range(9):
            ii, jj = i + tap // 3 - 1, j + tap % 3 - 1
            if 0 <= ii < 5 and 0 <= jj < 7:
                s += x[n, ii, jj] @ W[tap]
        assert np.allclose(y[n, i, j], s, rtol=1e-13, atol=1e-13)


import pytest


@pytest.mark.parametrize("kind", ["mse", "h1"])
def test_backprop_against_finite_differences(kind):
    L = 4
    rng = np.random.default_rng(1)
    _, _, npar = co.layout(L)
    params = rng.standard_normal(npar) * 0.3
    x = rng.standard_normal((2, 6, 5, 3))
    t = rng.standard_normal((2, 6, 5))
    g, loss = co.gradients(params, x, t, L, bf16_mode=False, kind=kind)
    assert abs(loss - co.loss_only(params, x, t, L, kind)) < 1e-12
    idx = rng.choice(npar, 40, replace=False)
    h = 1e-6
    for k in idx:
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        fd = (co.loss_only(pp, x, t, L, kind) - co.loss_only(pm, x, t, L, kind)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
