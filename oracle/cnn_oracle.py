"""CPU oracle of the learned-initial-guess network (test infrastructure: only
tests import it).

The reference has no counterpart (its pressure solve starts from zero,
MGSolver.cpp:573-590); this file restates, in numpy, the network that
impec2d_amd/csrc/k_cnn.hip computes, so that the GPU forward pass and the
GPU gradients can be checked:

  layer 0:        a0 = relu(conv3x3(x, W0) + b0)            3 -> 32
  layers 1..L-2:  a_l = relu(conv3x3(a_{l-1}, W_l) + b_l)    32 -> 32
  layer L-1:      y = conv3x3(a_{L-2}, W_{L-1}) + b_{L-1}    32 -> 1
  loss = mean((y - t)^2) ("mse"), or the gradient mismatch of y - t
         (squared jumps across inner faces) + 0.01 x its L2 norm ("h1")

conv3x3: y[p][co] = sum_{tap, ci} W[tap][ci][co] x[p + off(tap)][ci],
off(tap) = (tap // 3 - 1, tap % 3 - 1), zero padding; NHWC tensors.

With ``bf16=True`` the values the GPU stores in BF16 are rounded the same way
(round to nearest even): the mid-layer weights, every activation a_l and
every back-propagated activation gradient; sums are float64 here (float32 on
the GPU).  With ``bf16=False`` everything is float64, and that mode is pinned
against finite differences of the loss (tests/test_cnn_oracle.py).
"""
import numpy as np

CH = 32
TAPS = 9


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even); returns float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def layout(L):
    """(cin, cout) of every layer and the offsets of W_l and b_l in the flat
    parameter vector (W[tap][ci][co] then b[co] per layer)."""
    shapes, offs, off = [], [], 0
    for l in range(L):
        cin = 3 if l == 0 else CH
        cout = 1 if l == L - 1 else CH
        shapes.append((cin, cout))
        offs.append((off, off + TAPS * cin * cout))
        off += TAPS * cin * cout + cout
    return shapes, offs, off


def unpack(params, L):
    shapes, offs, _ = layout(L)
    Ws, bs = [], []
    for (cin, cout), (wo, bo) in zip(shapes, offs):
        Ws.append(np.asarray(params[wo:bo], dtype=np.float64).reshape(TAPS, cin, cout))
        bs.append(np.asarray(params[bo:bo + cout], dtype=np.float64))
    return Ws, bs


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))


def conv(x, W):
    """NHWC 3x3 convolution, zero padding (float64)."""
    n, h, w, _ = x.shape
    xp = _pad(x)
    y = np.zeros((n, h, w, W.shape[2]))
    for tap in range(TAPS):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        y += xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w, :] @ W[tap]
    return y


def conv_back_data(dz, W):
    """dx[q][ci] = sum_{tap, co} W[tap][ci][co] dz[q - off(tap)][co]."""
    n, h, w, _ = dz.shape
    dzp = _pad(dz)
    dx = np.zeros((n, h, w, W.shape[1]))
    for tap in range(TAPS):
        dy, dxx = tap // 3 - 1, tap % 3 - 1
        dx += dzp[:, 1 - dy:1 - dy + h, 1 - dxx:1 - dxx + w, :] @ W[tap].T
    return dx


def conv_back_weights(x, dz):
    """dW[tap][ci][co] = sum_p x[p + off(tap)][ci] dz[p][co]."""
    n, h, w, cin = x.shape
    xp = _pad(x)
    dW = np.zeros((TAPS, cin, dz.shape[3]))
    dzf = dz.reshape(-1, dz.shape[3])
    for tap in range(TAPS):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        dW[tap] = xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w, :].reshape(-1, cin).T @ dzf
    return dW


def _r(x, on):
    return bf16(x).astype(np.float64) if on else x


def forward(params, x, L, bf16_mode=True):
    """Returns y (n,h,w) and the activations a_0 .. a_{L-2}."""
    Ws, bs = unpack(params, L)
    acts = []
    a = np.maximum(conv(np.asarray(x, np.float64), Ws[0]) + bs[0], 0.0)
    a = _r(a, bf16_mode)
    acts.append(a)
    for l in range(1, L - 1):
        W = _r(Ws[l], bf16_mode)
        a = _r(np.maximum(conv(a, W) + bs[l], 0.0), bf16_mode)
        acts.append(a)
    y = (conv(a, Ws[L - 1]) + bs[L - 1])[..., 0]
    return y, acts


H1_L2 = 0.01


def loss_and_dy(y, t, kind):
    """kind "mse": mean (y - t)^2; kind "h1": (sum over inner faces of the
    squared jump of e = y - t, plus H1_L2 sum e^2) / npix.  Returns the loss
    and dLoss/dy."""
    e = y - t
    npix = e.size
    if kind == "mse":
        return float(np.mean(e * e)), 2.0 * e / npix
    jx = e[:, :, 1:] - e[:, :, :-1]
    jy = e[:, 1:, :] - e[:, :-1, :]
    loss = (np.sum(jx * jx) + np.sum(jy * jy) + H1_L2 * np.sum(e * e)) / npix
    g = 2.0 * H1_L2 * e
    g[:, :, :-1] -= 2.0 * jx
    g[:, :, 1:] += 2.0 * jx
    g[:, :-1, :] -= 2.0 * jy
    g[:, 1:, :] += 2.0 * jy
    return float(loss), g / npix


def gradients(params, x, t, L, bf16_mode=True, kind="h1"):
    """Gradient of the loss (loss_and_dy) in the parameter layout, and the loss."""
    x = np.asarray(x, np.float64)
    Ws, bs = unpack(params, L)
    y, acts = forward(params, x, L, bf16_mode)
    loss, g = loss_and_dy(y, t, kind)
    dy = g[..., None]
    grads = [None] * L
    grads[L - 1] = (conv_back_weights(acts[L - 2], dy), dy.sum(axis=(0, 1, 2)))
    dz = conv_back_data(dy, Ws[L - 1]) * (acts[L - 2] > 0)
    dz = _r(dz, bf16_mode)
    for l in range(L - 2, 0, -1):
        grads[l] = (conv_back_weights(acts[l - 1], dz), dz.sum(axis=(0, 1, 2)))
        W = _r(Ws[l], bf16_mode)
        dz = _r(conv_back_data(dz, W) * (acts[l - 1] > 0), bf16_mode)
    grads[0] = (conv_back_weights(x, dz), dz.sum(axis=(0, 1, 2)))
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1].ravel()]) for g in grads])
    return flat, loss


def loss_only(params, x, t, L, kind="h1"):
    y, _ = forward(params, x, L, bf16_mode=False)
    return loss_and_dy(y, t, kind)[0]
