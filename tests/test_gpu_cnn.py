"""The GPU network (k_cnn.hip: MFMA BF16 convolutions with FP32 accumulation,
own backward and Adam kernels) against the CPU oracle, and the learned
initial guess inside the pressure solve."""
import numpy as np
import pytest

from helpers import pair, params
from impec2d_amd import fields
from oracle import cnn_oracle as co

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("L,n,h,w,kind", [(8, 2, 32, 64, "h1"), (4, 1, 40, 33, "mse"), (3, 3, 8, 32, "h1"),
                                          (8, 1, 32, 32, "mse")])
def test_forward_and_gradients_match_oracle(gpu_lib, L, n, h, w, kind):
    from impec2d_amd import capi
    _say("cnn", L, n, h, w, kind)
    net = capi.Network(layers=L, capacity=4, crop=32, seed=5, mse_loss=(kind == "mse"))
    rng = np.random.default_rng(9)
    p = (rng.standard_normal(net.nparam) * 0.2).astype(np.float32)
    p = co.bf16(p)   # weights representable in BF16: the layers only differ by rounding of activations
    net.set_params(p)
    x = rng.standard_normal((n, h, w, 3)).astype(np.float32)
    t = rng.standard_normal((n, h, w)).astype(np.float32)
    y = net.forward(x)
    yo, _ = co.forward(p.astype(np.float64), x, L, bf16_mode=True)
    assert _rel(y, yo) < 1e-2, _rel(y, yo)
    g, loss = net.gradients(x, t)
    go, losso = co.gradients(p.astype(np.float64), x, t, L, bf16_mode=True, kind=kind)
    assert abs(loss - losso) <= 1e-2 * losso
    _, offs, _ = co.layout(L)
    for l, (wo, bo) in enumerate(offs):
        shapes = co.layout(L)[0][l]
        end = bo + shapes[1]
        e = _rel(g[wo:end], go[wo:end])
        assert e < 3e-2, (l, e)


def test_learned_guess_trains_and_keeps_tolerance(gpu_lib):
    """Pairs from a lognormal-K run (M=20), a few hundred Adam steps, then three
    warm starts on later steps: the previous pressure, the previous pressure
    plus the last step's change (extrapolation, no network), and the learned
    guess (previous pressure plus the network's increment and the last step's
    change, with minimal-residual step lengths).  Every solve reaches the
    tolerance; V-cycles are reported; the learned guess, minimising the same
    residual over a larger space of corrections, needs no more V-cycles than
    the extrapolation to within one per step on average."""
    from impec2d_amd import capi
    n = 128
    p = params(N1=n, N2=n, eps=1e-10, mu_inj=1.0 / 20)
    K = fields.lognormal_permeability(n, n, sigma=1.0, corr_len=0.05, seed=11)
    net = capi.Network(layers=8, capacity=2048, crop=32, seed=3)
    o, g = pair(p, K, None, initial_guess=capi.GUESS_PREVIOUS)
    g.set_warmstart(net)   # keeps the previous pressure for the pairs
    g.set_source(o.source("one_cell"))
    g.set_concentration(o.scalar())
    t = 0.0
    for k in range(40):
        t += g.step(t)
        if k >= 1:
            net.collect(g, 16)
    assert net.pairs() == 39 * 16
    l0 = net.train(1, batch=16, lr=1e-3)
    l1 = net.train(400, batch=16, lr=1e-3)
    _say("training loss %.4f -> %.4f" % (l0, l1))
    assert l1 < l0
    state = g.get_concentration()
    pstate = g.get_pressure()
    res = {}
    steps = 10
    for guess in (capi.GUESS_PREVIOUS, capi.GUESS_EXTRAPOLATE, capi.GUESS_LEARNED):
        o2, g2 = pair(p, K, None, initial_guess=guess)
        if guess == capi.GUESS_LEARNED:
            g2.set_warmstart(net)
        g2.set_source(o2.source("one_cell"))
        g2.set_concentration(state)
        g2.set_pressure(pstate)
        tt, its, steplen = t, [], []
        for _ in range(steps):
            tt += g2.step(tt)
            st = g2.stats()
            assert st.last_relres <= 1e-10
            its.append(st.last_its)
            steplen.append((round(st.last_guess_alpha, 3), round(st.last_guess_gamma, 3)))
        res[guess] = its
        if guess != capi.GUESS_PREVIOUS:
            _say("step lengths (network increment, last change)", steplen)
        g2.close()
    _say("V-cycles previous-p", res[capi.GUESS_PREVIOUS], "extrapolated", res[capi.GUESS_EXTRAPOLATE],
         "learned", res[capi.GUESS_LEARNED])
    assert sum(res[capi.GUESS_EXTRAPOLATE]) <= sum(res[capi.GUESS_PREVIOUS]) + 2
    assert sum(res[capi.GUESS_LEARNED]) <= sum(res[capi.GUESS_EXTRAPOLATE]) + steps
