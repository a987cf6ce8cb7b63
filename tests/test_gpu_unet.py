"""Device warm-start network (BF16 MFMA convolutions, FP32 accumulation) against
the CPU oracle with BF16 rounding at the same points.

Forward output and every layer's weight/bias gradient are compared in relative
L2 norm: the device sums in FP32 in another order, so a few BF16 roundings land
on the other side; a wrong tap, channel or transposition gives O(1) errors.
"""
import numpy as np
import pytest

from oracle import unet
from test_unet_oracle import smooth_field

pytestmark = pytest.mark.gpu

D, B = 32, 2


def dataset(seed, n):
    rng = np.random.default_rng(seed)
    V = smooth_field(rng, n, D, 3.0).astype(np.float32)
    mu = rng.uniform(-1.5, 1.5, n).astype(np.float32)
    T = (smooth_field(rng, n, D, 1.0) - 1.0).astype(np.float32)
    return V, mu, T


def device(cdft, V, mu, T):
    return (cdft.DeviceArray.from_host(V, np.float32), cdft.DeviceArray.from_host(mu, np.float32),
            cdft.DeviceArray.from_host(T, np.float32))


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


def trained_like_params(seed):
    """He weights plus small biases and a head bias, so every ReLU sees both signs."""
    p = unet.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 1)
    _, bo, ho, _ = unet.offsets()
    for i, (_, c, _) in enumerate(unet.LAYERS):
        p[bo[i]:bo[i] + c] = 0.05 * rng.standard_normal(c)
    p[ho + 16] = -0.5
    return p.astype(np.float32)


def test_forward_matches_oracle(cdft):
    V, mu, T = dataset(11, 3)
    dV, dmu, _ = device(cdft, V, mu, T)
    net = cdft.UNet(D)
    p = trained_like_params(12)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    idx = [2, 0]
    out = cdft.DeviceArray((B, D, D, D), np.float32)
    net.forward(dV, dmu, idx, out)
    ref, _ = unet.forward(p, V[idx], mu[idx], bf16=True)
    got = out.download()
    assert rel(got, ref) <= 1e-2, rel(got, ref)
    # a batch of one after a batch of two (buffers re-sized) gives the same member
    one = cdft.DeviceArray((1, D, D, D), np.float32)
    net.forward(dV, dmu, [0], one)
    assert rel(one.download()[0], got[1]) <= 1e-6


def test_gradients_match_oracle(cdft):
    V, mu, T = dataset(13, 2)
    dV, dmu, dT = device(cdft, V, mu, T)
    net = cdft.UNet(D)
    p = trained_like_params(14)
    net.set_params(p)
    loss, g = net.grad(dV, dmu, dT, [0, 1])
    lref, gref = unet.loss_and_grad(p, V, mu, T, bf16=True)
    assert abs(loss - lref) <= 1e-3 * lref
    wo, bo, ho, n = unet.offsets()
    for i, (cin, cout, _) in enumerate(unet.LAYERS):
        w = slice(wo[i], wo[i] + cout * 27 * cin)
        b = slice(bo[i], bo[i] + cout)
        assert rel(g[w], gref[w]) <= 3e-2, (i, "W", rel(g[w], gref[w]))
        assert rel(g[b], gref[b]) <= 3e-2, (i, "b", rel(g[b], gref[b]))
    h = slice(ho, n)
    assert rel(g[h], gref[h]) <= 1e-3, rel(g[h], gref[h])


def test_train_step_is_adam(cdft):
    """First Adam step with bias correction: w -= lr g / (|g| + 1e-8); the loss then falls on a fixed batch."""
    V, mu, T = dataset(15, 2)
    dV, dmu, dT = device(cdft, V, mu, T)
    net = cdft.UNet(D, seed=3)
    p0 = net.params()
    _, g = net.grad(dV, dmu, dT, [0, 1])
    lr = 1e-3
    l0 = net.train_step(dV, dmu, dT, [0, 1], lr)
    p1 = net.params()
    expect = p0 - lr * g / (np.abs(g) + 1e-8)
    assert np.max(np.abs(p1 - expect)) <= 1e-6
    losses = [l0] + [net.train_step(dV, dmu, dT, [0, 1], lr) for _ in range(40)]
    assert losses[-1] < 0.5 * losses[0], losses
