"""Pins the CPU oracle of the warm-start network before it checks the device kernels.

The convolution is checked against an FFT circular convolution (a different
algorithm), the BF16 rounding against torch's float32 -> bfloat16 cast and
hand-worked ties, pooling/upsampling by adjointness, and the whole backward
pass against central finite differences of the float64 forward pass, one
direction per layer so that a wrong index in any layer's gradient fails.
"""
import numpy as np
import pytest
import torch

from oracle import unet

D = 32


def smooth_field(rng, B, D, scale):
    k = np.fft.fftfreq(D)[:, None, None] ** 2 + np.fft.fftfreq(D)[None, :, None] ** 2 + \
        np.fft.fftfreq(D)[None, None, :] ** 2
    out = []
    for _ in range(B):
        f = np.fft.ifftn(np.fft.fftn(rng.standard_normal((D, D, D))) * np.exp(-k * 60.0)).real
        out.append(scale * f / np.max(np.abs(f)))
    return np.array(out)


def test_conv_matches_fft_convolution():
    rng = np.random.default_rng(1)
    n, cin, cout = 8, 3, 2
    X = rng.standard_normal((1, n, n, n, cin))
    W = rng.standard_normal((cout, 27, cin))
    Y = unet.conv(X, W)
    for co in range(cout):
        acc = np.zeros((n, n, n))
        for ci in range(cin):
            h = np.zeros((n, n, n))  # Y = h * X (convolution) with h[-off] = W[tap]
            for t, (dx, dy, dz) in enumerate(unet.TAPS):
                h[-dx % n, -dy % n, -dz % n] = W[co, t, ci]
            acc += np.fft.ifftn(np.fft.fftn(h) * np.fft.fftn(X[0, ..., ci])).real
        assert np.max(np.abs(acc - Y[0, ..., co])) < 1e-12 * np.max(np.abs(acc))


def test_bf16_rounding():
    one = 1.0
    assert unet.bf16_round(np.array([one]))[0] == 1.0
    assert unet.bf16_round(np.array([one + 2.0 ** -8]))[0] == 1.0  # tie -> even mantissa
    assert unet.bf16_round(np.array([one + 3 * 2.0 ** -8]))[0] == one + 2.0 ** -6  # tie -> even mantissa
    x = np.random.default_rng(2).standard_normal(100000).astype(np.float32) * 10.0 ** np.random.default_rng(3).integers(
        -30, 30, 100000)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(unet.bf16_round(x), ref)


def test_pool_up_adjoint():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 4, 4, 4, 3))
    y = rng.standard_normal((2, 8, 8, 8, 3))
    assert abs(np.sum(unet.up(x) * y) - np.sum(x * unet.up_bwd(y))) < 1e-12 * np.sum(np.abs(x * unet.up_bwd(y)))
    big = rng.standard_normal((1, 8, 8, 8, 2))
    assert np.allclose(unet.pool(big)[0, 1, 2, 3], big[0, 2:4, 4:6, 6:8].mean(axis=(0, 1, 2)))
    Wx = rng.standard_normal((3, 27, 2))
    dz = rng.standard_normal((1, 8, 8, 8, 3))
    # conv^T is the adjoint of conv
    assert abs(np.sum(unet.conv(big, Wx) * dz) - np.sum(big * unet.conv_bwd_data(dz, Wx))) < 1e-10


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    V = smooth_field(rng, 1, D, 3.0)
    mu = np.array([0.7])
    T = smooth_field(rng, 1, D, 1.0) - 1.0
    p = unet.init_params(6).astype(np.float64)
    wo, bo, ho, n = unet.offsets()
    p[ho + 16] = -0.5
    for i, (_, c, _) in enumerate(unet.LAYERS):
        p[bo[i]:bo[i] + c] = 0.05 * rng.standard_normal(c)
    loss, g = unet.loss_and_grad(p, V, mu, T, bf16=False)
    blocks = [(wo[i], bo[i] + unet.LAYERS[i][1]) for i in range(10)] + [(ho, n)]
    for lo, hi in blocks:
        v = np.zeros(n)
        v[lo:hi] = rng.standard_normal(hi - lo)
        h = 1e-6 / np.linalg.norm(v)  # small enough that no ReLU kink is crossed
        lp, _ = unet.forward(p + h * v, V, mu, bf16=False)
        lm, _ = unet.forward(p - h * v, V, mu, bf16=False)
        fd = (np.mean((lp - T) ** 2) - np.mean((lm - T) ** 2)) / (2 * h)
        an = g @ v
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-12, (lo, hi, fd, an)


def test_bf16_network_close_to_float64():
    """The rounding points perturb the output and gradient at the BF16 level, not more."""
    rng = np.random.default_rng(7)
    V = smooth_field(rng, 1, D, 3.0)
    mu = np.array([-0.4])
    T = smooth_field(rng, 1, D, 1.0) - 1.0
    p = unet.init_params(8)
    o64, _ = unet.forward(p, V, mu, bf16=False)
    o16, _ = unet.forward(p, V, mu, bf16=True)
    assert np.linalg.norm(o16 - o64) <= 3e-2 * np.linalg.norm(o64)
    _, g64 = unet.loss_and_grad(p, V, mu, T, bf16=False)
    _, g16 = unet.loss_and_grad(p, V, mu, T, bf16=True)
    assert np.linalg.norm(g16 - g64) <= 5e-2 * np.linalg.norm(g64)
