"""CPU oracle of the learned warm-start network (numpy).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the
``cdft_unet_*`` entry points; the product package never imports it.

The network has no counterpart in the reference; it is the 3-level 3D U-Net
that maps (beta V_ext, mu) on periodic D^3 grids to log rho:

  x0  = [beta V/3, (mu+3)/2, 0 x 6]                         (8 channels, level 0)
  y0  = relu(conv0(x0)), y1 = relu(conv1(y0))               (16, level 0)
  y2  = relu(conv2(pool(y1))), y3 = relu(conv3(y2))         (32, level 1)
  y4  = relu(conv4(pool(y3))), y5 = relu(conv5(y4))         (64, level 2)
  y6  = relu(conv6([up(y5), y3])), y7 = relu(conv7(y6))     (32, level 1)
  y8  = relu(conv8([up(y7), y1])), y9 = relu(conv9(y8))     (16, level 0)
  out = hb + hw . y9                                        (FP32)

conv is a 3x3x3 periodic correlation Y[p][co] = b[co] + sum_{tap,ci}
W[co][tap][ci] X[p + off(tap)][ci], off(tap) = (tap//9 - 1, tap//3 % 3 - 1,
tap % 3 - 1); pool is a 2x2x2 average, up a nearest-neighbour copy.  The loss
is mean((out - target)^2).  Parameters are one flat FP32 vector: per layer
W [cout][27][cin] then b [cout], then hw [16] and hb.

``bf16=True`` rounds (round to nearest even) where the device kernels store
BF16: the network input, the weights, every convolution output, pooling
outputs and every stored gradient; sums are in float64 here and FP32 on the
device.  ``bf16=False`` is the plain float64 network, whose backward pass is
pinned by finite differences of its forward pass and whose convolution is
pinned by an FFT circular convolution (tests/test_unet_oracle.py).
"""
import numpy as np

LAYERS = [(8, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2),
          (64, 64, 2), (96, 32, 1), (32, 32, 1), (48, 16, 0), (16, 16, 0)]

TAPS = [(t // 9 - 1, t // 3 % 3 - 1, t % 3 - 1) for t in range(27)]


def bf16_round(x):
    """float -> nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def offsets():
    """Offsets of each layer's W and b and of the head in the flat parameter vector, and its length."""
    o, wo, bo = 0, [], []
    for cin, cout, _ in LAYERS:
        wo.append(o)
        o += cout * 27 * cin
        bo.append(o)
        o += cout
    return wo, bo, o, o + 17


def unpack(params):
    wo, bo, ho, _ = offsets()
    W = [params[wo[i]:wo[i] + c * 27 * k].reshape(c, 27, k) for i, (k, c, _) in enumerate(LAYERS)]
    b = [params[bo[i]:bo[i] + c] for i, (_, c, _) in enumerate(LAYERS)]
    return W, b, params[ho:ho + 16], params[ho + 16]


def conv(X, W, b=None):
    """Periodic 3x3x3 correlation, X (B, D, D, D, cin), W (cout, 27, cin)."""
    Y = np.zeros(X.shape[:4] + (W.shape[0],))
    for t, (dx, dy, dz) in enumerate(TAPS):
        Y += np.roll(X, (-dx, -dy, -dz), axis=(1, 2, 3)) @ W[:, t, :].T
    return Y if b is None else Y + b


def conv_bwd_data(dZ, W):
    dX = np.zeros(dZ.shape[:4] + (W.shape[2],))
    for t, (dx, dy, dz) in enumerate(TAPS):
        dX += np.roll(dZ, (dx, dy, dz), axis=(1, 2, 3)) @ W[:, t, :]
    return dX


def conv_wgrad(X, dZ):
    co, ci = dZ.shape[-1], X.shape[-1]
    dW = np.zeros((co, 27, ci))
    Z2 = dZ.reshape(-1, co)
    for t, (dx, dy, dz) in enumerate(TAPS):
        dW[:, t, :] = Z2.T @ np.roll(X, (-dx, -dy, -dz), axis=(1, 2, 3)).reshape(-1, ci)
    return dW, Z2.sum(axis=0)


def pool(X):
    B, D = X.shape[0], X.shape[1]
    return X.reshape(B, D // 2, 2, D // 2, 2, D // 2, 2, -1).mean(axis=(2, 4, 6))


def up(X):
    return X.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def up_bwd(dY):
    B, D = dY.shape[0], dY.shape[1]
    return dY.reshape(B, D // 2, 2, D // 2, 2, D // 2, 2, -1).sum(axis=(2, 4, 6))


def net_input(V, mu):
    """V (B, D, D, D) beta V_ext, mu (B,) -> (B, D, D, D, 8) as the device forms it (FP32 arithmetic)."""
    V = np.asarray(V, dtype=np.float32)
    x = np.zeros(V.shape + (8,))
    x[..., 0] = V * np.float32(1.0 / 3.0)
    x[..., 1] = ((np.asarray(mu, dtype=np.float32) + np.float32(3.0)) * np.float32(0.5))[:, None, None, None]
    return x


def forward(params, V, mu, bf16=True):
    """Returns (out (B, D, D, D), cache of intermediates for backward)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W, b, hw, hb = unpack(np.asarray(params, dtype=np.float64))
    Wr = [r(w) for w in W]
    c = {"x0": r(net_input(V, mu)), "Wr": Wr}
    y = [None] * 10

    def layer(i, X):
        y[i] = r(np.maximum(conv(X, Wr[i], b[i]), 0.0))
        return y[i]

    layer(1, layer(0, c["x0"]))
    c["p1"] = r(pool(y[1]))
    layer(3, layer(2, c["p1"]))
    c["p2"] = r(pool(y[3]))
    layer(5, layer(4, c["p2"]))
    c["u2"] = np.concatenate([up(y[5]), y[3]], axis=-1)
    layer(7, layer(6, c["u2"]))
    c["u1"] = np.concatenate([up(y[7]), y[1]], axis=-1)
    layer(9, layer(8, c["u1"]))
    c["y"] = y
    out = y[9] @ hw + hb
    if bf16:
        out = out.astype(np.float32).astype(np.float64)
    return out, c


def loss_and_grad(params, V, mu, T, bf16=True):
    """Mean squared error of out against T and its gradient with respect to the flat parameters."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    params = np.asarray(params, dtype=np.float64)
    out, c = forward(params, V, mu, bf16)
    W, b, hw, hb = unpack(params)
    Wr, y = c["Wr"], c["y"]
    e = out - T
    n = e.size
    d = 2.0 * e / n
    g = np.zeros_like(params)
    wo, bo, ho, _ = offsets()
    g[ho:ho + 16] = np.tensordot(d, y[9], axes=(range(4), range(4)))
    g[ho + 16] = d.sum()

    def back(i, X, dZ, want_dx=True):
        dW, db = conv_wgrad(X, dZ)
        cin, cout, _ = LAYERS[i]
        g[wo[i]:wo[i] + cout * 27 * cin] = dW.ravel()
        g[bo[i]:bo[i] + cout] = db
        return r(conv_bwd_data(dZ, Wr[i])) if want_dx else None

    mask = [yi > 0 for yi in y]
    dz = r(d[..., None] * hw) * mask[9]
    dz = back(9, y[8], dz) * mask[8]
    du1 = back(8, c["u1"], dz)
    dz = r(up_bwd(du1[..., :32])) * mask[7]
    dskip1 = du1[..., 32:]
    dz = back(7, y[6], dz) * mask[6]
    du2 = back(6, c["u2"], dz)
    dz = r(up_bwd(du2[..., :64])) * mask[5]
    dskip3 = du2[..., 64:]
    dz = back(5, y[4], dz) * mask[4]
    dp2 = back(4, c["p2"], dz)
    dz = r(up(r(0.125 * dp2)) + dskip3) * mask[3]
    dz = back(3, y[2], dz) * mask[2]
    dp1 = back(2, c["p1"], dz)
    dz = r(up(r(0.125 * dp1)) + dskip1) * mask[1]
    dz = back(1, y[0], dz) * mask[0]
    back(0, c["x0"], dz, want_dx=False)
    return float(np.mean(e * e)), g


def init_params(seed, std_head=0.25):
    """He-normal weights (layer 0 sees 2 input channels), zero biases: same scheme as the device
    initialiser, not the same random stream."""
    rng = np.random.default_rng(seed)
    wo, bo, ho, n = offsets()
    p = np.zeros(n)
    for i, (cin, cout, _) in enumerate(LAYERS):
        sd = np.sqrt(2.0 / (27.0 * (2 if i == 0 else cin)))
        p[wo[i]:wo[i] + cout * 27 * cin] = sd * rng.standard_normal(cout * 27 * cin)
    p[ho:ho + 16] = std_head * rng.standard_normal(16)
    return p.astype(np.float32)
