"""FP64 CPU restatement of the warm-start network (test infrastructure).

The checker of the GPU network kernels (newrecon_amd/csrc/cnn.hip): forward
pass and analytic gradients in float64 numpy, same architecture, layouts and
conventions (channels-last activations, weights [cout][27][cin] with tap =
(dz+1)*9 + (dy+1)*3 + (dx+1), zero padding, ReLU, 2x average pooling, 2x
nearest upsampling, channel concatenation [upsampled, skip], residual output
pred = img + out[..., 0] / scale). Its gradients are pinned by finite
differences (tests/test_cnn_oracle.py); nothing in the product imports it.
"""
from __future__ import annotations

import numpy as np

# name, cin, cout, relu (see newrecon_amd.warmstart.LAYERS)
LAYERS = [
    ("e1a", 8, 16, True), ("e1b", 16, 16, True),
    ("e2a", 16, 32, True), ("e2b", 32, 32, True),
    ("e3a", 32, 64, True), ("e3b", 64, 64, True),
    ("d2", 96, 32, True), ("d1", 48, 16, True),
    ("out", 16, 16, False),
]

OFFS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def shifted(x, dz, dy, dx):
    """s[z,y,x] = x[z+dz, y+dy, x+dx] (zero outside)."""
    nz, ny, nx = x.shape[:3]
    out = np.zeros_like(x)
    zs, zd = slice(max(dz, 0), nz + min(dz, 0)), slice(max(-dz, 0), nz + min(-dz, 0))
    ys, yd = slice(max(dy, 0), ny + min(dy, 0)), slice(max(-dy, 0), ny + min(-dy, 0))
    xs, xd = slice(max(dx, 0), nx + min(dx, 0)), slice(max(-dx, 0), nx + min(-dx, 0))
    out[zd, yd, xd] = x[zs, ys, xs]
    return out


def conv(x, w, b=None):
    """x [nz,ny,nx,cin], w [cout,27,cin] -> [nz,ny,nx,cout]."""
    y = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, (dz, dy, dx) in enumerate(OFFS):
        y += shifted(x, dz, dy, dx) @ w[:, t, :].T
    return y if b is None else y + b


def conv_grads(x, w, dy):
    """(dx, dw, db) of y = conv(x, w) + b for output gradient dy."""
    dw = np.zeros_like(w)
    dx = np.zeros_like(x)
    flat_dy = dy.reshape(-1, dy.shape[-1])
    for t, (oz, oy, ox) in enumerate(OFFS):
        xs = shifted(x, oz, oy, ox).reshape(-1, x.shape[-1])
        dw[:, t, :] = flat_dy.T @ xs
        dx += shifted(dy @ w[:, t, :], -oz, -oy, -ox)
    return dx, dw, flat_dy.sum(axis=0)


def pool(x):
    nz, ny, nx, c = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def pool_grad(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_grad(dy):
    nz, ny, nx, c = dy.shape
    return dy.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))


def forward(params, img, scale):
    """params: {name: (w [cout,27,cin], b [cout])}; img [nz,ny,nx].
    Returns (pred, cache)."""
    relu = lambda a: np.maximum(a, 0.0)
    x0 = np.zeros(img.shape + (8,))
    x0[..., 0] = img * scale
    a = {}
    a["e1a"] = relu(conv(x0, *params["e1a"]))
    a["e1b"] = relu(conv(a["e1a"], *params["e1b"]))
    p1 = pool(a["e1b"])
    a["e2a"] = relu(conv(p1, *params["e2a"]))
    a["e2b"] = relu(conv(a["e2a"], *params["e2b"]))
    p2 = pool(a["e2b"])
    a["e3a"] = relu(conv(p2, *params["e3a"]))
    a["e3b"] = relu(conv(a["e3a"], *params["e3b"]))
    cat2 = np.concatenate([up(a["e3b"]), a["e2b"]], axis=-1)
    a["d2"] = relu(conv(cat2, *params["d2"]))
    cat1 = np.concatenate([up(a["d2"]), a["e1b"]], axis=-1)
    a["d1"] = relu(conv(cat1, *params["d1"]))
    out = conv(a["d1"], *params["out"])
    pred = img + out[..., 0] / scale
    cache = dict(a, x0=x0, p1=p1, p2=p2, cat1=cat1, cat2=cat2, out=out)
    return pred, cache


def loss_and_grads(params, img, target, scale):
    """mean((pred - target)^2) and its gradient for every parameter."""
    pred, c = forward(params, img, scale)
    n = img.size
    err = pred - target
    loss = float((err * err).sum() / n)
    g = {}
    dout = np.zeros(c["out"].shape)
    dout[..., 0] = 2.0 * err / (n * scale)
    d, *g["out"] = conv_grads(c["d1"], params["out"][0], dout)
    d = d * (c["d1"] > 0)
    dcat1, *g["d1"] = conv_grads(c["cat1"], params["d1"][0], d)
    dcat1 = dcat1 * (c["cat1"] > 0)
    dd2 = up_grad(dcat1[..., :32]) * (c["d2"] > 0)
    de1b = dcat1[..., 32:]
    dcat2, *g["d2"] = conv_grads(c["cat2"], params["d2"][0], dd2)
    dcat2 = dcat2 * (c["cat2"] > 0)
    de3b = up_grad(dcat2[..., :64]) * (c["e3b"] > 0)
    de2b = dcat2[..., 64:]
    de3a, *g["e3b"] = conv_grads(c["e3a"], params["e3b"][0], de3b)
    de3a = de3a * (c["e3a"] > 0)
    dp2, *g["e3a"] = conv_grads(c["p2"], params["e3a"][0], de3a)
    de2b = de2b + pool_grad(dp2) * (c["e2b"] > 0)
    de2a, *g["e2b"] = conv_grads(c["e2a"], params["e2b"][0], de2b)
    de2a = de2a * (c["e2a"] > 0)
    dp1, *g["e2a"] = conv_grads(c["p1"], params["e2a"][0], de2a)
    de1b = de1b + pool_grad(dp1) * (c["e1b"] > 0)
    de1a, *g["e1b"] = conv_grads(c["e1a"], params["e1b"][0], de1b)
    de1a = de1a * (c["e1a"] > 0)
    _, *g["e1a"] = conv_grads(c["x0"], params["e1a"][0], de1a)
    return loss, {k: (v[0], v[1]) for k, v in g.items()}, pred


def adam(p, m, v, g, lr, b1, b2, eps, step):
    """Bias-corrected Adam step (Kingma & Ba 2015), float64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    """He-normal weights, zero biases (float64)."""
    rng = np.random.default_rng(seed)
    out = {}
    for name, cin, cout, _ in LAYERS:
        std = np.sqrt(2.0 / (27 * cin))
        if name == "out":
            std *= 0.1
        out[name] = (rng.standard_normal((cout, 27, cin)) * std, np.zeros(cout))
    return out
