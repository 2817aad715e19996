"""Pins the FP64 network restatement (oracle/cnn_oracle.py): analytic
gradients of every layer against central finite differences of the loss, the
convolution against a direct per-voxel sum, and Adam against its closed form
for a constant gradient (bias-corrected first step = lr * sign(g))."""
import numpy as np
import pytest

from oracle import cnn_oracle as C


def case(seed=0, shape=(4, 8, 8)):
    rng = np.random.default_rng(seed)
    params = C.init_params(seed)
    # a larger output layer so every layer's gradient is well above round-off
    w, b = params["out"]
    params["out"] = (w * 10.0, rng.standard_normal(b.shape) * 0.01)
    params = {k: (w, rng.standard_normal(b.shape) * 0.05 if k != "out" else b)
              for k, (w, b) in params.items()}
    img = 0.05 + 0.01 * rng.standard_normal(shape)
    target = img + 0.005 * rng.standard_normal(shape)
    return params, img, target


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 4, 5, 8))
    w = rng.standard_normal((16, 27, 8))
    y = C.conv(x, w)
    for (z, yy, xx, co) in [(0, 0, 0, 0), (1, 2, 3, 5), (2, 3, 4, 15), (1, 0, 4, 7)]:
        s = 0.0
        for t, (dz, dy, dx) in enumerate(C.OFFS):
            zz, y2, x2 = z + dz, yy + dy, xx + dx
            if 0 <= zz < 3 and 0 <= y2 < 4 and 0 <= x2 < 5:
                s += w[co, t] @ x[zz, y2, x2]
        assert abs(y[z, yy, xx, co] - s) < 1e-12


@pytest.mark.parametrize("layer", [n for n, *_ in C.LAYERS])
def test_gradients_match_finite_differences(layer):
    params, img, target = case()
    scale = 20.0
    _, grads, _ = C.loss_and_grads(params, img, target, scale)
    rng = np.random.default_rng(hash(layer) % 1000)
    w, b = params[layer]
    gw, gb = grads[layer]
    checks = []
    for _ in range(4):  # weight entries with a non-negligible gradient
        idx = tuple(rng.integers(0, s) for s in w.shape)
        if layer == "e1a":
            idx = (idx[0], idx[1], 0)  # input channels 1..7 are zero
        checks.append(("w", idx))
    checks.append(("b", (int(rng.integers(0, b.shape[0])),)))
    for kind, idx in checks:
        arr = w if kind == "w" else b
        h = 1e-6
        old = arr[idx]
        arr[idx] = old + h
        lp, _, _ = C.loss_and_grads(params, img, target, scale)
        arr[idx] = old - h
        lm, _, _ = C.loss_and_grads(params, img, target, scale)
        arr[idx] = old
        fd = (lp - lm) / (2 * h)
        an = (gw if kind == "w" else gb)[idx]
        assert abs(fd - an) <= 1e-5 * max(abs(an), abs(fd)) + 1e-12, (layer, kind, idx, fd, an)


def test_adam_first_step_and_closed_form():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.1, 2.0])
    p1, m, v = C.adam(p, np.zeros(3), np.zeros(3), g, 1e-3, 0.9, 0.999, 0.0, 1)
    np.testing.assert_allclose(p1, p - 1e-3 * np.sign(g), rtol=0, atol=1e-15)
    # constant gradient: m_t = (1-b1^t) g, v_t = (1-b2^t) g^2 -> step lr*sign(g)
    q, m, v = p.copy(), np.zeros(3), np.zeros(3)
    for t in range(1, 6):
        q, m, v = C.adam(q, m, v, g, 1e-3, 0.9, 0.999, 0.0, t)
    np.testing.assert_allclose(q, p - 5e-3 * np.sign(g), rtol=0, atol=1e-13)
