"""Warm-start network kernels (BF16 MFMA, FP32 accumulation) against the FP64
CPU restatement (oracle/cnn_oracle.py).

The oracle runs on the BF16-rounded weights and input the GPU uses, so the
remaining difference is the BF16 storage of intermediate activations and
gradients (relative rounding 2^-9 per value). Bars: single convolutions
relative L2 <= 1e-2; end-to-end loss <= 1e-2 relative; every parameter
gradient relative L2 <= 3e-2. Exact: Adam against its FP32 formula (1e-6),
pooling/upsampling against numpy, and bit-identical weight gradients on
repeated runs.
"""
import numpy as np
import pytest
import torch

from oracle import cnn_oracle as C
from newrecon_amd import warmstart as W

pytestmark = pytest.mark.gpu


def bf16_round(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def rel_l2(a, ref):
    a, ref = np.asarray(a, np.float64), np.asarray(ref, np.float64)
    return float(np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-30))


def to_dev_bf16(a, dev):
    """float array -> int16 device tensor holding its bf16 bits."""
    u = np.ascontiguousarray(bf16_round(a).astype(np.float32)).view(np.uint32) >> 16
    return torch.from_numpy(u.astype(np.uint16).view(np.int16)).to(dev)


def from_dev_bf16(t):
    u = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


@pytest.mark.parametrize("cin,cout", [(8, 16), (16, 32), (64, 64), (96, 32), (32, 96),
                                      (16, 48)])
def test_conv3d_matches_fp64(dev, cin, cout):
    net = W.WarmStartNet(dev)
    L = net._L
    rng = np.random.default_rng(cin * 100 + cout)
    nz, ny, nx = 4, 6, 37  # ragged x: a partial wave tile
    x = bf16_round(rng.standard_normal((nz, ny, nx, cin)))
    w = bf16_round(rng.standard_normal((cout, 27, cin)) / np.sqrt(27 * cin))
    b = rng.standard_normal(cout).astype(np.float32)
    ref = np.maximum(C.conv(x, w, b.astype(np.float64)), 0.0)
    wf = torch.zeros(cout * L.ostr_cnn_kpad(cin), dtype=torch.int16, device=dev)
    wd = torch.zeros(cin * L.ostr_cnn_kpad(cout), dtype=torch.int16, device=dev)
    wt = torch.from_numpy(w.astype(np.float32)).to(dev)
    net._check(L.ostr_cnn_pack(W._ptr(wt), cin, cout, W._ptr(wf), W._ptr(wd), None))
    xd = to_dev_bf16(x.reshape(-1), dev)
    y = torch.zeros(nz * ny * nx * cout, dtype=torch.int16, device=dev)
    bd = torch.from_numpy(b).to(dev)
    net._check(L.ostr_cnn_conv3d(cin, cout, W._ptr(xd), cin, 0, W._ptr(wf), W._ptr(bd),
                                 W._ptr(y), cout, 0, None, 0, 0, 1, nz, ny, nx, None))
    torch.cuda.synchronize()
    got = from_dev_bf16(y).reshape(nz, ny, nx, cout)
    assert rel_l2(got, ref) <= 1e-2
    # the data-gradient packing: conv with wd equals the transposed convolution
    gy = bf16_round(rng.standard_normal((nz, ny, nx, cout)))
    ref_dx, _, _ = C.conv_grads(x, w, gy)
    gyd = to_dev_bf16(gy.reshape(-1), dev)
    dx = torch.zeros(nz * ny * nx * cin, dtype=torch.int16, device=dev)
    if (cout, cin) in [(16, 8)]:
        return  # the first layer's input gradient is never needed
    net._check(L.ostr_cnn_conv3d(cout, cin, W._ptr(gyd), cout, 0, W._ptr(wd), None, W._ptr(dx),
                                 cin, 0, None, 0, 0, 0, nz, ny, nx, None))
    torch.cuda.synchronize()
    assert rel_l2(from_dev_bf16(dx).reshape(nz, ny, nx, cin), ref_dx) <= 1e-2


@pytest.mark.parametrize("cin,cout", [(16, 16), (48, 16), (16, 48), (32, 32), (96, 32)])
def test_conv3d_z_march_matches_fp64(dev, cin, cout):
    """11 x 101 x 250 voxels: each work-group marches over several slices
    (its loader waves refill the LDS ring between them), and the last tiles
    in x, y and z are partial. x and y are channel views of wider buffers
    (offsets 8 and 4); a second call adds the ReLU mask of the data gradient
    (mask a channel view too) and no bias. Bar: relative L2 <= 1e-2 against
    the FP64 convolution of the same BF16 inputs, and the same bits on a repeat."""
    net = W.WarmStartNet(dev)
    L = net._L
    rng = np.random.default_rng(cin * 7 + cout)
    nz, ny, nx = 11, 101, 250
    xcs, xco, ycs, yco = cin + 16, 8, cout + 8, 4
    x = bf16_round(rng.standard_normal((nz, ny, nx, cin)))
    w = bf16_round(rng.standard_normal((cout, 27, cin)) / np.sqrt(27 * cin))
    b = rng.standard_normal(cout).astype(np.float32)
    wf = torch.zeros(cout * L.ostr_cnn_kpad(cin), dtype=torch.int16, device=dev)
    wd = torch.zeros(cin * L.ostr_cnn_kpad(cout), dtype=torch.int16, device=dev)
    wt = torch.from_numpy(w.astype(np.float32)).to(dev)
    net._check(L.ostr_cnn_pack(W._ptr(wt), cin, cout, W._ptr(wf), W._ptr(wd), None))
    xbuf = np.zeros((nz, ny, nx, xcs))
    xbuf[..., xco:xco + cin] = x
    xd = to_dev_bf16(xbuf.reshape(-1), dev)
    y = torch.zeros(nz * ny * nx * ycs, dtype=torch.int16, device=dev)
    bd = torch.from_numpy(b).to(dev)
    ref = C.conv(x, w, b.astype(np.float64))
    m = bf16_round(rng.standard_normal((nz, ny, nx, ycs)))
    md = to_dev_bf16(m.reshape(-1), dev)
    cases = ((W._ptr(bd), None, 1, np.maximum(ref, 0.0)),
             (None, W._ptr(md), 0, np.where(m[..., yco:yco + cout] > 0, ref - b, 0.0)))
    for bias, mask, relu, want in cases:
        outs = []
        for _ in range(2):
            y.zero_()
            net._check(L.ostr_cnn_conv3d(cin, cout, W._ptr(xd), xcs, xco, W._ptr(wf), bias,
                                         W._ptr(y), ycs, yco, mask, ycs, yco, relu, nz, ny, nx,
                                         None))
            torch.cuda.synchronize()
            outs.append(y.clone())
        assert torch.equal(outs[0], outs[1])
        got = from_dev_bf16(y).reshape(nz, ny, nx, ycs)
        assert not got[..., :yco].any() and not got[..., yco + cout:].any()
        assert rel_l2(got[..., yco:yco + cout], want) <= 1e-2


@pytest.mark.parametrize("cin,cout", [(8, 16), (16, 16), (48, 16), (16, 32), (32, 32),
                                      (32, 64), (64, 64), (96, 32)])
def test_weight_gradient_z_march_matches_fp64(dev, cin, cout):
    """The weight and bias gradients of every network shape at 13 x 37 x 150
    voxels: work-groups march over 1-2 slices, the last x, y and z tiles are
    partial, x and dY are channel views (offset 8 in wider buffers); the wider
    shapes split their output channels and accumulator tiles over groups. The
    products of BF16 values are exact in FP32, so the only error is the FP32
    accumulation over 72k voxels. Bar: relative L2 <= 1e-4 (weights and bias)
    against FP64, and the same bits on a repeat."""
    net = W.WarmStartNet(dev)
    L = net._L
    rng = np.random.default_rng(40 + cin + cout)
    nz, ny, nx = 13, 37, 150
    xcs, xco, dcs, dco = cin + 8, 8, cout + 16, 8
    x = bf16_round(rng.standard_normal((nz, ny, nx, cin)))
    gy = bf16_round(rng.standard_normal((nz, ny, nx, cout)))
    w = np.zeros((cout, 27, cin))
    _, ref_dw, ref_db = C.conv_grads(x, w, gy)
    xbuf = np.zeros((nz, ny, nx, xcs))
    xbuf[..., xco:xco + cin] = x
    dbuf = np.zeros((nz, ny, nx, dcs))
    dbuf[..., dco:dco + cout] = gy
    xd = to_dev_bf16(xbuf.reshape(-1), dev)
    gd = to_dev_bf16(dbuf.reshape(-1), dev)
    nbytes = L.ostr_cnn_wgrad_workspace(cin, cout, nz, ny, nx, 64)
    work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    outs = []
    for _ in range(2):
        dw = torch.full((cout * 27 * cin,), float("nan"), device=dev)
        db = torch.full((cout,), float("nan"), device=dev)
        net._check(L.ostr_cnn_wgrad(cin, cout, W._ptr(xd), xcs, xco, W._ptr(gd), dcs, dco, nz, ny,
                                    nx, W._ptr(dw), W._ptr(db), W._ptr(work), nbytes, 64, None))
        torch.cuda.synchronize()
        outs.append((dw.clone(), db.clone()))
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    dw, db = outs[0]
    assert rel_l2(dw.cpu().numpy().reshape(cout, 27, cin), ref_dw) <= 1e-4
    assert rel_l2(db.cpu().numpy(), ref_db) <= 1e-4


def oracle_params_from(net):
    return {k: (bf16_round(w), b.astype(np.float64)) for k, (w, b) in net.host_params().items()}


def test_network_gradients_match_fp64_oracle(dev):
    net = W.WarmStartNet(dev, seed=3)
    rng = np.random.default_rng(4)
    # larger output layer and biases so that every layer's gradient is sizeable
    hp = net.host_params()
    hp = {k: (w * (10.0 if k == "out" else 1.0),
              rng.standard_normal(b.shape).astype(np.float32) * 0.05) for k, (w, b) in hp.items()}
    hp["e1a"][0][:, :, 1:] = 0.0
    net.set_params(hp)
    shape = (8, 12, 40)
    img = (0.05 + 0.01 * rng.standard_normal(shape)).astype(np.float32)
    img = bf16_round(img * W.SCALE).astype(np.float32) / np.float32(W.SCALE)
    target = (img + 0.005 * rng.standard_normal(shape)).astype(np.float32)
    imgd, tgd = torch.from_numpy(img).to(dev), torch.from_numpy(target).to(dev)
    pred = net.forward(imgd)
    loss = net.loss_backward(tgd)
    torch.cuda.synchronize()
    ref_loss, ref_g, ref_pred = C.loss_and_grads(oracle_params_from(net), img.astype(np.float64),
                                                 target.astype(np.float64), W.SCALE)
    assert abs(loss - ref_loss) <= 1e-2 * ref_loss, (loss, ref_loss)
    assert rel_l2(pred.cpu().numpy() - img, ref_pred - img) <= 2e-2
    for name, *_ in W.LAYERS:
        gw = net.weight(name, net.grads).cpu().numpy()
        gb = net.bias(name, net.grads).cpu().numpy()
        assert rel_l2(gw, ref_g[name][0]) <= 3e-2, (name, rel_l2(gw, ref_g[name][0]))
        assert rel_l2(gb, ref_g[name][1]) <= 3e-2, (name, rel_l2(gb, ref_g[name][1]))


def test_weight_gradient_is_deterministic(dev):
    net = W.WarmStartNet(dev, seed=5)
    rng = np.random.default_rng(6)
    img = torch.from_numpy((0.05 + 0.01 * rng.standard_normal((8, 8, 32))).astype(np.float32)).to(dev)
    tgt = img + 0.003
    out = []
    for _ in range(2):
        net.forward(img)
        net.loss_backward(tgt)
        torch.cuda.synchronize()
        out.append(net.grads.clone())
    assert torch.equal(out[0], out[1])


def test_adam_kernel_matches_formula(dev):
    net = W.WarmStartNet(dev, seed=7)
    rng = np.random.default_rng(8)
    p0 = net.params.cpu().numpy().copy()
    g = rng.standard_normal(net.n_params).astype(np.float32) * 1e-3
    net.grads.copy_(torch.from_numpy(g))
    net.adam_step(lr=1e-3)
    net.grads.copy_(torch.from_numpy(g * 0.5))
    net.adam_step(lr=1e-3)
    torch.cuda.synchronize()
    p, m, v = p0.astype(np.float64), np.zeros(net.n_params), np.zeros(net.n_params)
    p, m, v = C.adam(p, m, v, g.astype(np.float64), 1e-3, 0.9, 0.999, 1e-8, 1)
    p, m, v = C.adam(p, m, v, 0.5 * g.astype(np.float64), 1e-3, 0.9, 0.999, 1e-8, 2)
    got = net.params.cpu().numpy()
    assert np.abs(got - p).max() <= 1e-6 * max(1.0, np.abs(p).max())


def test_pool_and_upsample_exact(dev):
    net = W.WarmStartNet(dev)
    L = net._L
    rng = np.random.default_rng(9)
    x = bf16_round(rng.standard_normal((4, 6, 8, 16)))
    xd = to_dev_bf16(x.reshape(-1), dev)
    y = torch.zeros(2 * 3 * 4 * 16, dtype=torch.int16, device=dev)
    net._check(L.ostr_cnn_pool2(W._ptr(xd), 16, 0, W._ptr(y), 16, 0, 16, 4, 6, 8, None))
    u = torch.zeros(4 * 6 * 8 * 32, dtype=torch.int16, device=dev)
    net._check(L.ostr_cnn_up2(W._ptr(y), 16, 0, W._ptr(u), 32, 16, 16, 4, 6, 8, None))
    torch.cuda.synchronize()
    ref = C.pool(x)
    got = from_dev_bf16(y).reshape(2, 3, 4, 16)
    np.testing.assert_array_equal(got, bf16_round(ref))
    np.testing.assert_array_equal(from_dev_bf16(u).reshape(4, 6, 8, 32)[..., 16:], C.up(got))


def test_pool_up_and_gradients_vector_and_scalar_paths(dev):
    """Pooling, upsampling and their gradients (mask and accumulate) run on
    8 channels per thread when channel offsets/strides are multiples of 8 and
    one channel per thread otherwise (offset 4 here): the same bits, and the
    float32 restatement (children summed in order, 1/8, ReLU mask, the
    accumulated sum rounded to BF16)."""
    net = W.WarmStartNet(dev)
    L = net._L
    rng = np.random.default_rng(12)
    nz, ny, nx, Cc = 4, 6, 10, 16
    fine = (nz, ny, nx, Cc)
    coarse = (nz // 2, ny // 2, nx // 2, Cc)

    def dev_view(a, cs, co):
        buf = np.zeros(a.shape[:3] + (cs,))
        buf[..., co:co + Cc] = a
        return to_dev_bf16(buf.reshape(-1), dev)

    def view(t, shape, cs, co):
        return from_dev_bf16(t).reshape(shape[:3] + (cs,))[..., co:co + Cc]

    xf = bf16_round(rng.standard_normal(fine))
    xc = bf16_round(rng.standard_normal(coarse))
    mf = bf16_round(rng.standard_normal(fine))
    mc = bf16_round(rng.standard_normal(coarse))
    acc0 = bf16_round(rng.standard_normal(fine))
    f32 = np.float32
    # float32 restatements
    ch = [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    s = np.zeros(coarse, f32)
    for a, b, c in ch:
        s = s + xf[a::2, b::2, c::2].astype(f32)
    want_pool = bf16_round(s * f32(0.125))
    want_up = np.repeat(np.repeat(np.repeat(xc, 2, 0), 2, 1), 2, 2)
    g = np.repeat(np.repeat(np.repeat(xc.astype(f32) * f32(0.125), 2, 0), 2, 1), 2, 2)
    g = np.where(mf > 0, g, f32(0))
    want_pool_bwd = bf16_round(acc0.astype(f32) + g)
    s = np.zeros(coarse, f32)
    for a, b, c in ch:
        s = s + xf[a::2, b::2, c::2].astype(f32)
    want_up_bwd = bf16_round(np.where(mc > 0, s, f32(0)))
    results = {}
    for cs, co in ((16, 0), (24, 4)):  # vector path, scalar path
        out = {}
        y = torch.zeros(int(np.prod(coarse[:3])) * cs, dtype=torch.int16, device=dev)
        net._check(L.ostr_cnn_pool2(W._ptr(dev_view(xf, cs, co)), cs, co, W._ptr(y), cs, co, Cc,
                                    nz, ny, nx, None))
        out["pool"] = view(y, coarse, cs, co)
        u = torch.zeros(int(np.prod(fine[:3])) * cs, dtype=torch.int16, device=dev)
        net._check(L.ostr_cnn_up2(W._ptr(dev_view(xc, cs, co)), cs, co, W._ptr(u), cs, co, Cc,
                                  nz, ny, nx, None))
        out["up"] = view(u, fine, cs, co)
        dx = dev_view(acc0, cs, co)
        mfd, dyc = dev_view(mf, cs, co), dev_view(xc, cs, co)
        net._check(L.ostr_cnn_pool2_bwd(W._ptr(dyc), cs, co, W._ptr(dx), cs, co, W._ptr(mfd), cs,
                                        co, 1, Cc, nz, ny, nx, None))
        out["pool_bwd"] = view(dx, fine, cs, co)
        dxc = torch.zeros(int(np.prod(coarse[:3])) * cs, dtype=torch.int16, device=dev)
        mcd, dyf = dev_view(mc, cs, co), dev_view(xf, cs, co)
        net._check(L.ostr_cnn_up2_bwd(W._ptr(dyf), cs, co, W._ptr(dxc), cs, co, W._ptr(mcd), cs,
                                      co, Cc, nz, ny, nx, None))
        out["up_bwd"] = view(dxc, coarse, cs, co)
        torch.cuda.synchronize()
        results[co] = out
    for k, want in (("pool", want_pool), ("up", want_up), ("pool_bwd", want_pool_bwd),
                    ("up_bwd", want_up_bwd)):
        np.testing.assert_array_equal(results[0][k], results[4][k], err_msg=k)
        np.testing.assert_array_equal(results[0][k], want, err_msg=k)


def test_training_reduces_loss(dev):
    """Adam steps on one sample with a learnable target (an affine map of the
    input) lower the loss several-fold (smoke test of the training loop)."""
    net = W.WarmStartNet(dev, seed=11)
    rng = np.random.default_rng(12)
    img = torch.from_numpy((0.05 + 0.01 * rng.standard_normal((8, 16, 32))).astype(np.float32)).to(dev)
    tgt = 1.3 * img - 0.01
    losses = []
    for _ in range(60):
        net.forward(img)
        losses.append(net.loss_backward(tgt))
        net.adam_step(lr=2e-3)
    assert losses[-1] < 0.2 * losses[0], losses
