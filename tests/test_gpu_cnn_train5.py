"""The warm-start network's training step at the stated training size:
256x256x48 voxels of 0.1x0.1x1 mm, 15 views over +-7.5 deg, I0 = 5e3 (one
seeded case, seed 100000, paired with its 50-iteration OSTR image from this
code), forward, backward and Adam through include/ostr_cnn.h.

Checked: the first layer's output on an interior block and the output layer's
weight gradient over the whole volume against the FP64 oracle
(oracle/cnn_oracle.py) on the same BF16 inputs; the weight gradients bit for
bit on a repeat; the loss falling over Adam steps."""
import numpy as np
import pytest
import torch

from oracle import cnn_oracle as C
from newrecon_amd import warmstart as W

pytestmark = pytest.mark.gpu


def rel_l2(a, ref):
    a, ref = np.asarray(a, np.float64), np.asarray(ref, np.float64)
    return float(np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-30))


def from_dev_bf16(t):
    u = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def train_case(dev):
    from newrecon_amd.geometry import OSTRConfig, scan_train5
    from newrecon_amd.ostr import OSTR
    scan = scan_train5()
    _, y, b = W.make_case(scan, scan.phantom_seed, dev)
    rec = OSTR(scan.geom, scan.angles, y, b, cfg=OSTRConfig(n_subsets=5), x0="bp", device=dev)
    bp = rec.x.clone()
    target = rec.iterate(50).clone()
    torch.cuda.synchronize()
    print(f"[train5] {scan.name}: bp {tuple(bp.shape)}, 50-iteration OSTR target", flush=True)
    assert bp.shape == (48, 256, 256) and torch.isfinite(target).all()
    return scan, bp, target


def test_forward_first_layer_matches_fp64(train_case, dev):
    _, bp, _ = train_case
    net = W.WarmStartNet(dev, seed=21)
    pred = net.forward(bp)
    torch.cuda.synchronize()
    assert torch.isfinite(pred).all()
    B = net._buffers(tuple(bp.shape))
    nz, ny, nx = bp.shape
    x0 = from_dev_bf16(B["x0"]).reshape(nz, ny, nx, 8)
    e1a = from_dev_bf16(B["e1a"]).reshape(nz, ny, nx, 16)
    hp = net.host_params()
    w = C_bf16(hp["e1a"][0])
    # an interior block with its one-voxel halo: conv there needs nothing outside
    z0, y0, x0_ = 20, 100, 120
    blk = x0[z0 - 1:z0 + 9, y0 - 1:y0 + 17, x0_ - 1:x0_ + 33]
    ref = np.maximum(C.conv(blk, w, hp["e1a"][1].astype(np.float64)), 0.0)[1:-1, 1:-1, 1:-1]
    got = e1a[z0:z0 + 8, y0:y0 + 16, x0_:x0_ + 32]
    assert rel_l2(got, ref) <= 1e-2, rel_l2(got, ref)


def C_bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def test_backward_deterministic_and_output_wgrad_matches_fp64(train_case, dev):
    _, bp, target = train_case
    net = W.WarmStartNet(dev, seed=22)
    grads = []
    for _ in range(2):
        net.forward(bp)
        net.loss_backward(target)
        torch.cuda.synchronize()
        grads.append(net.grads.clone())
    assert torch.equal(grads[0], grads[1])
    B = net._buffers(tuple(bp.shape))
    nz, ny, nx = bp.shape
    d1 = from_dev_bf16(B["d1"]).reshape(nz, ny, nx, 16)
    dout = from_dev_bf16(B["dout"]).reshape(nz, ny, nx, 16)
    assert not dout[..., 1:].any()
    _, dw, db = C.conv_grads(d1, np.zeros((16, 27, 16)), dout)
    gw = net.weight("out", net.grads).cpu().numpy()
    gb = net.bias("out", net.grads).cpu().numpy()
    print(f"[train5] out-layer weight gradient rel L2 {rel_l2(gw, dw):.2e}, "
          f"bias {rel_l2(gb, db):.2e}", flush=True)
    assert rel_l2(gw, dw) <= 1e-3 and rel_l2(gb, db) <= 1e-3


def test_adam_steps_lower_the_loss(train_case, dev):
    _, bp, target = train_case
    net = W.WarmStartNet(dev, seed=23)
    losses = []
    for _ in range(16):
        net.forward(bp)
        losses.append(net.loss_backward(target))
        net.adam_step(lr=2e-4)
    torch.cuda.synchronize()
    print("[train5] losses " + " ".join(f"{v:.3e}" for v in losses), flush=True)
    assert np.isfinite(losses).all() and losses[-1] < 0.95 * losses[0], losses
