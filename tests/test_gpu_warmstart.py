"""The trained warm start on the 200x100x20 / 9-view workload (seeds 1 / 1001)
against the FP64 converged image x* (tests/golden/xstar_small.npz), and the
fallback.

What the committed network gains (measured, asserted here): its image has a
lower FP64 objective than the backprojection it starts from and is closer to
x*, after k = 2, 5, 10 accelerated OSTR iterations the warm-started images
stay closer to x* than those from the backprojection, and the warm start (its
residual low-passed in-plane) reaches 1e-3 of x* in fewer iterations. The
smoothing kernel is checked against scipy's separable filter.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(dev):
    from newrecon_amd.geometry import scan_small
    from newrecon_amd.timetotol import XSTAR, WEIGHTS, _data_checksum, load_xstar
    from newrecon_amd.warmstart import WarmStartNet, make_case
    scan = scan_small()
    xs, meta = load_xstar(XSTAR)
    _, y, b = make_case(scan, scan.phantom_seed, dev)
    assert _data_checksum(y) == meta["y_sha256"], "counts differ from those of x*"
    net = WarmStartNet(dev)
    net.load(WEIGHTS)
    return scan, y, b, torch.from_numpy(xs).to(dev), net


def rel(sysm, a, ref):
    return sysm.rel_rmse(a, ref)


def test_warm_start_is_closer_and_stays_closer(setup, dev):
    from newrecon_amd.ostr import OSTR
    from newrecon_amd.warmstart import recon_config, warm_started
    scan, y, b, xs, net = setup
    cfg = recon_config(scan)
    plain = OSTR(scan.geom, scan.angles, y, b, cfg=cfg, x0="bp", device=dev)
    warm = OSTR(scan.geom, scan.angles, y, b, cfg=cfg, x0="bp", device=dev)
    accepted, obj_bp, obj_net = warm_started(warm, net)
    assert accepted and obj_net < obj_bp
    assert rel(warm.sys, warm.x, xs) < 0.8 * rel(plain.sys, plain.x, xs)
    done = 0
    for k in (2, 5, 10):
        plain.iterate(k - done)
        warm.iterate(k - done)
        done = k
        assert rel(warm.sys, warm.estimate, xs) < rel(plain.sys, plain.estimate, xs), k


def test_fallback_keeps_backprojection_for_a_bad_network(setup, dev):
    """A network whose output raises the objective is rejected: the
    reconstruction starts from the backprojection (plain OSTR)."""
    from newrecon_amd.ostr import OSTR
    from newrecon_amd.warmstart import WarmStartNet, recon_config, warm_started
    scan, y, b, xs, _ = setup
    bad = WarmStartNet(dev, seed=1)
    hp = bad.host_params()
    w, bb = hp["out"]
    hp["out"] = (w * 200.0, bb + 1.0)  # large offsets everywhere
    bad.set_params(hp)
    rec = OSTR(scan.geom, scan.angles, y, b, cfg=recon_config(scan), x0="bp", device=dev)
    bp = rec.x.clone()
    accepted, obj_bp, obj_net = warm_started(rec, bad)
    assert not accepted and obj_net >= obj_bp
    assert torch.equal(rec.x, bp)


def test_smoothed_residual_matches_scipy(setup, dev):
    """pred = img + S(out)/scale with S the in-plane Gaussian (edges repeated)."""
    from scipy.ndimage import convolve1d
    from newrecon_amd.ostr import OSTR
    from newrecon_amd.warmstart import SCALE, gaussian_taps, pad4, recon_config
    scan, y, b, xs, net = setup
    rec = OSTR(scan.geom, scan.angles, y, b, cfg=recon_config(scan), x0="bp", device=dev)
    img = pad4(rec.x)
    for sigma in (0.7, 1.0, 2.0):
        got = net.forward(img, clamp=False, smooth_sigma=sigma).double().cpu().numpy()
        B = net._buffers(tuple(img.shape))
        out = B["out"].view(torch.bfloat16).view(-1, 16)[:, 0].double().cpu().numpy()
        out = out.reshape(img.shape)
        taps, r = gaussian_taps(sigma)
        assert r == min(4, int(np.ceil(3 * sigma))) and abs(float(taps.sum()) - 1) < 1e-6
        s = convolve1d(convolve1d(out, taps.astype(np.float64), axis=2, mode="nearest"),
                       taps.astype(np.float64), axis=1, mode="nearest")
        want = img.double().cpu().numpy() + s / SCALE
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 1e-6, (sigma, err)


def test_warm_start_reaches_the_tolerance_in_fewer_iterations(setup, dev):
    from newrecon_amd.ostr import OSTR
    from newrecon_amd.warmstart import recon_config, warm_started
    scan, y, b, xs, net = setup
    cfg = recon_config(scan)
    iters = {}
    for name in ("plain", "warm"):
        rec = OSTR(scan.geom, scan.angles, y, b, cfg=cfg, x0="bp", device=dev)
        if name == "warm":
            accepted, _, _ = warm_started(rec, net)
            assert accepted
        k = 0
        while k < 4000:
            rec.iterate(1)
            k += 1
            if rel(rec.sys, rec.estimate, xs) < 1e-3:
                break
        iters[name] = k
    print(f"iterations to 1e-3: {iters}")
    assert iters["plain"] < 4000 and iters["warm"] < iters["plain"], iters
