"""The trained warm start on the 200x100x20 / 9-view workload (seeds 1 / 1001)
against the FP64 converged image x* (tests/golden/xstar_small.npz), and the
fallback.

What the committed network gains (measured, asserted here): its image has a
lower FP64 objective than the backprojection it starts from and is closer to
x*, and after k = 2, 5, 10 accelerated OSTR iterations the warm-started
images stay closer to x* than those from the backprojection. (Its effect on
the time to 1e-3 is reported by bench.py; it is not asserted.)
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
