"""Batched OSTR (several uneven cases in one launch per kernel) against the
single-case fused kernels, and the autotuner's cache behaviour.

Bar: bit-identical images per case (every voxel and bin is computed by the
same device code in the same order; the tile shape and the persistent grid
only change which work-group computes it), for several tile shapes, both
grid sizings, FP32 and BF16 subset sinograms, over steps where the cases are
at different subsets (M = 3 / 5 / 5).
"""
import numpy as np
import pytest
import torch

from newrecon_amd.batch import BatchOSTR, TileCfg
from newrecon_amd.geometry import OSTRConfig, scan_batch
from newrecon_amd.ostr import OSTR, OSTRError
from newrecon_amd.phantom import breast_phantom, poisson_counts

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cases(dev):
    """The uneven batch table at 1/8 scale: 8 cases, 2..10 slices, 9/15/25 views."""
    from newrecon_amd.ostr import DBTSystem
    out = []
    for s in scan_batch(scale=8):
        x = breast_phantom(s.geom, s.phantom_seed, (5, 10), 3)
        sysm = DBTSystem(s.geom, s.angles, device=dev)
        l = sysm.fproj(torch.from_numpy(x).to(dev), 0, s.n_views).cpu().numpy()
        sysm.close()
        y = poisson_counts(l, s.i0, s.noise_seed)
        out.append((s, y, np.full_like(y, s.i0)))
    return out


def make(cases, dev, prj_dtype, prior="hyperbola", r=False):
    recs = []
    for s, y, b in cases:
        cfg = OSTRConfig(n_subsets=s.recon.n_subsets, beta=1e3, delta=5e-4, prior=prior)
        rr = (0.01 * y).astype(np.float32) if r else None
        recs.append(OSTR(s.geom, s.angles, y, b, r=rr, cfg=cfg, x0="bp", prj_dtype=prj_dtype,
                         device=dev))
    return recs


def run_single(recs, steps):
    for rec in recs:
        rec.precompute()
        for k in range(steps):
            rec._subset(k % rec.M)
    torch.cuda.synchronize()
    return [rec.x.clone() for rec in recs]


TILES = [
    dict(fwd_rows=4, fwd_groups_per_cu=0, upd_rows=8, upd_waves=4, upd_groups_per_cu=0),
    dict(fwd_rows=1, fwd_groups_per_cu=8, upd_rows=4, upd_waves=4, upd_groups_per_cu=4),
    dict(fwd_rows=8, fwd_groups_per_cu=32, upd_rows=8, upd_waves=2, upd_groups_per_cu=16),
    dict(fwd_rows=2, fwd_groups_per_cu=1, upd_rows=8, upd_waves=4, upd_groups_per_cu=1),
]


@pytest.mark.parametrize("prj_dtype", [torch.float32, torch.bfloat16])
def test_batch_is_bit_identical_to_single_case(cases, dev, prj_dtype):
    steps = 6  # crosses the end of the M = 3 and M = 5 subset cycles
    ref = run_single(make(cases, dev, prj_dtype), steps)
    for tile in TILES:
        print(f"batch prj={prj_dtype} tile={tile}", flush=True)
        recs = make(cases, dev, prj_dtype)
        bt = BatchOSTR(recs, tile=tile)
        bt.step(steps)
        torch.cuda.synchronize()
        for k, (got, want) in enumerate(zip(bt.images(), ref)):
            assert torch.equal(got, want), (tile, k, (got - want).abs().max().item())
        assert [r.iteration for r in recs] == [2, 1, 1, 2, 1, 1, 2, 1]
        bt.close()


def test_batch_lange_with_scatter(cases, dev):
    ref = run_single(make(cases[:3], dev, torch.float32, prior="lange", r=True), 4)
    recs = make(cases[:3], dev, torch.float32, prior="lange", r=True)
    bt = BatchOSTR(recs, tile=TILES[1])
    bt.step(4)
    torch.cuda.synchronize()
    for got, want in zip(bt.images(), ref):
        assert torch.equal(got, want)


def test_batch_argument_errors(cases, dev):
    recs = make(cases[:2], dev, torch.float32)
    bt = BatchOSTR(recs)
    with pytest.raises(OSTRError, match="fwd_rows"):
        bt.fproj_h(TileCfg(fwd_rows=3, fwd_groups_per_cu=0, upd_rows=8, upd_waves=4,
                           upd_groups_per_cu=0))
    with pytest.raises(OSTRError, match="update shape"):
        bt.update(TileCfg(fwd_rows=4, fwd_groups_per_cu=0, upd_rows=4, upd_waves=2,
                          upd_groups_per_cu=0))
    recs[1].x_next = recs[1].x  # aliasing is refused before anything is launched
    with pytest.raises(OSTRError, match="alias"):
        bt.update()
    with pytest.raises(ValueError):
        BatchOSTR([make(cases[:1], dev, torch.float32)[0], make(cases[:1], dev,
                   torch.bfloat16)[0]])


def test_autotune_choice_is_cached_and_reproducible(cases, dev, tmp_path):
    from newrecon_amd.autotune import tune
    cache = str(tmp_path / "tune.json")
    recs = make(cases[:4], dev, torch.bfloat16)
    bt = BatchOSTR(recs)
    first = tune(bt, cache_path=cache, reps=2)
    assert not first["cached"] and len(first["fwd_ms"]) > 1
    again = tune(bt, cache_path=cache)
    assert again["cached"] and again["tile"] == first["tile"]
    assert bt.tile.as_dict() == first["tile"]
    # the tuned shapes give the same images as the single-case kernels
    ref = run_single(make(cases[:4], dev, torch.bfloat16), 3)
    bt.step(3)
    torch.cuda.synchronize()
    for got, want in zip(bt.images(), ref):
        assert torch.equal(got, want)


def test_batch_full_scale_bit_identical(dev):
    """The batch table at full scale (30..80 slices, 800x400..1400x700 voxels,
    9/15/25 views of 2048x1664 bins, BF16 subset sinograms): two batched steps
    equal the single-case kernels bit for bit. The volumes are smooth seeded
    random fields (upsampled on the GPU) and the counts their noise-free
    transmission, which exercises the same kernels as the phantom data at a
    fraction of the set-up time."""
    from newrecon_amd.ostr import DBTSystem
    data = []
    for s in scan_batch():
        g = s.geom
        gen = torch.Generator(device=dev).manual_seed(s.phantom_seed)
        coarse = torch.rand((max(g.nz // 4, 1), g.ny // 16, g.nx // 16), generator=gen,
                            device=dev)
        x = torch.nn.functional.interpolate(coarse[None, None], size=g.vol_shape,
                                            mode="trilinear")[0, 0].contiguous() * 0.05
        sysm = DBTSystem(g, s.angles, device=dev)
        l = sysm.fproj(x, 0, s.n_views)
        sysm.close()
        y = (s.i0 * torch.exp(-l)).cpu().numpy()
        del x, l
        data.append((s, y, np.full_like(y, s.i0)))
    print("full-scale batch: data ready", flush=True)
    steps = 2
    ref = run_single(make(data, dev, torch.bfloat16), steps)
    recs = make(data, dev, torch.bfloat16)
    bt = BatchOSTR(recs, tile=TILES[0])
    bt.step(steps)
    torch.cuda.synchronize()
    for k, (got, want) in enumerate(zip(bt.images(), ref)):
        assert torch.equal(got, want), (k, (got - want).abs().max().item())
    bt.close()
