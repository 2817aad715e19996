"""2x2 supersampled forward (ostr_fproj_4r, the fprojectCB1_4R_GPU_SART
counterpart) vs the CPU oracle on identical inputs: sub-bin sums and merged
bins <= 1e-6 of the oracle's peak (the walker shares the oracle's expression
graph), misses exactly 0 on both sides, partial tiles at the detector edge."""
import numpy as np
import pytest
import torch

from newrecon_amd.geometry import Geometry, arc_angles, regroup_index
from newrecon_amd.ostr import DBTSystem, OSTRError

pytestmark = pytest.mark.gpu


def rel_max(a, ref):
    return float(np.abs(np.asarray(a, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)


def test_fproj_4r_matches_oracle(small_case, dev, oracle_lib):
    s, x, _, _ = small_case
    idx = regroup_index(s.n_views, 3)
    sysr = DBTSystem(s.geom, s.angles, index=idx, device=dev)
    for m in range(3):
        a0, a1 = 3 * m, 3 * m + 3
        ref, ref4 = oracle_lib.fproj_4r(s.geom, s.angles, x, a0, a1, index=idx)
        prj, prj4 = sysr.fproj_4r(T(x, dev), a0, a1, keep_4r=True)
        prj, prj4 = prj.cpu().numpy(), prj4.cpu().numpy()
        assert rel_max(prj4, ref4) <= 1e-6, rel_max(prj4, ref4)
        assert rel_max(prj, ref) <= 1e-6, rel_max(prj, ref)
        assert np.array_equal(prj4 == 0, ref4 == 0)
        # without the sub-bin buffer: the same merged bins
        alone = sysr.fproj_4r(T(x, dev), a0, a1).cpu().numpy()
        np.testing.assert_array_equal(alone, prj)


def test_fproj_4r_partial_tiles_and_reference_name(dev, oracle_lib):
    # 2*nbx = 74 sub-columns (a partial 64-lane tile), 2*nby = 26 sub-rows (partial 4-row tile)
    g = Geometry.half_cone(20, 12, 6, 1.0, 1.0, 2.0, 37, 13, 1.3, 1.3)
    ang = arc_angles(5, 7.5)
    rng = np.random.default_rng(31)
    x = rng.random(g.vol_shape).astype(np.float32)
    sysg = DBTSystem(g, ang, device=dev)
    ref, ref4 = oracle_lib.fproj_4r(g, ang, x, 1, 5)
    prj = torch.full((4, g.nby, g.nbx), -1.0, device=dev)
    prj4 = torch.full((4, 2 * g.nby, 2 * g.nbx), -1.0, device=dev)
    sysg.fprojectCB1_4R_GPU_SART(T(x, dev), prj, prj4, 1, 5)
    assert rel_max(prj4.cpu().numpy(), ref4) <= 1e-6
    assert rel_max(prj.cpu().numpy(), ref) <= 1e-6
    assert (prj4 >= 0).all() and (prj >= 0).all()   # every element written
    with pytest.raises(OSTRError):
        sysg.fproj_4r(T(x, dev), 0, 6)
