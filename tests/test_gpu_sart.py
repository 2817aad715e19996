"""OS-SART kernels (through the C ABI) vs the CPU oracle on identical inputs.

Bars: the SART forward and row sum <= 1e-6 of the oracle's peak (the walker
shares the oracle's expression graph); SART_prj_diff bit-exact (IEEE division
on both sides); the backprojection update and the Huber prior (FP32 on the
GPU, double terms in the reference) <= 1e-5 of the peak; the OS-SART loop,
fused or launch by launch, per-iteration relative RMSE <= 1e-4 (the parity
bar of the path); the fused loop equal to the launch-by-launch loop bit for
bit.
"""
import numpy as np
import pytest
import torch

from newrecon_amd.geometry import GAMMA_MODES, Geometry, arc_angles, regroup_index
from newrecon_amd.ostr import DBTSystem, OSTRError
from newrecon_amd.sart import OSSART

pytestmark = pytest.mark.gpu


def rel_max(a, ref):
    return float(np.abs(np.asarray(a, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))


def rel_rmse(a, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - ref) / max(np.linalg.norm(ref), 1e-30))


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)


@pytest.fixture(scope="module")
def sart_case(small_case, dev):
    s, x, y, b = small_case
    idx = regroup_index(s.n_views, 3)
    sysr = DBTSystem(s.geom, s.angles, index=idx, device=dev)
    p = np.log(b / np.maximum(y, 1.0)).astype(np.float32)
    return s, x, p, idx, sysr


@pytest.mark.parametrize("cos_weight", [False, True])
def test_fproj_sart_matches_oracle(sart_case, dev, oracle_lib, cos_weight):
    s, x, _, idx, sysr = sart_case
    for m in range(3):
        a0, a1 = 3 * m, 3 * m + 3
        prj, norm = sysr.fproj_sart(T(x, dev), a0, a1, cos_weight=cos_weight)
        ref = oracle_lib.fproj(s.geom, s.angles, x, a0, a1, index=idx, cos_weight=cos_weight)
        rn = oracle_lib.gamma(s.geom, s.angles, a0, a1, GAMMA_MODES["row_sum"], index=idx)
        prj, norm = prj.cpu().numpy(), norm.cpu().numpy()
        assert rel_max(prj, ref) <= 1e-6, rel_max(prj, ref)
        assert rel_max(norm, rn) <= 1e-6, rel_max(norm, rn)
        assert np.array_equal(prj == 0, ref == 0) and np.array_equal(norm == 0, rn == 0)


def test_reference_named_forward(sart_case, dev):
    s, x, _, _, sysr = sart_case
    xt = T(x, dev)
    prj, norm = sysr.sinogram(3), sysr.sinogram(3)
    sysr.fprojectCB_1R_GPU_SART_cos(xt, prj, norm, 3, 6)
    p2, n2 = sysr.fproj_sart(xt, 3, 6, cos_weight=True)
    assert torch.equal(prj, p2) and torch.equal(norm, n2)
    sysr.fprojectCB_1R_GPU_SART(xt, prj, norm, 3, 6)
    p3, _ = sysr.fproj_sart(xt, 3, 6, cos_weight=False)
    assert torch.equal(prj, p3)


def test_prj_diff_bit_exact(sart_case, dev, oracle_lib):
    s, x, p, idx, sysr = sart_case
    rng = np.random.default_rng(11)
    est = rng.random((3,) + s.geom.view_shape).astype(np.float32)
    norm = (rng.random((3,) + s.geom.view_shape) * 40).astype(np.float32)
    norm[:, ::5, :] = 0.0
    ps = p[idx[:3]]
    ref = oracle_lib.sart_prj_diff(ps, est, norm)
    got = sysr.SART_prj_diff(torch.empty(est.shape, device=dev), T(ps, dev), T(est, dev),
                             T(norm, dev), 0, 3).cpu().numpy()
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("with_prior", [False, True])
def test_bproj_sart_matches_oracle(sart_case, dev, oracle_lib, with_prior):
    s, x, _, idx, sysr = sart_case
    rng = np.random.default_rng(12)
    diff = (rng.standard_normal((3,) + s.geom.view_shape) * 1e-3).astype(np.float32)
    x0 = (x + rng.random(x.shape) * 1e-3).astype(np.float32)
    prior = (rng.standard_normal(x.shape) * 1e-2).astype(np.float32) if with_prior else None
    beta = 0.05 if with_prior else 0.0
    for m in range(3):
        a0, a1 = 3 * m, 3 * m + 3
        ref = oracle_lib.bproj_sart(s.geom, s.angles, diff, a0, a1, x0, 0.9, index=idx,
                                    prior=prior, beta=beta)
        img = T(x0, dev)
        if with_prior:
            sysr.bprojectCB_4B_GPU_SART_R(img, T(diff, dev), T(prior, dev), a0, a1, 0.9, beta)
        else:
            sysr.bprojectCB_4B_GPU_SART(img, T(diff, dev), a0, a1, 0.9)
        got = img.cpu().numpy()
        assert rel_max(got, ref) <= 1e-5, (m, rel_max(got, ref))


def test_huber_prior_matches_oracle(sart_case, dev, oracle_lib):
    """The GPU evaluates the Huber terms in FP32 (the reference in double):
    within 1e-5 of the peak, accumulating onto the buffer's contents."""
    s, x, _, _, sysr = sart_case
    rng = np.random.default_rng(13)
    start = (rng.standard_normal(x.shape) * 1e-3).astype(np.float32)
    for delta in (5e-4, 2e-3):
        ref = oracle_lib.prior_huber(s.geom, x, delta, prior=start)
        got = sysr.prior_GPU_SART(T(start, dev), T(x, dev), delta).cpu().numpy()
        err = rel_max(got, ref)
        print(f"Huber prior delta={delta}: max error / peak {err:.2e}")
        assert err <= 1e-5, err


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("beta,cos_weight", [(0.0, False), (0.0, True), (-2e-3, True)])
def test_os_sart_per_iteration(sart_case, dev, oracle_lib, beta, cos_weight, fused):
    s, x, p, _, _ = sart_case
    n_it = 4
    ref, snaps = oracle_lib.run_sart(s.geom, s.angles, p, n_subsets=3, n_iters=n_it, lam=0.8,
                                     beta=beta, delta=1e-3, cos_weight=cos_weight,
                                     snapshots=True)
    rec = OSSART(s.geom, s.angles, p, n_subsets=3, lam=0.8, beta=beta, delta=1e-3,
                 cos_weight=cos_weight, fused=fused, device=dev)
    errs = []
    rec.iterate(n_it, callback=lambda it, xi: errs.append(rel_rmse(xi.cpu().numpy(),
                                                                   snaps[it - 1])))
    print("OS-SART per-iteration rel RMSE", errs)
    assert max(errs) <= 1e-4, errs
    # and the loop reconstructs: the image moves towards the phantom
    assert rel_rmse(rec.x.cpu().numpy(), x) < rel_rmse(np.zeros_like(x), x)


@pytest.mark.parametrize("beta", [0.0, -2e-3])
def test_fused_equals_reference_launchers(sart_case, dev, beta):
    """Two fused launches per subset give the five launchers' result bit for bit."""
    s, _, p, _, _ = sart_case
    a = OSSART(s.geom, s.angles, p, n_subsets=3, lam=0.8, beta=beta, delta=1e-3,
               cos_weight=True, fused=True, device=dev)
    b = OSSART(s.geom, s.angles, p, n_subsets=3, lam=0.8, beta=beta, delta=1e-3,
               cos_weight=True, fused=False, device=dev)
    for _ in range(3):
        a.iterate(1)
        b.iterate(1)
        assert torch.equal(a.x, b.x)


def test_sart_errors(dev):
    g = Geometry.half_cone(8, 6, 4, 1.0, 1.0, 2.0, 16, 12, 1.0, 1.0)
    ang = arc_angles(66, 0.5)
    sysg = DBTSystem(g, ang, device=dev)
    img = sysg.volume()
    diff = sysg.sinogram(66)
    with pytest.raises(OSTRError):
        sysg.bproj_sart(diff, 0, 65, img, 1.0)
    with pytest.raises(OSTRError):
        sysg.prior_sart(img, 0.0)
    with pytest.raises(OSTRError):
        sysg.fproj_sart(img, 0, 67)
    with pytest.raises(ValueError):
        OSSART(g, ang, np.zeros((66,) + g.view_shape, np.float32), n_subsets=1, device=dev)
    rows = torch.zeros((65, g.nby, g.nbx), device=dev)
    with pytest.raises(OSTRError):
        sysg.sart_update_fused(img, img, rows, 0, 2, 1.0)
    with pytest.raises(OSTRError):
        sysg.sart_update_fused(img, sysg.volume(), rows, 0, 2, 1.0, beta=1.0, delta=0.0)
    with pytest.raises(OSTRError):
        sysg.sart_update_fused(img, sysg.volume(), rows, 0, 65, 1.0)
    with pytest.raises(OSTRError):
        sysg.sart_fproj_diff(img, rows, 0, 2, diff_rows=rows)
    with pytest.raises(OSTRError):
        sysg.prior_sart(img, 1e-3, prior=img)


def test_sart_ragged_sizes_and_empty_ranges(dev, oracle_lib):
    """A volume and detector that fill no tile (13x7x5 voxels, 37x9 bins, 3 views):
    the SART forward, the update with prior and the fused loop against the
    oracle; empty slot ranges are accepted and change nothing."""
    from newrecon_amd.geometry import scan_tiny
    s = scan_tiny(nx=13, ny=7, nz=5, nbx=37, nby=9, n_views=3)
    g = s.geom
    rng = np.random.default_rng(17)
    x = (rng.random(g.vol_shape) * 0.02 + 0.005).astype(np.float32)
    p = oracle_lib.fproj(g, s.angles, x, 0, 3, cos_weight=False)
    S = DBTSystem(g, s.angles, device=dev)
    prj, norm = S.fproj_sart(T(x, dev), 0, 3)
    assert rel_max(prj.cpu().numpy(), p) <= 1e-6
    ref, snaps = oracle_lib.run_sart(g, s.angles, p, n_subsets=3, n_iters=3, lam=0.9, beta=-1e-3,
                                     delta=1e-3, snapshots=True)
    for fused in (True, False):
        rec = OSSART(g, s.angles, p, n_subsets=3, lam=0.9, beta=-1e-3, delta=1e-3, fused=fused,
                     device=dev)
        errs = []
        rec.iterate(3, callback=lambda it, xi: errs.append(rel_rmse(xi.cpu().numpy(),
                                                                       snaps[it - 1])))
        assert max(errs) <= 1e-4, (fused, errs)
    img = T(x, dev)
    before = img.clone()
    S.bproj_sart(S.sinogram(1), 2, 2, img, 1.0)             # empty range
    S.fproj_sart(img, 1, 1, S.sinogram(1), S.sinogram(1))   # empty range
    assert torch.equal(img, before)
