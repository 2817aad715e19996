"""HIP kernels (through the C ABI) vs the CPU oracle on identical inputs.

Bars: per-kernel max error relative to the oracle's peak <= 1e-5 (FP32; the
walker and elementwise ops share the oracle's expression graph, the SBP and
prior use hardware reciprocal / rsqrt within 2 ulp); OSTR per-iteration
relative RMSE <= 1e-4 (the parity bar of the path); exact integer equality for
the nnz count and the regroup permutation.
"""
import numpy as np
import pytest
import torch

from newrecon_amd.geometry import OSTRConfig, regroup_index, scan_small, scan_tiny
from newrecon_amd.ostr import OSTR, DBTSystem, OSTRError

pytestmark = pytest.mark.gpu


def rel_max(a, ref):
    return float(np.abs(np.asarray(a, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))


def rel_rmse(a, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - ref) / max(np.linalg.norm(ref), 1e-30))


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)


@pytest.fixture(scope="module")
def sysm(small_case, dev):
    s = small_case[0]
    return DBTSystem(s.geom, s.angles, device=dev)


def test_fproj_matches_oracle(small_case, sysm, dev, oracle_lib):
    s, x, y, b = small_case
    for cos_w in (True, False):
        ref = oracle_lib.fproj(s.geom, s.angles, x, 0, s.n_views, cos_weight=cos_w)
        got = sysm.fproj(T(x, dev), 0, s.n_views, cos_weight=cos_w).cpu().numpy()
        assert rel_max(got, ref) <= 1e-6, rel_max(got, ref)
        # rays that miss the volume are exactly zero on both sides
        assert np.array_equal(got == 0, ref == 0)


def test_fproj_subset_slots_follow_index(small_case, dev, oracle_lib):
    s, x, _, _ = small_case
    idx = regroup_index(s.n_views, 3)
    sysr = DBTSystem(s.geom, s.angles, index=idx, device=dev)
    for m in range(3):
        a0, a1 = 3 * m, 3 * m + 3
        ref = oracle_lib.fproj(s.geom, s.angles, x, a0, a1, index=idx)
        got = sysr.fproj(T(x, dev), a0, a1).cpu().numpy()
        assert rel_max(got, ref) <= 1e-6


@pytest.mark.parametrize("mode", ["ray_length", "row_sum"])
def test_gamma_matches_oracle(small_case, sysm, mode, oracle_lib):
    s = small_case[0]
    ref = oracle_lib.gamma(s.geom, s.angles, 0, s.n_views, {"ray_length": 0, "row_sum": 1}[mode])
    got = sysm.gamma(0, s.n_views, mode).cpu().numpy()
    assert rel_max(got, ref) <= 1e-6


def test_nnz_count_exact(small_case, sysm, oracle_lib):
    s = small_case[0]
    for a0, a1 in ((0, 3), (0, s.n_views)):
        assert sysm.count_nnz(a0, a1) == oracle_lib.count_nnz(s.geom, s.angles, a0, a1)


def test_sbp_matches_oracle(small_case, sysm, dev, oracle_lib):
    s, x, y, b = small_case
    rng = np.random.default_rng(5)
    prj = rng.standard_normal((s.n_views,) + s.geom.view_shape).astype(np.float32)
    base = rng.standard_normal(s.geom.vol_shape).astype(np.float32)
    ref = oracle_lib.bproj_sbp(s.geom, s.angles, prj, 0, s.n_views, img=base)
    got = sysm.bproj_sbp(T(prj, dev), 0, s.n_views, img=T(base, dev)).cpu().numpy()
    assert rel_max(got, ref) <= 1e-5, rel_max(got, ref)
    # sub-range of slots with a regrouped order accumulates the same way
    idx = regroup_index(s.n_views, 3)
    sysr = DBTSystem(s.geom, s.angles, index=idx, device=dev)
    ref = oracle_lib.bproj_sbp(s.geom, s.angles, prj[:3], 3, 6, index=idx)
    got = sysr.bproj_sbp(T(prj[:3], dev), 3, 6).cpu().numpy()
    assert rel_max(got, ref) <= 1e-5


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16])
def test_sbp_detector_edges_match_oracle(dev, oracle_lib, dtype):
    """A detector smaller than the volume's shadow in x and y, so one launch
    holds waves whose 8 rows all take the unmasked straight run of tap loads,
    waves with some rows off the detector, and waves with columns off it (the
    masked per-row loads); FP16 h also takes the last-row pair rule. Against the
    oracle at the same bar as the full-detector case."""
    from newrecon_amd.geometry import Geometry, arc_angles
    g = Geometry.half_cone(150, 80, 12, 0.5, 0.5, 1.0, 96, 60, 0.56, 0.56)
    ang = arc_angles(5, 7.5)
    rng = np.random.default_rng(17)
    prj = rng.standard_normal((len(ang),) + g.view_shape).astype(np.float32)
    prj_t = torch.from_numpy(prj).to(dev).to(dtype)
    prj_q = prj_t.float().cpu().numpy()
    S = DBTSystem(g, ang, device=dev)
    ref = oracle_lib.bproj_sbp(g, ang, prj_q, 0, len(ang))
    got = S.bproj_sbp(prj_t, 0, len(ang)).cpu().numpy()
    assert np.count_nonzero(ref) < ref.size  # some voxels see no detector at all
    assert rel_max(got, ref) <= 1e-5, rel_max(got, ref)


def test_matched_backprojector_is_the_adjoint(small_case, sysm, dev, oracle_lib):
    """<A x, y> = <x, A^T y> for the GPU pair (FP32), and A^T agrees with the
    oracle's exact geometric transpose (voxel-driven clip, FP64)."""
    s, x, _, _ = small_case
    rng = np.random.default_rng(11)
    yv = rng.standard_normal((s.n_views,) + s.geom.view_shape).astype(np.float32)
    xv = rng.random(s.geom.vol_shape).astype(np.float32)
    Ax = sysm.fproj(T(xv, dev), 0, s.n_views).cpu().numpy().astype(np.float64)
    Aty = sysm.bproj_matched(T(yv, dev), 0, s.n_views).cpu().numpy()
    lhs = float((Ax * yv).sum())
    rhs = float((xv.astype(np.float64) * Aty).sum())
    # y is zero-mean, so the sums cancel heavily: bound the FP32 difference by
    # the magnitude of the terms, sum |Ax| |y|
    scale = float((np.abs(Ax) * np.abs(yv)).sum())
    assert abs(lhs - rhs) <= 1e-7 * scale, (lhs, rhs, scale)
    ref = oracle_lib.bproj_clip(s.geom, s.angles, yv.astype(np.float64), 0, s.n_views)
    assert rel_max(Aty, ref) <= 1e-5
    # row-major input gives the same result; subsets through the index table too
    rows = sysm.bproj_matched(T(yv.transpose(0, 2, 1), dev), 0, s.n_views, rows=True)
    assert np.array_equal(rows.cpu().numpy(), Aty)
    idx = regroup_index(s.n_views, 3)
    sysr = DBTSystem(s.geom, s.angles, index=idx, device=dev)
    got = sysr.bproj_matched(T(yv[:3], dev), 3, 6).cpu().numpy()
    ref = oracle_lib.bproj_clip(s.geom, s.angles, yv[:3].astype(np.float64), 3, 6, index=idx)
    assert rel_max(got, ref) <= 1e-5


def test_bp_init_matches_oracle(small_case, sysm, dev, oracle_lib):
    """ReconDBT/BP.m one-step image: from line integrals and from counts.

    BP.m's per-view term G^T q / G^T 1 is discontinuous where a view's only
    ray through a voxel grazes its boundary: G^T 1 is then a tiny positive
    number (the voxel counts for that view) or exactly 0 (it does not), and
    FP32 vs FP64 rounding decides which. Such voxels (a few per 1e5) may
    differ by one view's share, max/V; all others agree to 1e-5."""
    s, x, y, b = small_case
    p = oracle_lib.fproj(s.geom, s.angles, x, 0, s.n_views)
    pc = np.log(b.astype(np.float64) / np.maximum(y, 1.0))
    for got, ref in (
            (sysm.bp_init(p=T(p, dev)).cpu().numpy(),
             oracle_lib.bp_init(s.geom, s.angles, p.astype(np.float64))),
            (sysm.bp_init(y=T(y, dev), b=T(b, dev)).cpu().numpy(),
             oracle_lib.bp_init(s.geom, s.angles, pc))):
        peak = np.abs(ref).max()
        d = np.abs(got - ref)
        off = d > 1e-5 * peak
        assert off.sum() <= 1e-4 * ref.size, off.sum()
        assert d.max() <= peak / s.n_views


def test_row_major_layout_entry_points(small_case, sysm, dev, oracle_lib):
    """to_rows is an exact transpose; SBP of the row-major copy equals SBP of
    the reference layout; regroup(to_rows) = transpose of regroup."""
    s, x, y, b = small_case
    rng = np.random.default_rng(8)
    prj = rng.standard_normal((s.n_views,) + s.geom.view_shape).astype(np.float32)
    rows = sysm.to_rows(T(prj, dev)).cpu().numpy()
    assert np.array_equal(rows, prj.transpose(0, 2, 1))
    a = sysm.bproj_sbp(T(prj, dev), 0, s.n_views).cpu().numpy()
    r = sysm.bproj_sbp_rows(T(rows, dev), 0, s.n_views).cpu().numpy()
    assert np.array_equal(a, r)
    reg = sysm.regroup_prj(T(y, dev), 3, to_rows=True).cpu().numpy()
    assert np.array_equal(reg, oracle_lib.regroup(y, 3).transpose(0, 2, 1))


@pytest.mark.parametrize("potential,code,delta", [
    ("q2", 0, 5e-4), ("lange", 1, 5e-4), ("hyperbola", 2, 5e-4),
    # outside [2^-60, 2^60]: the kernels take omega through t/delta
    ("lange", 1, 2.0 ** -62), ("hyperbola", 2, 2.0 ** -62), ("hyperbola", 2, 2.0 ** 62)])
def test_prior_matches_oracle(small_case, sysm, dev, potential, code, delta, oracle_lib):
    s, x, _, _ = small_case
    rng = np.random.default_rng(6)
    xn = (x + 1e-3 * rng.standard_normal(x.shape)).astype(np.float32)
    rdd0 = rng.random(x.shape).astype(np.float32)
    rd0 = rng.standard_normal(x.shape).astype(np.float32)
    r_dd, r_d = oracle_lib.prior(s.geom, xn, code, delta, rdd=rdd0, rd=rd0)
    g_dd, g_d = sysm.prior(T(xn, dev), potential, delta, rdd=T(rdd0, dev), rd=T(rd0, dev))
    assert rel_max(g_dd.cpu().numpy(), r_dd) <= 1e-5
    assert rel_max(g_d.cpu().numpy(), r_d) <= 1e-5


def test_elementwise_ops(small_case, sysm, dev, oracle_lib):
    s, x, y, b = small_case
    rng = np.random.default_rng(7)
    r = (rng.random(y.shape) * 50).astype(np.float32)
    yy = y.copy()
    yy.ravel()[::97] = 0.0
    got = sysm.compute_yry(torch.empty_like(T(yy, dev)), T(yy, dev), T(r, dev)).cpu().numpy()
    assert np.array_equal(got, oracle_lib.yry(yy, r))
    l = rng.random(y.shape).astype(np.float32) * 3
    got = sysm.compute_h(torch.empty_like(T(y, dev)), T(y, dev), T(b, dev), T(l, dev),
                         T(r, dev)).cpu().numpy()
    assert rel_max(got, oracle_lib.compute_h(y, b, l, r)) <= 1e-6
    nv = x.size
    capl, rd, d, rdd = (rng.standard_normal(x.shape).astype(np.float32) for _ in range(4))
    d = np.abs(d) + 0.1
    rdd = np.abs(rdd)
    est = np.abs(rng.standard_normal(x.shape)).astype(np.float32)
    got = sysm.update_est(T(est, dev), T(capl, dev), T(rd, dev), T(d, dev), T(rdd, dev), 1e3,
                          3.0).cpu().numpy()
    ref = oracle_lib.update_est(est, capl, rd, d, rdd, 1e3, 3.0)
    assert np.array_equal(got, ref)
    reg = sysm.regroup_prj(T(y, dev), 3).cpu().numpy()
    assert np.array_equal(reg, oracle_lib.regroup(y, 3))


def test_precompute_d_matches_oracle(small_case, sysm, dev, oracle_lib):
    s, x, y, b = small_case
    for mode, code in (("row_sum", 1), ("ray_length", 0)):
        gam = oracle_lib.gamma(s.geom, s.angles, 0, s.n_views, code)
        gc = oracle_lib.yry(y) * gam
        ref = oracle_lib.bproj_sbp(s.geom, s.angles, gc, 0, s.n_views)
        got = sysm.precompute_d(T(y, dev), None, mode).cpu().numpy()
        assert rel_max(got, ref) <= 1e-5


@pytest.mark.parametrize("fused", [False, True])
def test_ostr_small_per_iteration_parity(small_case, dev, fused, oracle_lib):
    """Plain FP32 OSTR on the 200x100x20 / 9-view workload: <= 1e-4 rel RMSE
    against the FP32 oracle after every one of 10 iterations."""
    s, x, y, b = small_case
    cfg = OSTRConfig(n_subsets=3, n_iters=10, beta=1e3, delta=5e-4, prior="hyperbola",
                     gamma_mode="row_sum", guard=True, fused=fused)
    _, snaps = oracle_lib.run(s.geom, s.angles, y, b, n_subsets=3, n_iters=10, beta=1e3,
                              delta=5e-4, potential=2, gamma_mode=1, guard=True, snapshots=True)
    rec = OSTR(s.geom, s.angles, y, b, cfg=cfg, device=dev)
    errs = []
    rec.iterate(10, callback=lambda it, xg: errs.append(rel_rmse(xg.cpu().numpy(), snaps[it - 1])))
    assert max(errs) <= 1e-4, errs


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("M", [1, 3])
def test_ostr_matched_per_iteration_parity(small_case, dev, M, fused, oracle_lib):
    """OS-SQS with the exact transpose (d = G^T(yry G 1), capL = M G^T h):
    <= 1e-4 rel RMSE against the FP32 oracle (clip transpose) after every
    iteration, and the FP64 objective decreases monotonically for M = 1."""
    s, x, y, b = small_case
    n = 6
    cfg = OSTRConfig(n_subsets=M, beta=1e3, delta=5e-4, prior="hyperbola",
                     backprojector="matched", fused=fused)
    _, snaps = oracle_lib.run(s.geom, s.angles, y, b, n_subsets=M, n_iters=n, beta=1e3,
                              delta=5e-4, potential=2, snapshots=True, backprojector=1)
    rec = OSTR(s.geom, s.angles, y, b, cfg=cfg, device=dev)
    assert rec.fused == fused
    errs, objs = [], []

    def cb(it, xg):
        errs.append(rel_rmse(xg.cpu().numpy(), snaps[it - 1]))
        objs.append(rec.objective())
    rec.iterate(n, callback=cb)
    assert max(errs) <= 1e-4, errs
    if M == 1:
        assert all(b2 <= b1 for b1, b2 in zip(objs, objs[1:])), objs


def test_accelerated_sqs_matches_fp64_solver(small_case, dev, oracle_lib):
    """Momentum OSTR (M = 1, exact transpose, gradient restart) on the GPU in
    FP32 against the FP64 accelerated solver of oracle/converged.py from the
    same backprojected start: <= 1e-4 rel RMSE after 30 iterations."""
    from oracle.converged import SQS, solve
    s, x, y, b = small_case
    n = 30
    sqs = SQS(s.geom, s.angles, y, b, beta=1e3, delta=5e-4, nthreads=8)
    cfg = OSTRConfig(n_subsets=1, beta=1e3, delta=5e-4, prior="hyperbola",
                     backprojector="matched", momentum=True)
    rec = OSTR(s.geom, s.angles, y, b, cfg=cfg, x0="bp", device=dev)
    x0 = rec.x.double().cpu().numpy()
    ref, info = solve(sqs, x0, n)
    got = rec.iterate(n).cpu().numpy()
    assert rec.restarts == info["restarts"]
    assert rel_rmse(got, ref) <= 1e-4, rel_rmse(got, ref)


def test_fused_equals_unfused(small_case, dev):
    s, x, y, b = small_case
    out = []
    for fused in (False, True):
        cfg = OSTRConfig(n_subsets=3, n_iters=3, fused=fused)
        out.append(OSTR(s.geom, s.angles, y, b, cfg=cfg, device=dev).iterate(3).cpu().numpy())
    assert rel_max(out[1], out[0]) <= 1e-6


@pytest.mark.parametrize("prior,M,momentum,delta", [
    ("hyperbola", 3, False, 5e-4), ("lange", 1, True, 5e-4), ("q2", 3, False, 5e-4),
    ("hyperbola", 3, False, 2.0 ** -62), ("lange", 3, False, 2.0 ** 62)])
def test_fused_matched_equals_unfused(small_case, dev, prior, M, momentum, delta):
    """The fused matched update (exact transpose + prior + update in one
    launch) equals ostr_bproj_matched + ostr_prior + ostr_update_est bit for bit,
    also for a delta whose square leaves the normal range."""
    s, x, y, b = small_case
    out = []
    for fused in (False, True):
        cfg = OSTRConfig(n_subsets=M, prior=prior, backprojector="matched", momentum=momentum,
                         fused=fused, delta=delta)
        out.append(OSTR(s.geom, s.angles, y, b, cfg=cfg, x0="bp", device=dev).iterate(4))
    assert torch.equal(out[0], out[1])


@pytest.mark.parametrize("prior,gmode", [("lange", "ray_length"), ("q2", "row_sum")])
def test_ostr_reference_modes(small_case, dev, prior, gmode, oracle_lib):
    """The reference's own settings: Lange prior (Recon_Method 1) with gamma =
    ray length (Tomo_OSTR.cu:1172), and Q2 (Recon_Method 0)."""
    s, x, y, b = small_case
    code = {"q2": 0, "lange": 1}[prior]
    cfg = OSTRConfig(n_subsets=3, n_iters=4, beta=1e3, delta=5e-4, prior=prior,
                     gamma_mode=gmode, guard=True, fused=True)
    ref, _ = oracle_lib.run(s.geom, s.angles, y, b, n_subsets=3, n_iters=4, beta=1e3, delta=5e-4,
                            potential=code, gamma_mode={"ray_length": 0, "row_sum": 1}[gmode])
    got = OSTR(s.geom, s.angles, y, b, cfg=cfg, device=dev).iterate(4).cpu().numpy()
    assert rel_rmse(got, ref) <= 1e-4


def test_tiny_and_edge_cases(dev, oracle_lib):
    # ragged detector/volume sizes that do not fill a work-group tile
    s = scan_tiny(nx=13, ny=7, nz=5, nbx=37, nby=9, n_views=3)
    from newrecon_amd.phantom import breast_phantom
    x = breast_phantom(s.geom, 3, (2, 3), 1)
    sysm = DBTSystem(s.geom, s.angles, device=dev)
    ref = oracle_lib.fproj(s.geom, s.angles, x, 0, 3)
    got = sysm.fproj(T(x, dev), 0, 3).cpu().numpy()
    assert rel_max(got, ref) <= 1e-6
    # empty view range is a no-op
    out = torch.full((1,) + s.geom.view_shape, 7.0, device=dev)
    sysm.fproj(T(x, dev), 1, 1, prj=out)
    assert float(out.min()) == 7.0
    # invalid ranges are reported, not executed
    with pytest.raises(OSTRError):
        sysm.fproj(T(x, dev), 2, 5)
    with pytest.raises(OSTRError):
        sysm.prior(T(x, dev), "lange", 0.0)
