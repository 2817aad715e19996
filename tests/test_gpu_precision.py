"""FP64 reductions on the GPU and the mixed-precision error budget.

* ostr_objective (FP64 data term + penalty) against the same quantities
  evaluated from the oracle's FP32 projections in FP64, and the oracle's FP64
  penalty; deterministic (bitwise equal on a second call).
* ostr_diff_norms against numpy FP64.
* BF16 / FP16 image and h storage (rays accumulated in FP32): relative error
  per iteration against the FP64 oracle OSTR on identical inputs, next to the
  FP32 path's own error, on the 200x100x20 / 9-view workload.
"""
import numpy as np
import pytest
import torch

from newrecon_amd.geometry import OSTRConfig
from newrecon_amd.ostr import OSTR, DBTSystem

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)


@pytest.mark.parametrize("potential,code", [("hyperbola", 2), ("lange", 1), ("q2", 0)])
def test_objective_matches_oracle(small_case, dev, potential, code, oracle_lib):
    s, x, y, b = small_case
    g = s.geom
    sysm = DBTSystem(g, s.angles, device=dev)
    data, pen = sysm.objective(T(x, dev), T(y, dev), T(b, dev), None, potential, 5e-4)
    l = oracle_lib.fproj(g, s.angles, x, 0, s.n_views).astype(np.float64)
    yb = b.astype(np.float64) * np.exp(-l)
    yy = y.astype(np.float64)
    ref_data = np.where(yy == 0, yb, yb - yy * np.log(np.where(yy == 0, 1, yb))).sum()
    ref_pen = oracle_lib.penalty(g, x.astype(np.float64), code, 5e-4)
    # sums of ~3e5 (data) and ~1e7 (penalty) FP64 terms in different orders
    assert data == pytest.approx(ref_data, rel=1e-11)
    assert pen == pytest.approx(ref_pen, rel=1e-10)
    # deterministic
    assert sysm.objective(T(x, dev), T(y, dev), T(b, dev), None, potential, 5e-4) == (data, pen)
    # and consistent with the all-FP64 oracle objective (FP64 rays)
    full = oracle_lib.objective(g, s.angles, x, y, b, None, code, 1e3, 5e-4)
    assert data + 1e3 * pen == pytest.approx(full, rel=1e-7)


def test_diff_norms(small_case, dev):
    s, x, _, _ = small_case
    rng = np.random.default_rng(3)
    a = (x + 1e-3 * rng.standard_normal(x.shape)).astype(np.float32)
    sysm = DBTSystem(s.geom, s.angles, device=dev)
    dd, bb = sysm.diff_norms(T(a, dev), T(x, dev))
    assert dd == pytest.approx(((a.astype(np.float64) - x) ** 2).sum(), rel=1e-12)
    assert bb == pytest.approx((x.astype(np.float64) ** 2).sum(), rel=1e-12)


@pytest.fixture(scope="module")
def fp64_reference(small_case, oracle_lib):
    s, x, y, b = small_case
    _, snaps = oracle_lib.run(s.geom, s.angles, y, b, n_subsets=3, n_iters=10, beta=1e3,
                              delta=5e-4, potential=2, gamma_mode=1, guard=True,
                              dtype=np.float64, snapshots=True)
    return snaps


@pytest.mark.parametrize("dtype,bound", [(torch.float32, 1e-4), (torch.float16, 5e-3),
                                         (torch.bfloat16, 3e-2)])
def test_mixed_precision_error_budget(small_case, dev, fp64_reference, oracle_lib, dtype,
                                      bound):
    """Per-iteration relative error against the FP64 oracle OSTR, image and h
    stored in `dtype`, rays and updates computed in FP32."""
    s, x, y, b = small_case
    cfg = OSTRConfig(n_subsets=3, n_iters=10, beta=1e3, delta=5e-4, fused=True)
    rec = OSTR(s.geom, s.angles, y, b, cfg=cfg, device=dev, img_dtype=dtype, prj_dtype=dtype)
    errs, objs = [], []

    def cb(it, xg):
        xf = xg.float().cpu().numpy().astype(np.float64)
        ref = fp64_reference[it - 1]
        errs.append(float(np.linalg.norm(xf - ref) / np.linalg.norm(ref)))

    rec.iterate(10, callback=cb)
    print(f"{dtype}: per-iteration rel error vs FP64 oracle: " +
          " ".join(f"{e:.2e}" for e in errs))
    assert max(errs) <= bound, errs
    # objective gap of the final image against the FP64 oracle's
    f_gpu = rec.objective()
    f_ref = oracle_lib.objective(s.geom, s.angles, fp64_reference[-1], y, b, None, 2, 1e3, 5e-4)
    gap = abs(f_gpu - f_ref) / abs(f_ref)
    print(f"{dtype}: objective {f_gpu:.10e} vs FP64 {f_ref:.10e} (rel gap {gap:.2e})")
    assert gap <= bound


@pytest.fixture(scope="module")
def fp64_accelerated(small_case, dev):
    """The FP64 accelerated SQS solver (oracle/converged.py) from the GPU's
    backprojected start: its iterates z_k for k = 1..30."""
    from oracle.converged import SQS, solve
    s, x, y, b = small_case
    cfg = OSTRConfig(n_subsets=1, beta=1e3, delta=5e-4, prior="hyperbola",
                     backprojector="matched", momentum=True)
    x0 = OSTR(s.geom, s.angles, y, b, cfg=cfg, x0="bp", device=dev).x.double().cpu().numpy()
    sqs = SQS(s.geom, s.angles, y, b, beta=1e3, delta=5e-4, nthreads=8)
    snaps = {}
    solve(sqs, x0, 30, each=lambda k, z: snaps.__setitem__(k, z.copy()))
    return x0, snaps


@pytest.mark.parametrize("img,h,bound", [
    (torch.float32, torch.float32, 1e-4), (torch.float32, torch.float16, 1e-3),
    (torch.float32, torch.bfloat16, 5e-3), (torch.float16, torch.float16, 3e-2),
    (torch.bfloat16, torch.bfloat16, 0.3)])
def test_mixed_precision_accelerated_matched(small_case, dev, fp64_accelerated, img, h, bound):
    """The warm-start iteration (one subset, exact transpose, momentum with the
    device restart test) with image and h stored in 16 bits, rays and updates
    in FP32, the FP64 reductions unchanged: relative error per iteration
    against the FP64 solver from the same start."""
    s, x, y, b = small_case
    x0, snaps = fp64_accelerated
    cfg = OSTRConfig(n_subsets=1, beta=1e3, delta=5e-4, prior="hyperbola",
                     backprojector="matched", momentum=True)
    rec = OSTR(s.geom, s.angles, y, b, cfg=cfg, x0=x0.astype(np.float32), device=dev,
               img_dtype=img, prj_dtype=h)
    errs = []

    def cb(it, xg):
        ref = snaps[it]
        xf = xg.float().cpu().numpy().astype(np.float64)
        errs.append(float(np.linalg.norm(xf - ref) / np.linalg.norm(ref)))

    rec.iterate(30, callback=cb)
    assert rec.h.dtype == h and rec.estimate.dtype == img
    print(f"accelerated matched, image {img}, h {h}: rel error vs FP64 solver "
          + " ".join(f"{e:.1e}" for e in errs[::3]), flush=True)
    assert max(errs) <= bound, errs


def test_fp64_denominator(small_case, dev, oracle_lib):
    """d summed over views in FP64 (SBP and exact transpose): equal bit for bit
    to the GPU's own per-view FP32 terms summed in double in view order and
    rounded once, and within the FP32 bar of the FP64 oracle's d."""
    s, x, y, b = small_case
    g = s.geom
    nv = s.n_views
    S = DBTSystem(g, s.angles, device=dev)
    yd = T(y, dev)
    # the per-view terms: gamma (row sum) * yry, row-major, one view per launch
    gc = S.gamma(0, nv, "row_sum")
    S.compute_gamma_yry(gc, S.compute_yry(torch.empty_like(yd), yd), gc)
    rows = S.to_rows(gc)
    acc = np.zeros(g.vol_shape, np.float64)
    for v in range(nv):
        acc = acc + S.bproj_sbp_rows(rows[v:v + 1], v, v + 1).cpu().numpy().astype(np.float64)
    d64 = S.precompute_d(yd, None, "row_sum", fp64=True).cpu().numpy()
    assert np.array_equal(d64, acc.astype(np.float32))
    gam = oracle_lib.gamma(g, s.angles, 0, nv, 1, dtype=np.float64)
    c = oracle_lib.yry(y.astype(np.float64), dtype=np.float64)
    ref = oracle_lib.bproj_sbp(g, s.angles, gam * c, 0, nv, dtype=np.float64)
    d32 = S.precompute_d(yd, None, "row_sum").cpu().numpy()
    peak = np.abs(ref).max()
    e32, e64 = np.abs(d32 - ref).max() / peak, np.abs(d64 - ref).max() / peak
    print(f"SBP d: FP32 sum {e32:.2e}, FP64 sum {e64:.2e} of the FP64 oracle's peak")
    assert e64 <= 1e-5
    # exact transpose, through the OSTR driver
    refm = oracle_lib.bproj_clip(g, s.angles, c * oracle_lib.fproj(
        g, s.angles, np.ones(g.vol_shape), 0, nv, dtype=np.float64), 0, nv, nthreads=8)
    for fp64 in (False, True):
        cfg = OSTRConfig(n_subsets=3, backprojector="matched", d_fp64=fp64)
        rec = OSTR(g, s.angles, y, b, cfg=cfg, device=dev)
        rec.precompute()
        got = rec.d.cpu().numpy()
        err = np.abs(got - refm).max() / np.abs(refm).max()
        print(f"matched d (fp64 sum {fp64}): {err:.2e}")
        assert err <= 1e-5
