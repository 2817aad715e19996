"""Parity at the clinical size: 1400x600x60 voxels of 0.1x0.1x1 mm, 15 views of
2048x1664 bins, 5 subsets, hyperbola prior (the workload bench.py times).

The oracle (C++/OpenMP) takes about a second per subset step at this size on
16 threads, so these checks compare full-size arrays on the same seeded
phantom and counts, to the bars of the small-case tests:
* the forward of subset 0 through the index table (fproj, cos-weighted)
  bit-identical to the oracle's walk, and the in-volume segment count of one
  view exact;
* d = SBP(gamma . yry) (row-sum gamma) against the oracle, <= 1e-5 of the peak;
* one fused OSTR iteration (5 subsets of fproj_h + update_fused) against the
  oracle's launch-by-launch subset steps from the same start and the same d,
  <= 1e-4 relative RMSE (the parity bar of the path), and the change the
  iteration makes to within 1e-3 of the oracle's;
* one fused OS-SART pass (cos-weighted, Huber prior) against the oracle's
  OS-SART loop on the reference's SART launchers, <= 1e-4 relative RMSE;
* <A x, y> = <x, A^T y> for the forward projector and the exact transpose over
  one full view (FP64 sums of the FP32 results);
* the 2x2 supersampled forward of one view (sub-bins and merged bins) against
  the oracle, <= 1e-6 of the peak with the same zero pattern;
* HIP-graph replay of fused iterations equal to the eager loop bit for bit.
"""
import os

import numpy as np
import pytest
import torch

from newrecon_amd.geometry import regroup_index, scan_clinical
from newrecon_amd.ostr import OSTR, DBTSystem

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def rel_max(a, ref):
    return float(np.abs(np.asarray(a, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))


def rel_rmse(a, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - ref) / max(np.linalg.norm(ref), 1e-30))


@pytest.fixture(scope="module")
def clinical(dev, oracle_lib):
    """Seeded phantom, its noisy counts (oracle forward of all 15 views, Poisson
    from the scan's noise seed) and the blank scan."""
    from newrecon_amd.phantom import breast_phantom, poisson_counts
    s = scan_clinical()
    x = breast_phantom(s.geom, s.phantom_seed, s.n_fg, s.n_calc)
    print(f"[clinical] phantom {x.shape}", flush=True)
    l = oracle_lib.fproj(s.geom, s.angles, x, 0, s.n_views, nthreads=THREADS)
    y = poisson_counts(l, s.i0, s.noise_seed)
    del l
    b = np.full_like(y, s.i0)
    print("[clinical] counts", flush=True)
    return s, x, y, b


def test_clinical_forward_matches_oracle(clinical, dev, oracle_lib):
    s, x, _, _ = clinical
    idx = regroup_index(s.n_views, s.recon.n_subsets)
    vs = s.n_views // s.recon.n_subsets
    sysm = DBTSystem(s.geom, s.angles, index=idx, device=dev)
    got = sysm.fproj(torch.from_numpy(x).to(dev), 0, vs).cpu().numpy()
    ref = oracle_lib.fproj(s.geom, s.angles, x, 0, vs, index=idx, nthreads=THREADS)
    err = rel_max(got, ref)
    print(f"[clinical] forward of subset 0: max error / peak {err:.3g}, "
          f"{int((got != ref).sum())} of {ref.size} bins differ", flush=True)
    # the walk follows the oracle's operations one for one: bit-identical
    assert np.array_equal(got, ref), err
    assert sysm.count_nnz(0, 1) == oracle_lib.count_nnz(s.geom, s.angles, 0, 1, index=idx,
                                                        nthreads=THREADS)


def test_clinical_ostr_iteration_matches_oracle(clinical, dev, oracle_lib):
    s, x, y, b = clinical
    g, cfg = s.geom, s.recon
    cfg.fused = True
    M = cfg.n_subsets
    vs = s.n_views // M
    x0 = (0.5 * x).astype(np.float32)
    # GPU: d, then one fused iteration (fproj_h + update_fused per subset)
    rec = OSTR(g, s.angles, y, b, cfg=cfg, x0=x0, device=dev)
    rec.precompute()
    d_gpu = rec.d.cpu().numpy()
    x1 = rec.iterate(1).cpu().numpy()
    del rec
    torch.cuda.empty_cache()
    # oracle: d = SBP(gamma . yry), natural view order, row-sum gamma
    gam = oracle_lib.gamma(g, s.angles, 0, s.n_views, 1, nthreads=THREADS)
    d_ref = oracle_lib.bproj_sbp(g, s.angles, oracle_lib.yry(y) * gam, 0, s.n_views,
                                 nthreads=THREADS)
    del gam
    err_d = rel_max(d_gpu, d_ref)
    print(f"[clinical] d: max error / peak {err_d:.3g}", flush=True)
    assert err_d <= 1e-5, err_d
    # oracle: the subset steps of the reference driver, from the same start
    idx = regroup_index(s.n_views, M)
    xr = x0
    for m in range(M):
        a0, a1 = m * vs, (m + 1) * vs
        views = idx[a0:a1]
        l = oracle_lib.fproj(g, s.angles, xr, a0, a1, index=idx, nthreads=THREADS)
        h = oracle_lib.compute_h(np.ascontiguousarray(y[views]), np.ascontiguousarray(b[views]),
                                 l)
        capl = oracle_lib.bproj_sbp(g, s.angles, h, a0, a1, index=idx, nthreads=THREADS)
        rdd, rd = oracle_lib.prior(g, xr, 2, cfg.delta, nthreads=THREADS)
        xr = oracle_lib.update_est(xr, capl, rd, d_ref, rdd, cfg.beta, float(M),
                                   guard=cfg.guard)
        print(f"[clinical] oracle subset {m}", flush=True)
    err = rel_rmse(x1, xr)
    err_step = rel_rmse(x1 - x0, xr - x0)
    print(f"[clinical] one iteration: rel RMSE {err:.3g}, of the change {err_step:.3g}",
          flush=True)
    assert err <= 1e-4, err
    assert err_step <= 1e-3, err_step


def test_clinical_matched_is_the_adjoint(clinical, dev):
    s, x, _, _ = clinical
    sysm = DBTSystem(s.geom, s.angles, device=dev)
    rng = np.random.default_rng(17)
    yv = rng.standard_normal((1,) + s.geom.view_shape).astype(np.float32)
    xt = torch.from_numpy(x).to(dev)
    yt = torch.from_numpy(yv).to(dev)
    v = 7  # the central view
    Ax = sysm.fproj(xt, v, v + 1).double()
    Aty = sysm.bproj_matched(yt, v, v + 1).double()
    lhs = float((Ax * yt.double()).sum())
    rhs = float((xt.double() * Aty).sum())
    scale = float((Ax.abs() * yt.double().abs()).sum())
    print(f"[clinical] adjoint: |lhs - rhs| / sum|Ax||y| = {abs(lhs - rhs) / scale:.3g}",
          flush=True)
    assert abs(lhs - rhs) <= 1e-6 * scale, (lhs, rhs, scale)


def test_clinical_os_sart_pass_matches_oracle(clinical, dev, oracle_lib):
    from newrecon_amd.sart import OSSART
    s, x, y, b = clinical
    p = np.log(b / np.maximum(y, 1.0)).astype(np.float32)
    kw = dict(n_subsets=s.recon.n_subsets, lam=1.0, beta=-2e-3, delta=1e-3, cos_weight=True)
    ref, _ = oracle_lib.run_sart(s.geom, s.angles, p, n_iters=1, nthreads=THREADS, **kw)
    rec = OSSART(s.geom, s.angles, p, device=dev, **kw)
    got = rec.iterate(1).cpu().numpy()
    err = rel_rmse(got, ref)
    print(f"[clinical] OS-SART pass: rel RMSE {err:.3g}", flush=True)
    assert err <= 1e-4, err


def test_clinical_fproj_4r_matches_oracle(clinical, dev, oracle_lib):
    s, x, _, _ = clinical
    sysm = DBTSystem(s.geom, s.angles, device=dev)
    v = 3
    ref, ref4 = oracle_lib.fproj_4r(s.geom, s.angles, x, v, v + 1, nthreads=THREADS)
    prj, prj4 = sysm.fproj_4r(torch.from_numpy(x).to(dev), v, v + 1, keep_4r=True)
    prj, prj4 = prj.cpu().numpy(), prj4.cpu().numpy()
    e4, e = rel_max(prj4, ref4), rel_max(prj, ref)
    print(f"[clinical] 4R forward of one view: sub-bins {e4:.3g}, merged {e:.3g}", flush=True)
    assert e4 <= 1e-6, e4
    assert e <= 1e-6, e
    assert np.array_equal(prj4 == 0, ref4 == 0)


def test_clinical_graph_replay_equals_eager(clinical, dev):
    """Five clinical iterations (5 subsets each, fused) replayed from a HIP
    graph equal the eager loop bit for bit."""
    s, x, y, b = clinical
    cfg = s.recon
    cfg.fused = True
    x0 = (0.5 * x).astype(np.float32)
    out = []
    for graph in (False, True):
        rec = OSTR(s.geom, s.angles, y, b, cfg=cfg, x0=x0, device=dev)
        out.append(rec.iterate(5, graph=graph).clone())
        del rec
    torch.cuda.synchronize()
    assert torch.equal(out[0], out[1])
