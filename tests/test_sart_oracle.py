"""Pins the oracle's OS-SART restatement (oracle/ostr_oracle.cpp, OS-SART
section) against what the mathematics fixes, independently of its code:

* SART_prj_diff: the numpy formula, with the norm == 0 bins set to 0.
* the SART backprojection update: SBP(diff) and the sensitivity are scipy's
  bilinear interpolation (map_coordinates, order 1, zero outside the detector)
  of diff and of an all-ones sinogram, so the update is checked against
  max(0, x + lam * (textbook(diff) [+ beta prior]) / textbook(1)); a constant
  diff c moves every voxel with a sensitivity by exactly lam c.
* the Huber prior: a numpy 27-neighbour loop, and the finite-difference
  gradient of its potential sum_pairs w delta^2 (sqrt(1 + (t/delta)^2) - 1).
* the driver: the noise-free fixed point, FP32 vs FP64 agreement, and the data
  residual falling over the first passes from a zero image.
"""
import math

import numpy as np
import pytest
from scipy.ndimage import map_coordinates

from newrecon_amd.geometry import GAMMA_MODES, Geometry, arc_angles


def tiny_geom(nx=6, ny=5, nz=4, dx=2.0, dz=3.0, nbx=20, nby=14, db=3.0, src=200.0):
    return Geometry.half_cone(nx, ny, nz, dx, dx, dz, nbx, nby, db, db,
                              source_to_detector=src)


def sbp_textbook(g, ang, prj):
    """SBP via scipy bilinear interpolation at the projected voxel centres."""
    out = np.zeros(g.vol_shape)
    zc = g.z_p0 + (np.arange(g.nz) + 0.5) * g.dz
    yc = g.y_p0 + (np.arange(g.ny) + 0.5) * g.dy
    xc = g.x_p0 + (np.arange(g.nx) + 0.5) * g.dx
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    for a in range(len(ang)):
        th = float(ang[a])
        s = np.array([g.source_r * math.sin(th), g.source_y, g.source_r * math.cos(th)])
        t = (-g.detector_r - s[2]) / (Z - s[2])
        x2 = s[0] + (X - s[0]) * t
        y2 = s[1] + (Y - s[1]) * t
        ci = (x2 - g.x_d0) / g.dbx - 0.5
        cj = (y2 - g.y_d0) / g.dby - 0.5
        val = map_coordinates(prj[a], [ci.ravel(), cj.ravel()], order=1, mode="grid-constant",
                              cval=0.0).reshape(X.shape)
        dw = (s @ s) / ((s[0] - X) ** 2 + (s[1] - Y) ** 2 + (s[2] - Z) ** 2)
        out += val * g.dz * dw
    return out


def test_prj_diff_formula(oracle_lib):
    rng = np.random.default_rng(1)
    p, est = rng.random(500), rng.random(500)
    norm = rng.random(500) + 0.5
    norm[::7] = 0.0
    want = np.where(norm != 0, (p - est) / np.where(norm != 0, norm, 1.0), 0.0)
    np.testing.assert_allclose(oracle_lib.sart_prj_diff(p, est, norm, dtype=np.float64), want,
                               rtol=1e-15)
    got32 = oracle_lib.sart_prj_diff(p, est, norm)
    assert np.all(got32[::7] == 0)
    np.testing.assert_allclose(got32, want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("with_prior", [False, True])
def test_update_matches_textbook_bilinear(oracle_lib, with_prior):
    # a detector narrower than the volume's shadow, so some taps are masked
    g = tiny_geom(dx=4.0, nbx=8, nby=5)
    ang = arc_angles(3, 7.5)
    rng = np.random.default_rng(2)
    diff = rng.standard_normal((3,) + g.view_shape)
    x = rng.random(g.vol_shape) * 0.05
    prior = rng.standard_normal(g.vol_shape) if with_prior else None
    lam, beta = 0.7, (0.3 if with_prior else 0.0)
    got = oracle_lib.bproj_sart(g, ang, diff, 0, 3, x, lam, prior=prior, beta=beta,
                                dtype=np.float64)
    num = sbp_textbook(g, ang, diff)
    sens = sbp_textbook(g, ang, np.ones_like(diff))
    assert (sens == 0).any() and (sens > 0).any()
    if with_prior:
        num = num + beta * prior
    u = np.where(sens != 0, num / np.where(sens != 0, sens, 1.0), 0.0)
    want = np.maximum(0.0, x + lam * u)
    assert (want == 0).any()  # the clamp is exercised
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12)


def test_constant_diff_moves_by_lambda_c(oracle_lib):
    g = tiny_geom(dx=4.0, nbx=8, nby=5)
    ang = arc_angles(5, 6.0)
    x = np.full(g.vol_shape, 0.5)
    got = oracle_lib.bproj_sart(g, ang, np.full((5,) + g.view_shape, 0.04), 0, 5, x, 2.5,
                                dtype=np.float64)
    sens = sbp_textbook(g, ang, np.ones((5,) + g.view_shape))
    np.testing.assert_allclose(got[sens > 0], 0.6, rtol=1e-13)
    np.testing.assert_array_equal(got[sens == 0], 0.5)


def huber_numpy(x, delta):
    """The reference's weights are 1 / (float32 distance): sqrt(2) and sqrt(3)
    rounded to float32 (Tomo_OSTR.cu:301)."""
    nz, ny, nx = x.shape
    out = np.zeros_like(x)
    xp = np.pad(x, 1, mode="constant", constant_values=np.nan)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                dist = float(np.float32(math.sqrt(di * di + dj * dj + dk * dk))) or 1.0
                nb = xp[1 + dk:1 + dk + nz, 1 + dj:1 + dj + ny, 1 + di:1 + di + nx]
                t = np.where(np.isnan(nb), 0.0, x - np.nan_to_num(nb))
                out += (t / dist) / np.sqrt(1.0 + (t / delta) ** 2)
    return out


def test_huber_prior_matches_numpy_loop(oracle_lib):
    g = tiny_geom()
    rng = np.random.default_rng(3)
    x = rng.random(g.vol_shape) * 1e-3
    start = rng.standard_normal(g.vol_shape)
    got = oracle_lib.prior_huber(g, x, 5e-4, prior=start, dtype=np.float64)
    np.testing.assert_allclose(got, start + huber_numpy(x, 5e-4), rtol=1e-12, atol=1e-15)
    got32 = oracle_lib.prior_huber(g, x.astype(np.float32), 5e-4)
    ref = huber_numpy(x.astype(np.float32).astype(np.float64), 5e-4)
    np.testing.assert_allclose(got32, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())


def test_huber_prior_is_potential_gradient(oracle_lib):
    # sum over voxels and in-volume neighbours of w delta^2 (sqrt(1+(t/delta)^2) - 1)
    # counts each pair twice, so its gradient is 2 * prior.
    g = tiny_geom(nx=4, ny=3, nz=3)
    rng = np.random.default_rng(4)
    x = rng.random(g.vol_shape) * 2e-3
    delta = 1e-3

    def phi(v):
        nz, ny, nx = v.shape
        vp = np.pad(v, 1, mode="constant", constant_values=np.nan)
        tot = 0.0
        for dk in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    dist = math.sqrt(di * di + dj * dj + dk * dk) or 1.0
                    nb = vp[1 + dk:1 + dk + nz, 1 + dj:1 + dj + ny, 1 + di:1 + di + nx]
                    t = np.where(np.isnan(nb), 0.0, v - np.nan_to_num(nb))
                    tot += np.sum(delta ** 2 * (np.sqrt(1 + (t / delta) ** 2) - 1) / dist)
        return tot

    got = oracle_lib.prior_huber(g, x, delta, dtype=np.float64)
    h = 1e-8
    for idx in [(0, 0, 0), (1, 1, 2), (2, 2, 3), (1, 0, 1)]:
        e = np.zeros_like(x)
        e[idx] = h
        fd = (phi(x + e) - phi(x - e)) / (2 * h)
        assert got[idx] == pytest.approx(fd / 2, rel=1e-6, abs=1e-9)


def test_huber_delta_zero_rejected(oracle_lib):
    g = tiny_geom()
    with pytest.raises(ValueError):
        oracle_lib.prior_huber(g, np.zeros(g.vol_shape), 0.0)


@pytest.mark.parametrize("cos_weight", [False, True])
def test_driver_noise_free_fixed_point(oracle_lib, cos_weight):
    g = tiny_geom(nbx=40, nby=30)
    ang = arc_angles(6, 7.5)
    rng = np.random.default_rng(5)
    x = rng.random(g.vol_shape) * 0.05 + 0.01
    p = oracle_lib.fproj(g, ang, x, 0, 6, cos_weight=cos_weight, dtype=np.float64)
    got, _ = oracle_lib.run_sart(g, ang, p, x0=x, n_subsets=3, n_iters=2,
                                 cos_weight=cos_weight, dtype=np.float64)
    np.testing.assert_allclose(got, x, rtol=1e-12, atol=1e-15)


def test_driver_residual_falls_and_fp32_tracks_fp64(oracle_lib):
    g = tiny_geom(nbx=40, nby=30)
    ang = arc_angles(6, 7.5)
    rng = np.random.default_rng(6)
    x = rng.random(g.vol_shape) * 0.05 + 0.01
    p = oracle_lib.fproj(g, ang, x, 0, 6, cos_weight=False, dtype=np.float64)
    _, snaps = oracle_lib.run_sart(g, ang, p, n_subsets=2, n_iters=4, lam=0.8,
                                   dtype=np.float64, snapshots=True)
    res = [np.linalg.norm(oracle_lib.fproj(g, ang, s, 0, 6, cos_weight=False,
                                           dtype=np.float64) - p) for s in snaps]
    assert all(b < a for a, b in zip(res, res[1:]))
    assert res[-1] < 0.5 * np.linalg.norm(p)
    x32, _ = oracle_lib.run_sart(g, ang, p, n_subsets=2, n_iters=4, lam=0.8)
    np.testing.assert_allclose(x32, snaps[-1], rtol=1e-4, atol=1e-6)


def test_driver_matches_launcher_sequence(oracle_lib):
    """The driver's subset order and buffers: slot m*Vs+j holds view j*M+m."""
    from newrecon_amd.geometry import regroup_index
    g = tiny_geom(nbx=40, nby=30)
    ang = arc_angles(6, 7.5)
    rng = np.random.default_rng(7)
    p = rng.random((6,) + g.view_shape) * 0.3
    M, vs = 3, 2
    idx = regroup_index(6, M)
    x = np.zeros(g.vol_shape)
    for _ in range(2):
        for m in range(M):
            sl = idx[m * vs:(m + 1) * vs]
            est = oracle_lib.fproj(g, ang, x, m * vs, (m + 1) * vs, index=idx,
                                   cos_weight=True, dtype=np.float64)
            norm = oracle_lib.gamma(g, ang, m * vs, (m + 1) * vs, GAMMA_MODES["row_sum"],
                                    index=idx, dtype=np.float64)
            diff = oracle_lib.sart_prj_diff(p[sl], est, norm, dtype=np.float64)
            pri = oracle_lib.prior_huber(g, x, 2e-3, dtype=np.float64)
            x = oracle_lib.bproj_sart(g, ang, diff, m * vs, (m + 1) * vs, x, 0.9, index=idx,
                                      prior=pri, beta=-0.01, dtype=np.float64)
    got, _ = oracle_lib.run_sart(g, ang, p, n_subsets=M, n_iters=2, lam=0.9, beta=-0.01,
                                 delta=2e-3, cos_weight=True, dtype=np.float64)
    np.testing.assert_array_equal(got, x)


def test_driver_rejects_bad_configs(oracle_lib):
    g = tiny_geom()
    ang = arc_angles(6, 7.5)
    p = np.zeros((6,) + g.view_shape)
    with pytest.raises(ValueError):
        oracle_lib.run_sart(g, ang, p, n_subsets=4)
    with pytest.raises(ValueError):
        oracle_lib.run_sart(g, ang, p, n_subsets=3, beta=1.0, delta=0.0)
