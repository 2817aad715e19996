"""Pins the CPU oracle (oracle/ostr_oracle.cpp) against what the mathematics
fixes, independently of the oracle's own code:

* Siddon walker: analytic ray/box chord lengths, exact single-voxel segment
  lengths, uniform-volume integrals, dense-sampling brute force, linearity,
  mirror symmetry, nnz against per-voxel analytic intersections.
* gamma: |S - D| (ray length) and chord length (row sum).
* matched backprojector: the adjoint identity <Ax, y> = <x, A^T y> in FP64.
* SBP: constant-sinogram closed form and scipy's bilinear interpolation
  (map_coordinates, order 1) as the textbook check of the taps.
* priors: a numpy 27-neighbour loop, constant-image closed form, and RD equal
  to the finite-difference gradient of the penalty.
* elementwise ops and regroup: numpy formulas and the reference's printed
  regroup order (ostr.ipynb raw line 1274, tests/golden/reference_outputs.json).
* the OSTR driver: the noise-free fixed point, and FP32 vs FP64 agreement.
Each check is chosen so a dropped term, a wrong sign/index or a transposed
operand fails it.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy.ndimage import map_coordinates

from newrecon_amd.geometry import Geometry, arc_angles, regroup_index, scan_small

HERE = os.path.dirname(os.path.abspath(__file__))


def tiny_geom(nx=6, ny=5, nz=4, dx=2.0, dz=3.0, nbx=20, nby=14, db=3.0, src=200.0):
    return Geometry.half_cone(nx, ny, nz, dx, dx, dz, nbx, nby, db, db,
                              source_to_detector=src)


def box_chord(g, s, e):
    """Analytic length of segment s->e (FP64) inside the volume box."""
    s, e = np.asarray(s, float), np.asarray(e, float)
    d = e - s
    lo = np.array([g.x_p0, g.y_p0, g.z_p0])
    hi = lo + np.array([g.nx * g.dx, g.ny * g.dy, g.nz * g.dz])
    t0, t1 = -np.inf, np.inf
    for k in range(3):
        if d[k] == 0:
            if not (lo[k] < s[k] < hi[k]):
                return 0.0
            continue
        a, b = (lo[k] - s[k]) / d[k], (hi[k] - s[k]) / d[k]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    return max(t1 - t0, 0.0) * np.linalg.norm(d)


def voxel_chord(g, s, e, ix, iy, iz):
    s, e = np.asarray(s, float), np.asarray(e, float)
    d = e - s
    lo = np.array([g.x_p0 + ix * g.dx, g.y_p0 + iy * g.dy, g.z_p0 + iz * g.dz])
    hi = lo + np.array([g.dx, g.dy, g.dz])
    t0, t1 = -np.inf, np.inf
    for k in range(3):
        a, b = (lo[k] - s[k]) / d[k], (hi[k] - s[k]) / d[k]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    return max(t1 - t0, 0.0) * np.linalg.norm(d)


def random_rays(g, n, rng):
    """Rays from a source above the volume to points below it, crossing it."""
    out = []
    while len(out) < n:
        s = [rng.uniform(-40, 40), rng.uniform(-30, 5), rng.uniform(150, 200)]
        e = [rng.uniform(g.x_p0 - 2, -g.x_p0 + 2), rng.uniform(g.y_p0 - 2, 2),
             rng.uniform(-20, -1)]
        if box_chord(g, s, e) > 1e-3:
            out.append((s, e))
    return out


def test_chord_length_matches_slab_clip(oracle_lib):
    g = tiny_geom()
    rng = np.random.default_rng(1)
    for s, e in random_rays(g, 200, rng):
        val, hit, length = oracle_lib.walk(g, s, e)
        assert hit
        assert val == pytest.approx(box_chord(g, s, e), rel=1e-12, abs=1e-12)
        assert length == pytest.approx(np.linalg.norm(np.subtract(e, s)), rel=1e-15)


def test_single_voxel_segments_exact(oracle_lib):
    g = tiny_geom()
    rng = np.random.default_rng(2)
    rays = random_rays(g, 40, rng)
    for iz in range(g.nz):
        for iy in range(g.ny):
            for ix in range(g.nx):
                img = np.zeros(g.vol_shape)
                img[iz, iy, ix] = 1.0
                for s, e in rays[:8]:
                    val, _, _ = oracle_lib.walk(g, s, e, img=img)
                    assert val == pytest.approx(voxel_chord(g, s, e, ix, iy, iz), abs=1e-11)


def test_dense_sampling_brute_force(oracle_lib):
    g = tiny_geom()
    rng = np.random.default_rng(3)
    img = rng.random(g.vol_shape)
    for s, e in random_rays(g, 25, rng):
        s, e = np.asarray(s), np.asarray(e)
        n = 200_000
        t = (np.arange(n) + 0.5) / n
        p = s[None, :] + t[:, None] * (e - s)[None, :]
        ix = np.floor((p[:, 0] - g.x_p0) / g.dx).astype(int)
        iy = np.floor((p[:, 1] - g.y_p0) / g.dy).astype(int)
        iz = np.floor((p[:, 2] - g.z_p0) / g.dz).astype(int)
        ok = (ix >= 0) & (ix < g.nx) & (iy >= 0) & (iy < g.ny) & (iz >= 0) & (iz < g.nz)
        brute = img[iz[ok], iy[ok], ix[ok]].sum() * np.linalg.norm(e - s) / n
        val, _, _ = oracle_lib.walk(g, s, e, img=img)
        assert val == pytest.approx(brute, rel=2e-4)


def test_uniform_volume_projection_is_cos_times_chord(oracle_lib):
    g = tiny_geom()
    ang = arc_angles(3, 7.5)
    c = 0.37
    prj = oracle_lib.fproj(g, ang, np.full(g.vol_shape, c), 0, 3, dtype=np.float64)
    for a in range(3):
        th = float(ang[a])
        s = [g.source_r * math.sin(th), g.source_y, g.source_r * math.cos(th)]
        for i in range(0, g.nbx, 3):
            for j in range(0, g.nby, 2):
                e = [g.x_d0 + (i + 0.5) * g.dbx, g.y_d0 + (j + 0.5) * g.dby, -g.detector_r]
                ref = c * box_chord(g, s, e) * math.cos(th)
                assert prj[a, i, j] == pytest.approx(ref, rel=1e-12, abs=1e-12)


def test_projection_linear_and_mirror_symmetric(oracle_lib):
    g = tiny_geom(nx=8, nbx=24)
    ang = arc_angles(5, 7.5)  # symmetric: ang[k] = -ang[4-k]
    rng = np.random.default_rng(4)
    x, z = rng.random(g.vol_shape), rng.random(g.vol_shape)
    A = lambda v: oracle_lib.fproj(g, ang, v, 0, 5, dtype=np.float64)
    np.testing.assert_allclose(A(2.0 * x - 3.0 * z), 2.0 * A(x) - 3.0 * A(z), rtol=1e-11,
                               atol=1e-12)
    # mirroring the volume in x and negating the angle mirrors the detector in x
    px, pm = A(x), A(x[:, :, ::-1])
    np.testing.assert_allclose(pm[::-1, ::-1, :], px, rtol=1e-10, atol=1e-12)


def test_gamma_modes(oracle_lib):
    g = tiny_geom()
    ang = arc_angles(3, 7.5)
    rl = oracle_lib.gamma(g, ang, 0, 3, 0, dtype=np.float64)
    rs = oracle_lib.gamma(g, ang, 0, 3, 1, dtype=np.float64)
    for a in range(3):
        th = float(ang[a])
        s = np.array([g.source_r * math.sin(th), g.source_y, g.source_r * math.cos(th)])
        for i in range(g.nbx):
            for j in range(g.nby):
                e = np.array([g.x_d0 + (i + 0.5) * g.dbx, g.y_d0 + (j + 0.5) * g.dby,
                              -g.detector_r])
                chord = box_chord(g, s, e)
                if chord > 1e-9:
                    assert rl[a, i, j] == pytest.approx(np.linalg.norm(e - s), rel=1e-14)
                    assert rs[a, i, j] == pytest.approx(chord, rel=1e-11)
                elif chord == 0.0:
                    assert rl[a, i, j] == 0.0 and rs[a, i, j] == 0.0


def test_nnz_counts_positive_voxel_intersections(oracle_lib):
    g = tiny_geom()
    ang = arc_angles(3, 7.5)
    brute = 0
    for a in range(3):
        th = float(ang[a])
        s = [g.source_r * math.sin(th), g.source_y, g.source_r * math.cos(th)]
        for i in range(g.nbx):
            for j in range(g.nby):
                e = [g.x_d0 + (i + 0.5) * g.dbx, g.y_d0 + (j + 0.5) * g.dby, -g.detector_r]
                for iz in range(g.nz):
                    for iy in range(g.ny):
                        for ix in range(g.nx):
                            brute += voxel_chord(g, s, e, ix, iy, iz) > 1e-9
    assert oracle_lib.count_nnz(g, ang, 0, 3) == brute


def test_adjoint_identity_fp64(oracle_lib):
    g = tiny_geom(nx=7, ny=6, nz=5)
    ang = arc_angles(3, 7.5)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(g.vol_shape)
    y = rng.standard_normal((3,) + g.view_shape)
    Ax = oracle_lib.fproj(g, ang, x, 0, 3, dtype=np.float64)
    Aty = oracle_lib.bproj_matched(g, ang, y, 0, 3, dtype=np.float64)
    lhs, rhs = float((Ax * y).sum()), float((x * Aty).sum())
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # and A^T is not the SBP backprojector (the reference's is unmatched)
    B = oracle_lib.bproj_sbp(g, ang, y, 0, 3, dtype=np.float64)
    assert not np.allclose(B, Aty, rtol=1e-3)


def test_clip_transpose_matches_brute_force(oracle_lib):
    """Exact geometric transpose: every bin's ray clipped against every voxel
    (no candidate selection) in numpy FP64."""
    g = tiny_geom(nx=4, ny=3, nz=3, nbx=12, nby=9)
    ang = arc_angles(3, 7.5)
    rng = np.random.default_rng(12)
    y = rng.standard_normal((3,) + g.view_shape)
    got = oracle_lib.bproj_clip(g, ang, y, 0, 3)
    ref = np.zeros(g.vol_shape)
    f32 = np.float32
    for a in range(3):
        th = float(ang[a])
        # the forward projector's float rays: source rounded to float
        s = [float(f32(g.source_r) * f32(math.sin(th))), g.source_y,
             float(f32(g.source_r) * f32(math.cos(th)))]
        for i in range(g.nbx):
            for j in range(g.nby):
                e = [g.x_d0 + (i + 0.5) * g.dbx, g.y_d0 + (j + 0.5) * g.dby, -g.detector_r]
                for iz in range(g.nz):
                    for iy in range(g.ny):
                        for ix in range(g.nx):
                            ref[iz, iy, ix] += float(f32(math.cos(th))) * y[a, i, j] * \
                                voxel_chord(g, s, e, ix, iy, iz)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12)


def test_clip_transpose_vs_walk_transpose(small_case, oracle_lib):
    """The walk's relative stop test (Tomo_OSTR.cu:815) credits the last
    < 1e-6*max_l of a ray to the voxel before the exit, so the walk is the
    exact system matrix only up to ~1e-6 of a ray length per ray."""
    s = small_case[0]
    rng = np.random.default_rng(13)
    y = rng.standard_normal((s.n_views,) + s.geom.view_shape)
    walk = oracle_lib.bproj_matched(s.geom, s.angles, y, 0, s.n_views)
    clip = oracle_lib.bproj_clip(s.geom, s.angles, y, 0, s.n_views)
    diff = np.abs(walk - clip).max()
    assert diff > 0  # they are different operators ...
    # ... by at most ~1e-6 of a ray (660 mm) per ray through a voxel
    assert diff <= 1e-6 * 660 * np.abs(y).max() * 4 * s.n_views


def test_bp_init_constant_volume(oracle_lib):
    """BP.m's one-step image of the exact line integrals of a uniform volume
    is that uniform value wherever every ray through a voxel is unmasked."""
    # bins (0.7 mm) finer than voxels (2 mm): every voxel is crossed by rays
    g = tiny_geom(nx=6, ny=5, nz=4, nbx=60, nby=40, db=0.7)
    ang = arc_angles(3, 7.5)
    c = 0.0456
    p = oracle_lib.fproj(g, ang, np.full(g.vol_shape, c), 0, 3, dtype=np.float64)
    img = oracle_lib.bp_init(g, ang, p)
    # p comes from FP64 rays, the transpose uses the projector's float rays and
    # float cos: agreement to ~1e-8
    np.testing.assert_allclose(img[:, 1:-1, 1:-1], c, rtol=1e-6)
    # and it is linear in p
    img2 = oracle_lib.bp_init(g, ang, 2.5 * p)
    np.testing.assert_allclose(img2, 2.5 * img, rtol=1e-12, atol=1e-15)


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


def test_sbp_matches_textbook_bilinear(oracle_lib):
    g = tiny_geom(nbx=16, nby=10)
    ang = arc_angles(3, 7.5)
    rng = np.random.default_rng(6)
    prj = rng.standard_normal((3,) + g.view_shape)
    got = oracle_lib.bproj_sbp(g, ang, prj, 0, 3, dtype=np.float64)
    np.testing.assert_allclose(got, sbp_textbook(g, ang, prj), rtol=1e-10, atol=1e-12)


def test_sbp_constant_sinogram_closed_form(oracle_lib):
    g = tiny_geom(nbx=40, nby=30)
    ang = arc_angles(3, 7.5)
    got = oracle_lib.bproj_sbp(g, ang, np.full((3,) + g.view_shape, 2.5), 0, 3,
                               dtype=np.float64)
    # voxels whose four taps all fall on the detector (the half-cone detector
    # ends at the chest wall y = 0, so rows near y = 0 see partial taps)
    for iz, iy, ix in [(0, 0, 0), (3, 2, 5), (2, 1, 3)]:
        v = np.array([g.x_p0 + (ix + 0.5) * g.dx, g.y_p0 + (iy + 0.5) * g.dy,
                      g.z_p0 + (iz + 0.5) * g.dz])
        ref = 0.0
        for th in ang.astype(float):
            s = np.array([g.source_r * math.sin(th), g.source_y, g.source_r * math.cos(th)])
            ref += 2.5 * g.dz * (s @ s) / ((s - v) @ (s - v))
        assert got[iz, iy, ix] == pytest.approx(ref, rel=1e-12)


def prior_numpy(x, potential, delta, rho):
    nz, ny, nx = x.shape
    rdd, rd = np.zeros_like(x), np.zeros_like(x)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if dk == dj == di == 0:
                    continue
                w = 1.0 / math.sqrt(di * di + dj * dj + (dk * rho) ** 2)
                xp = np.pad(x, 1, mode="constant", constant_values=np.nan)
                nb = xp[1 + dk:1 + dk + nz, 1 + dj:1 + dj + ny, 1 + di:1 + di + nx]
                t = np.where(np.isnan(nb), 0.0, x - np.nan_to_num(nb))
                q = t / delta
                om = {0: np.ones_like(t), 1: 1 / (1 + np.abs(q)), 2: 1 / np.sqrt(1 + q * q)}[
                    potential]
                rdd += w * om
                rd += w * om * t
    return rdd, rd


@pytest.mark.parametrize("potential", [0, 1, 2])
def test_prior_matches_numpy_loop(oracle_lib, potential):
    g = tiny_geom()
    rng = np.random.default_rng(7)
    x = rng.random(g.vol_shape) * 1e-3
    rho = g.dz / g.dx
    rdd, rd = oracle_lib.prior(g, x, potential, 5e-4, rho=rho, dtype=np.float64)
    e_dd, e_d = prior_numpy(x, potential, 5e-4, rho)
    np.testing.assert_allclose(rdd, e_dd, rtol=1e-12)
    np.testing.assert_allclose(rd, e_d, rtol=1e-11, atol=1e-18)
    f_dd, f_d = oracle_lib.prior(g, x.astype(np.float32), potential, 5e-4, rho=rho)
    np.testing.assert_allclose(f_dd, e_dd, rtol=2e-6)
    np.testing.assert_allclose(f_d, e_d, rtol=2e-5, atol=1e-6 * np.abs(e_d).max())


def test_prior_constant_image_closed_form(oracle_lib):
    g = tiny_geom(dx=0.5, dz=1.0)
    rho = 2.0
    rdd, rd = oracle_lib.prior(g, np.full(g.vol_shape, 0.05), 2, 5e-4, rho=rho,
                               dtype=np.float64)
    total = (4 + 2 / rho + 4 / math.sqrt(2) + 8 / math.sqrt(1 + rho * rho)
             + 8 / math.sqrt(2 + rho * rho))
    np.testing.assert_allclose(rdd, total, rtol=1e-14)  # out-of-volume neighbours count
    assert np.abs(rd).max() == 0.0


@pytest.mark.parametrize("potential", [0, 1, 2])
def test_prior_gradient_is_penalty_gradient(oracle_lib, potential):
    g = tiny_geom(nx=4, ny=3, nz=3)
    rng = np.random.default_rng(8)
    x = rng.random(g.vol_shape) * 2e-3
    delta, rho = 5e-4, g.dz / g.dx
    _, rd = oracle_lib.prior(g, x, potential, delta, rho=rho, dtype=np.float64)
    h = 1e-7
    for idx in [(0, 0, 0), (1, 1, 2), (2, 2, 3)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = (oracle_lib.penalty(g, xp, potential, delta, rho)
              - oracle_lib.penalty(g, xm, potential, delta, rho)) / (2 * h)
        assert rd[idx] == pytest.approx(fd, rel=1e-5)


def test_elementwise_closed_forms(oracle_lib):
    rng = np.random.default_rng(9)
    y = rng.poisson(50, 1000).astype(np.float64)
    r = rng.random(1000) * 80
    c = oracle_lib.yry(y, r, dtype=np.float64)
    dif = np.where(y - r <= 0, y, y - r)
    np.testing.assert_array_equal(c, np.where(y == 0, 0, dif * dif / np.where(y == 0, 1, y)))
    b = np.full(1000, 1e4)
    l = rng.random(1000) * 5
    h = oracle_lib.compute_h(y, b, l, r, dtype=np.float64)
    yb = b * np.exp(-l)
    np.testing.assert_allclose(h, (y / (yb + r) - 1) * yb, rtol=1e-14)
    # h is the derivative of the negative log-likelihood (yb + r) - y log(yb + r) in l
    nll = lambda ll: (b * np.exp(-ll) + r) - y * np.log(b * np.exp(-ll) + r)
    np.testing.assert_allclose(h, (nll(l + 1e-6) - nll(l - 1e-6)) / 2e-6, rtol=1e-5,
                               atol=1e-6)
    x = rng.random(1000)
    capl, rd, rdd = rng.standard_normal(1000), rng.standard_normal(1000), rng.random(1000)
    d = rng.random(1000) + 0.1
    got = oracle_lib.update_est(x, capl, rd, d, rdd, 7.0, 3.0, dtype=np.float64)
    np.testing.assert_allclose(got, np.maximum(0, x - (3 * capl + 7 * rd) / (d + 14 * rdd)),
                               rtol=1e-14)


def test_regroup_known_answer_from_reference():
    fx = json.load(open(os.path.join(HERE, "golden", "reference_outputs.json")))
    order = regroup_index(fx["n_views"], fx["n_subsets"])
    assert order.tolist() == fx["regroup_order"]
    assert fx["all_b_size"] == [fx["n_views"] * fx["bins_x"] * fx["bins_y"]]


def test_regroup_data_follows_index(oracle_lib):
    y = np.arange(25 * 6, dtype=np.float32).reshape(25, 2, 3)
    out = oracle_lib.regroup(y, 5)
    idx = regroup_index(25, 5)
    np.testing.assert_array_equal(out, y[idx])


def test_update_guard_and_nan_quirk(oracle_lib):
    x = np.array([0.5, 0.5], np.float32)
    z = np.zeros(2, np.float32)
    assert np.isnan(oracle_lib.update_est(x, z, z, z, z, 0.0, 1.0, guard=False)).all()
    np.testing.assert_array_equal(oracle_lib.update_est(x, z, z, z, z, 0.0, 1.0, guard=True), x)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noise_free_fixed_point(oracle_lib, dtype):
    """y = b exp(-cos A x_true) exactly => h = 0 => x_true is a fixed point
    (beta = 0, guard for voxels no ray sees)."""
    s = scan_small()
    g = s.geom
    from newrecon_amd.phantom import breast_phantom
    x = breast_phantom(g, 11, (3, 4), 2).astype(dtype)
    l = oracle_lib.fproj(g, s.angles, x, 0, 9, dtype=dtype)
    b = np.full_like(l, 1e4)
    # y = b e^{-l} with the oracle's own exp: compute_h(0, b, l) = -(b e^{-l})
    y = -oracle_lib.compute_h(np.zeros_like(l), b, l, dtype=dtype)
    out, _ = oracle_lib.run(g, s.angles, y, b, x0=x, n_subsets=3, n_iters=2, beta=0.0,
                            delta=5e-4, potential=2, gamma_mode=1, guard=True, dtype=dtype)
    np.testing.assert_array_equal(out, x)


def test_fp32_oracle_close_to_fp64(small_case, oracle_lib):
    s, x, y, b = small_case
    a, _ = oracle_lib.run(s.geom, s.angles, y, b, n_subsets=3, n_iters=10)
    d, _ = oracle_lib.run(s.geom, s.angles, y, b, n_subsets=3, n_iters=10, dtype=np.float64)
    rel = np.linalg.norm(a - d) / np.linalg.norm(d)
    assert rel < 1e-4, rel


def test_invalid_configs_rejected(oracle_lib):
    s = scan_small()
    y = np.ones((9,) + s.geom.view_shape, np.float32)
    with pytest.raises(ValueError):
        oracle_lib.run(s.geom, s.angles, y, y, n_subsets=2)  # 9 % 2 != 0
    with pytest.raises(ValueError):
        oracle_lib.run(s.geom, s.angles, y, y, n_subsets=3, delta=0.0, potential=1)
    with pytest.raises(ValueError):
        oracle_lib.prior(s.geom, np.zeros(s.geom.vol_shape, np.float32), 2, 0.0)


def _tiny_case(oracle_lib, seed=5):
    from newrecon_amd.geometry import scan_tiny
    from newrecon_amd.phantom import breast_phantom, poisson_counts
    s = scan_tiny()
    x = breast_phantom(s.geom, 3, (2, 3), 1).astype(np.float64)
    l = oracle_lib.fproj(s.geom, s.angles, x, 0, s.n_views, dtype=np.float64)
    y = poisson_counts(l, s.i0, seed).astype(np.float64)
    return s, x, y, np.full_like(y, s.i0)


def test_matched_update_direction_is_objective_gradient(oracle_lib):
    """With the exact transpose, G^T h + beta RD is the gradient of the FP64
    objective (data term + beta * penalty): directional finite differences
    along random directions. An SBP-style or uncos-weighted transpose, or a
    wrong sign of h, fails this."""
    s, x, y, b = _tiny_case(oracle_lib)
    g, beta, delta = s.geom, 1e3, 5e-4
    x = x + 0.002  # keep every voxel off the x >= 0 bound
    l = oracle_lib.fproj(g, s.angles, x, 0, s.n_views, dtype=np.float64)
    h = oracle_lib.compute_h(y, b, l, dtype=np.float64)
    grad = oracle_lib.bproj_clip(g, s.angles, h, 0, s.n_views, cos_weight=True)
    _, rd = oracle_lib.prior(g, x, 2, delta, dtype=np.float64)
    grad = grad + beta * rd
    rng = np.random.default_rng(2)
    for _ in range(3):
        v = rng.standard_normal(x.shape)
        eps = 1e-6
        fp = oracle_lib.objective(g, s.angles, x + eps * v, y, b, potential=2, beta=beta,
                                  delta=delta)
        fm = oracle_lib.objective(g, s.angles, x - eps * v, y, b, potential=2, beta=beta,
                                  delta=delta)
        fd = (fp - fm) / (2 * eps)
        an = float((grad * v).sum())
        assert abs(fd - an) <= 1e-5 * abs(an) + 1e-3, (fd, an)


def test_matched_sqs_decreases_objective(oracle_lib):
    """M = 1 with the exact transpose is a separable-quadratic-surrogate
    descent: the FP64 objective never increases."""
    s, x, y, b = _tiny_case(oracle_lib, seed=6)
    _, snaps = oracle_lib.run(s.geom, s.angles, y, b, n_subsets=1, n_iters=12, beta=1e3,
                              delta=5e-4, potential=2, dtype=np.float64, snapshots=True,
                              backprojector=1)
    objs = [oracle_lib.objective(s.geom, s.angles, z, y, b, potential=2, beta=1e3, delta=5e-4)
            for z in snaps]
    assert all(o2 <= o1 + 1e-12 * abs(o1) for o1, o2 in zip(objs, objs[1:])), objs
    assert objs[-1] < objs[0]
