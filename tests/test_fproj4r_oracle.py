"""Pins the oracle's 2x2 supersampled forward projection (oracle/ostr_oracle.cpp,
project_4r; Tomo_OSTR.cu:1453-1550) independently of its code:

* sub-bin geometry: every sub-bin value is the single-ray walk from the source
  to the analytic sub-bin centre x_d0 + (i + 0.5) Bsize / 2 (exact equality);
* uniform volume: each sub-bin value is the analytic chord length of its ray
  through the volume box, misses are 0;
* merge: each bin is the mean of its 2x2 block, summed in the stated order
  (exact), in the detector-row-major layout.
"""
import math

import numpy as np

from newrecon_amd.geometry import Geometry, arc_angles


def tiny_geom(nx=6, ny=5, nz=4, dx=2.0, dz=3.0, nbx=9, nby=7, db=3.0, src=200.0):
    return Geometry.half_cone(nx, ny, nz, dx, dx, dz, nbx, nby, db, db,
                              source_to_detector=src)


def box_chord(g, s, e):
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


def source(g, th):
    return np.array([g.source_r * math.sin(th), g.source_y, g.source_r * math.cos(th)])


def sub_centre(g, ii, jj):
    return np.array([g.x_d0 + (ii + 0.5) * g.dbx / 2, g.y_d0 + (jj + 0.5) * g.dby / 2,
                     -g.detector_r])


def test_sub_bins_are_walks_to_sub_bin_centres(oracle_lib):
    g = tiny_geom()
    ang = arc_angles(3, 7.5)
    rng = np.random.default_rng(21)
    x = rng.random(g.vol_shape)
    _, prj4 = oracle_lib.fproj_4r(g, ang, x, 0, 3, dtype=np.float64)
    for a in range(3):
        th = float(ang[a])
        for jj in range(0, 2 * g.nby, 3):
            for ii in range(0, 2 * g.nbx, 2):
                s, hit, _ = oracle_lib.walk(g, source(g, th), sub_centre(g, ii, jj), img=x)
                assert prj4[a, jj, ii] == (s if hit else 0.0)


def test_uniform_volume_sub_bins_are_chords(oracle_lib):
    g = tiny_geom()
    ang = arc_angles(3, 7.5)
    _, prj4 = oracle_lib.fproj_4r(g, ang, np.ones(g.vol_shape), 0, 3, dtype=np.float64)
    n_hit = 0
    for a in range(3):
        S = source(g, float(ang[a]))
        for jj in range(2 * g.nby):
            for ii in range(2 * g.nbx):
                want = box_chord(g, S, sub_centre(g, ii, jj))
                n_hit += want > 0
                assert abs(prj4[a, jj, ii] - want) <= 1e-9 * max(want, 1.0)
    assert 0 < n_hit < prj4.size  # hits and misses both occur


def test_merge_is_ordered_2x2_mean(oracle_lib):
    g = tiny_geom()
    ang = arc_angles(4, 7.5)
    rng = np.random.default_rng(22)
    x = rng.random(g.vol_shape).astype(np.float32)
    prj, prj4 = oracle_lib.fproj_4r(g, ang, x, 1, 4)
    assert prj.shape == (3, g.nby, g.nbx) and prj4.shape == (3, 2 * g.nby, 2 * g.nbx)
    want = (((prj4[:, 0::2, 0::2] + prj4[:, 0::2, 1::2]) + prj4[:, 1::2, 0::2])
            + prj4[:, 1::2, 1::2]) / np.float32(4)
    np.testing.assert_array_equal(prj, want)
    # and the slots follow the angle range [1, 4)
    p64, _ = oracle_lib.fproj_4r(g, ang[1:], x.astype(np.float64), 0, 3, dtype=np.float64)
    np.testing.assert_allclose(prj, p64, rtol=1e-5, atol=1e-6)
