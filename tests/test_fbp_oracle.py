"""Pins the DBT FBP pieces by closed forms (newrecon_amd/fbp.py taps,
oracle/fbp.py restatement):

* 'ramp' taps are tau times the discrete band-limited ramp of Kak & Slaney
  (1/(4 tau^2) at 0, -1/(pi k tau)^2 at odd k, 0 at even k);
* a Hann window reaching 0 at Nyquist is, in space, the 3-tap kernel
  (1/4, 1/2, 1/4): the 'hann' taps equal 0.5 h[k] + 0.25 (h[k-1] + h[k+1]);
* the ramp has no DC response: sum of taps -> 0 as the rows grow;
* the cosine weight is 1 under the source and D/|S - E| elsewhere;
* the restated filter equals a brute-force double loop.
"""
import math

import numpy as np

from newrecon_amd.fbp import filter_taps, ramlak
from newrecon_amd.geometry import Geometry, arc_angles
from oracle import fbp as F


def test_ramp_taps_closed_form():
    n, tau = 9, 0.14
    t = filter_taps(n, tau, "ramp")
    k = np.arange(-(n - 1), n)
    want = np.where(k == 0, 1 / (4 * tau * tau),
                    np.where(k % 2 != 0, -1 / (math.pi * np.maximum(np.abs(k), 1) * tau) ** 2, 0.0))
    np.testing.assert_allclose(t, want * tau, rtol=1e-15)


def test_hann_is_three_tap_smoothing_of_ramp():
    n, tau = 64, 0.5
    h = ramlak(n, tau) * tau
    t = filter_taps(n, tau, "hann")
    want = 0.5 * h[1:-1] + 0.25 * (h[:-2] + h[2:])
    np.testing.assert_allclose(t[1:-1], want, rtol=1e-9, atol=1e-12 * np.abs(h).max())


def test_ramp_dc_response_vanishes():
    sums = [abs(filter_taps(n, 1.0, "ramp").sum()) for n in (16, 64, 256)]
    assert sums[0] > sums[1] > sums[2] and sums[2] < 2e-3


def test_cos_weights():
    g = Geometry.half_cone(8, 6, 4, 1.0, 1.0, 1.0, 31, 9, 1.0, 1.0, source_to_detector=100.0,
                           source_y=-4.5)
    w = F.cos_weights(g, 0.0)
    # bin 15 along x is centred at x = 0, bin 4 along y at y = -4.5: under the source
    assert abs(w[15, 4] - 1.0) < 1e-15
    s = np.array([0.0, g.source_y, g.source_r])
    e = np.array([g.x_d0 + 3.5 * g.dbx, g.y_d0 + 7.5 * g.dby, -g.detector_r])
    assert abs(w[3, 7] - (s[2] + g.detector_r) / np.linalg.norm(s - e)) < 1e-15


def test_filter_matches_brute_force():
    g = Geometry.half_cone(8, 6, 4, 1.0, 1.0, 1.0, 11, 5, 1.0, 1.0, source_to_detector=100.0)
    ang = arc_angles(2, 7.5)
    rng = np.random.default_rng(3)
    p = rng.random((2, 11, 5))
    taps = filter_taps(11, 1.0, "hann75")
    got = F.fbp_filter(g, ang, p, taps)
    c = (len(taps) - 1) // 2
    for a in range(2):
        w = F.cos_weights(g, float(ang[a]))
        for i in range(11):
            for j in range(5):
                want = sum(taps[k] * w[i + k - c, j] * p[a, i + k - c, j]
                           for k in range(len(taps)) if 0 <= i + k - c < 11)
                assert abs(got[a, i, j] - want) < 1e-12
