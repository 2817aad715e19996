"""CPU restatement of the DBT FBP (newrecon_amd/fbp.py) in FP64 numpy (test
infrastructure only): cosine weight D / |S - E| per bin, linear convolution
of each detector row along x with the given odd-length centred taps, then the
oracle's SBP. Only tests/ use this module."""
from __future__ import annotations

import math

import numpy as np

from oracle import oracle as O


def cos_weights(g, theta):
    s = np.array([g.source_r * math.sin(theta), g.source_y, g.source_r * math.cos(theta)])
    D = s[2] + g.detector_r
    ex = g.x_d0 + (np.arange(g.nbx) + 0.5) * g.dbx
    ey = g.y_d0 + (np.arange(g.nby) + 0.5) * g.dby
    EX, EY = np.meshgrid(ex, ey, indexing="ij")
    return D / np.sqrt((EX - s[0]) ** 2 + (EY - s[1]) ** 2 + D * D)


def fbp_filter(g, angles, p, taps):
    """q[a][i][j] = sum_k taps[k] (w p)[a][i + k - c][j], samples off the detector 0."""
    taps = np.asarray(taps, np.float64)
    n, nt = g.nbx, len(taps)
    c = (nt - 1) // 2
    q = np.zeros((len(angles), g.nbx, g.nby))
    for a in range(len(angles)):
        pw = cos_weights(g, float(angles[a])) * np.asarray(p[a], np.float64)
        for i in range(n):
            ks = np.arange(max(0, c - i), min(nt - 1, c + n - 1 - i) + 1)
            q[a, i, :] = taps[ks] @ pw[i + ks - c, :]
    return q


def fbp(g, angles, p, taps):
    q = fbp_filter(g, angles, p, taps)
    return O.bproj_sbp(g, angles, q, 0, len(angles), dtype=np.float64)
