"""Filtered backprojection for DBT as an initial image.

ReconDBT/fbp_dbt.m resamples the views onto a virtual rotating detector and
calls IRT's feldkamp (FDK with IRT's filter windows); IRT is a third-party
package absent here. This is the textbook FDK form on the flat DBT detector,
on this build's kernels: each view is cosine-weighted (D / |S - E|) and
filtered along detector x, the tube-motion direction, with the band-limited
ramp (Ram-Lak) times a window (ostr_fbp_filter), then backprojected with the
voxel-driven SBP, whose depth weight |S|^2 / |S - V|^2 is FDK's distance
weight. Scale pi / (V dz): the 1/V average over views of a half-turn
parallel-beam FBP normalised to the arc as fbp_dbt.m does (its 180/orbit
factor), and 1/dz to cancel the SBP's slice thickness factor.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .geometry import Geometry
from .ostr import DBTSystem

WINDOWS = ("ramp", "hann", "hann50", "hann75", "hann80")


def ramlak(n: int, tau: float) -> np.ndarray:
    """Discrete band-limited ramp filter h[k], k = -(n-1)..n-1 (Kak & Slaney,
    Principles of Computerized Tomographic Imaging, eq. 3.61): 1/(4 tau^2) at 0,
    -1/(pi k tau)^2 at odd k, 0 at even k."""
    k = np.arange(-(n - 1), n, dtype=np.float64)
    h = np.zeros(2 * n - 1)
    h[k == 0] = 1.0 / (4.0 * tau * tau)
    odd = (k.astype(np.int64) % 2) != 0
    h[odd] = -1.0 / (math.pi * k[odd] * tau) ** 2
    return h


def filter_taps(n: int, tau: float, window: str = "hann") -> np.ndarray:
    """Odd-length centred taps (length 2n - 1) of the windowed ramp for
    detector rows of n bins of size tau, including the tau of the convolution
    sum. 'hannXX': Hann window reaching 0 at XX% of the Nyquist frequency."""
    if window not in WINDOWS:
        raise ValueError(f"unknown window {window!r}")
    h = ramlak(n, tau) * tau
    if window == "ramp":
        return h
    cut = 1.0 if window == "hann" else int(window[4:]) / 100.0
    L = 1 << int(math.ceil(math.log2(4 * n)))
    buf = np.zeros(L)
    buf[:n] = h[n - 1:]            # k = 0..n-1
    buf[L - (n - 1):] = h[:n - 1]  # k = -(n-1)..-1
    f = np.abs(np.fft.fftfreq(L, d=tau))
    fc = cut / (2.0 * tau)
    win = np.where(f <= fc, 0.5 * (1.0 + np.cos(np.pi * f / fc)), 0.0)
    out = np.real(np.fft.ifft(np.fft.fft(buf) * win))
    return np.concatenate([out[L - (n - 1):], out[:n]])


def fbp_dbt(geom: Geometry, angles, proj, window: str = "hann", device=None):
    """FBP image of line integrals proj [V][nbx][nby] (natural view order)."""
    S = DBTSystem(geom, angles, device=device)
    dev, nv = S.device, S.n_views
    p = torch.as_tensor(proj).to(device=dev, dtype=torch.float32).contiguous()
    scale = math.pi / (nv * geom.dz)
    taps = torch.from_numpy((filter_taps(geom.nbx, geom.dbx, window) * scale)
                            .astype(np.float32)).to(dev)
    q = S.fbp_filter(p, taps, 0, nv)
    return S.bproj_sbp(q, 0, nv)
