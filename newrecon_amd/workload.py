"""Synthetic measurement for a named scan, made with the HIP projector.

The line integrals of the seeded phantom are the cos-weighted Siddon
projections of this library (the model the reconstruction uses), so noise-free
data are exactly consistent with it; Poisson noise is drawn on the host with
numpy PCG64 from the scan's noise seed.
"""
from __future__ import annotations

import numpy as np
import torch

from .geometry import Scan
from .ostr import DBTSystem
from .phantom import breast_phantom, poisson_counts


def make_measurement(scan: Scan, device=None, noise=True):
    """Returns (x_true np.float32, y torch f32 [V][nbx][nby] on device, b likewise)."""
    g = scan.geom
    x_true = breast_phantom(g, scan.phantom_seed, scan.n_fg, scan.n_calc)
    sysm = DBTSystem(g, scan.angles, device=device)
    dev = sysm.device
    xt = torch.from_numpy(x_true).to(dev)
    l = sysm.fproj(xt, 0, scan.n_views)
    if noise:
        y = torch.from_numpy(poisson_counts(l.cpu().numpy(), scan.i0, scan.noise_seed)).to(dev)
    else:
        y = scan.i0 * torch.exp(-l)
    b = torch.full_like(y, scan.i0)
    sysm.close()
    del xt, l
    return x_true, y, b
