"""Seeded synthetic DBT phantoms and transmission data.

The reference reconstructs clinical scanner data it reads from disk
(Tomo_OSTR_gold.cpp:107-179); this module makes the synthetic stand-in:

* a half-ellipsoid breast of adipose tissue (mu = 0.0456/mm) cut by the chest
  wall plane y = 0,
* n_fg fibroglandular ellipsoids (mu = 0.0507/mm) with Gaussian-blurred edges
  (sigma = one in-plane voxel), blended into the adipose background,
* n_calc calcification spheres of radius 0.2-0.4 mm (mu = 1.2/mm) with their
  partial volume from 4x4x4 supersampling,

all drawn from numpy's PCG64 with a fixed seed, and Poisson counts
y = Poisson(I0 exp(-l)) for the cos-weighted line integrals l of the same
forward model the reconstruction uses. Sizes in mm, arrays (nz, ny, nx) f32.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erfc

from .geometry import Geometry

MU_ADIPOSE = 0.0456
MU_FIBRO = 0.0507
MU_CALC = 1.2


def _axes(g: Geometry):
    x = (g.x_p0 + (np.arange(g.nx) + 0.5) * g.dx).astype(np.float32)
    y = (g.y_p0 + (np.arange(g.ny) + 0.5) * g.dy).astype(np.float32)
    z = (g.z_p0 + (np.arange(g.nz) + 0.5) * g.dz).astype(np.float32)
    return x, y, z


def _box(c, r, origin, step, n):
    lo = max(int(math.floor((c - r - origin) / step)), 0)
    hi = min(int(math.ceil((c + r - origin) / step)) + 1, n)
    return lo, hi


def breast_phantom(g: Geometry, seed: int, n_fg=(10, 20), n_calc=5) -> np.ndarray:
    """Seeded breast phantom on the volume grid of g."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x, y, z = _axes(g)
    lx, ly, lz = g.nx * g.dx, g.ny * g.dy, g.nz * g.dz
    cx, cz = g.x_p0 + lx / 2, g.z_p0 + lz / 2
    ax, ay, az = 0.46 * lx, 0.92 * ly, 0.48 * lz
    y_wall = g.y_p0 + ly  # chest wall plane

    vol = np.zeros(g.vol_shape, np.float32)
    # breast mask, slab by slab to bound memory
    qx = ((x - cx) / ax) ** 2
    qy = ((y - y_wall) / ay) ** 2
    inplane = qy[:, None] + qx[None, :]
    for k in range(g.nz):
        qz = ((z[k] - cz) / az) ** 2
        vol[k][inplane + qz <= 1.0] = MU_ADIPOSE
    mask = vol > 0

    def inside_breast(px, py, pz, margin=0.8):
        return ((px - cx) / ax) ** 2 + ((py - y_wall) / ay) ** 2 + ((pz - cz) / az) ** 2 \
            <= margin ** 2

    # fibroglandular blobs: weight = max over blobs of a Gaussian-edged indicator
    weight = np.zeros(g.vol_shape, np.float32)
    n_blobs = int(rng.integers(n_fg[0], n_fg[1] + 1))
    scale = min(lx, ly) / 100.0  # blob size follows the breast size
    sigma = g.dx
    placed = 0
    while placed < n_blobs:
        px, py, pz = rng.uniform(g.x_p0, g.x_p0 + lx), rng.uniform(g.y_p0, y_wall), \
            rng.uniform(g.z_p0, g.z_p0 + lz)
        sa = rng.uniform(2.0, 8.0, size=2) * scale
        sc = rng.uniform(0.15, 0.35) * lz
        rot = rng.uniform(0, math.pi)
        if not inside_breast(px, py, pz):
            continue
        placed += 1
        reach = max(sa) + 3 * sigma
        x0, x1 = _box(px, reach, g.x_p0, g.dx, g.nx)
        y0, y1 = _box(py, reach, g.y_p0, g.dy, g.ny)
        z0, z1 = _box(pz, sc + 3 * g.dz, g.z_p0, g.dz, g.nz)
        if x0 >= x1 or y0 >= y1 or z0 >= z1:
            continue
        X = x[x0:x1][None, None, :] - px
        Y = y[y0:y1][None, :, None] - py
        Z = z[z0:z1][:, None, None] - pz
        c, s = math.cos(rot), math.sin(rot)
        u = (c * X + s * Y) / sa[0]
        v = (-s * X + c * Y) / sa[1]
        w = Z / sc
        rho = np.sqrt(u * u + v * v + w * w)
        # signed distance to the surface ~ (rho - 1) * smallest in-plane axis
        sd = (rho - 1.0) * min(sa)
        wt = (0.5 * erfc(sd / (sigma * math.sqrt(2.0)))).astype(np.float32)
        sub = weight[z0:z1, y0:y1, x0:x1]
        np.maximum(sub, wt, out=sub)
    vol += np.where(mask, weight * (MU_FIBRO - MU_ADIPOSE), 0).astype(np.float32)

    # calcifications with supersampled partial volume
    ss = 4
    offs = (np.arange(ss) + 0.5) / ss - 0.5
    placed = 0
    while placed < n_calc:
        px, py, pz = rng.uniform(g.x_p0, g.x_p0 + lx), rng.uniform(g.y_p0, y_wall), \
            rng.uniform(g.z_p0, g.z_p0 + lz)
        rad = rng.uniform(0.2, 0.4)
        if not inside_breast(px, py, pz, margin=0.7):
            continue
        placed += 1
        x0, x1 = _box(px, rad, g.x_p0, g.dx, g.nx)
        y0, y1 = _box(py, rad, g.y_p0, g.dy, g.ny)
        z0, z1 = _box(pz, rad, g.z_p0, g.dz, g.nz)
        if x0 >= x1 or y0 >= y1 or z0 >= z1:
            continue
        X = (x[x0:x1][:, None] + offs[None, :] * g.dx).reshape(-1)
        Y = (y[y0:y1][:, None] + offs[None, :] * g.dy).reshape(-1)
        Z = (z[z0:z1][:, None] + offs[None, :] * g.dz).reshape(-1)
        inside = ((Z[:, None, None] - pz) ** 2 + (Y[None, :, None] - py) ** 2
                  + (X[None, None, :] - px) ** 2) <= rad * rad
        frac = inside.reshape(z1 - z0, ss, y1 - y0, ss, x1 - x0, ss).mean(axis=(1, 3, 5))
        sub = vol[z0:z1, y0:y1, x0:x1]
        sub[:] = sub * (1 - frac) + MU_CALC * frac
    return vol.astype(np.float32)


def poisson_counts(line_integrals: np.ndarray, i0: float, seed: int) -> np.ndarray:
    """y = Poisson(I0 exp(-l)) as float32 (b = I0, r = 0)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lam = i0 * np.exp(-line_integrals.astype(np.float64))
    return rng.poisson(lam).astype(np.float32)
