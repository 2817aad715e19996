"""Scan geometry and run configuration of the OSTR DBT path.

`Geometry` carries the scalars the reference keeps in process-global
``__constant__`` memory (Tomo_OSTR.cu:81-115, uploaded at :2653-2690 from the
datafile keys read at :2578-2650); here they are a per-handle value passed to
``ostr_create``. `OSTRConfig` carries the remaining datafile keys of the
reconstruction itself (Beta, Delta, iteration_number, subset_number,
Recon_Method) plus the switches that make the reference's quirks explicit.

Coordinates (mm): volume voxel (ix,iy,iz) spans [x_p0+ix*dx, x_p0+(ix+1)*dx]
etc.; the flat detector is stationary in the plane z = -detector_r with bin
(i,j) centred at (x_d0+(i+0.5)*dbx, y_d0+(j+0.5)*dby); the source of view t
sits at (source_r*sin t, source_y, source_r*cos t) (Tomo_OSTR.cu:1340-1358).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import List, Optional

import numpy as np


class OstrGeom(ctypes.Structure):
    """ctypes mirror of ``ostr_geom`` in include/ostr_hip.h."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
        ("x_p0", ctypes.c_float), ("y_p0", ctypes.c_float), ("z_p0", ctypes.c_float),
        ("nbx", ctypes.c_int32), ("nby", ctypes.c_int32),
        ("dbx", ctypes.c_float), ("dby", ctypes.c_float),
        ("x_d0", ctypes.c_float), ("y_d0", ctypes.c_float),
        ("detector_r", ctypes.c_float), ("source_r", ctypes.c_float),
        ("source_y", ctypes.c_float),
    ]


@dataclasses.dataclass(frozen=True)
class Geometry:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    x_p0: float
    y_p0: float
    z_p0: float
    nbx: int
    nby: int
    dbx: float
    dby: float
    x_d0: float
    y_d0: float
    detector_r: float
    source_r: float
    source_y: float

    @property
    def n_vox(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def n_bin(self) -> int:
        return self.nbx * self.nby

    @property
    def vol_shape(self):
        """numpy shape of a volume: (nz, ny, nx), x fastest."""
        return (self.nz, self.ny, self.nx)

    @property
    def view_shape(self):
        """numpy shape of one view: (nbx, nby), nby fastest."""
        return (self.nbx, self.nby)

    def to_c(self) -> OstrGeom:
        return OstrGeom(**dataclasses.asdict(self))

    def validate(self) -> None:
        if min(self.nx, self.ny, self.nz, self.nbx, self.nby) <= 0:
            raise ValueError("grid sizes must be positive")
        if max(self.nx, self.ny, self.nz) >= 32768:
            raise ValueError("volume dimensions must be < 32768 (16-bit voxel walk)")
        if min(self.dx, self.dy, self.dz, self.dbx, self.dby) <= 0:
            raise ValueError("voxel and bin sizes must be positive")

    @staticmethod
    def half_cone(nx, ny, nz, dx, dy, dz, nbx, nby, dbx, dby,
                  source_to_detector=660.0, cor_above_detector=0.0,
                  source_y=0.0) -> "Geometry":
        """Reference half-cone placement (ostr.py:32-37): chest wall at y=0,
        volume y <= 0 resting on z = z_p0 = 0, detector centred in x and
        extending to y <= 0."""
        return Geometry(
            nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz,
            x_p0=-nx * dx / 2.0, y_p0=-ny * dy, z_p0=0.0,
            nbx=nbx, nby=nby, dbx=dbx, dby=dby,
            x_d0=-nbx * dbx / 2.0, y_d0=-nby * dby,
            detector_r=cor_above_detector,
            source_r=source_to_detector - cor_above_detector,
            source_y=source_y,
        )


POTENTIALS = {"q2": 0, "lange": 1, "hyperbola": 2}
GAMMA_MODES = {"ray_length": 0, "row_sum": 1}
BACKPROJECTORS = {"sbp": 0, "matched": 1}


@dataclasses.dataclass
class OSTRConfig:
    """Reconstruction keys (Tomo_OSTR.cu:2578-2650) plus quirk switches."""

    n_subsets: int = 3            # subset_number
    n_iters: int = 10             # iteration_number
    beta: float = 1e3             # Beta
    delta: float = 5e-4           # Delta
    prior: str = "hyperbola"      # Recon_Method 0 -> "q2", 1 -> "lange"
    gamma_mode: str = "row_sum"   # "ray_length" reproduces Tomo_OSTR.cu:1172
    guard: bool = True            # False reproduces the 0/0 NaN of update_est
    fused: bool = True            # fused kernels vs. launcher-by-launcher path
    # "sbp": the reference's bilinear backprojector for d and capL (not the
    # transpose of the forward model, so the iteration is not a descent method
    # and does not converge); "matched": the exact transpose G^T of
    # G = cos(t) A, d = G^T(yry * G 1) (gamma_mode unused): OS-SQS proper.
    backprojector: str = "sbp"
    # Nesterov momentum on each full OS pass (accelerated OSTR) with the
    # gradient restart test; off: plain OSTR.
    momentum: bool = False
    # the denominator d = B(gamma yry) summed over views in FP64 (rounded to
    # FP32 once) instead of in FP32
    d_fp64: bool = False

    @property
    def potential(self) -> int:
        return POTENTIALS[self.prior]

    @property
    def gamma_code(self) -> int:
        return GAMMA_MODES[self.gamma_mode]


@dataclasses.dataclass
class Scan:
    """A named workload: geometry, view angles, dose and phantom recipe."""

    name: str
    geom: Geometry
    angles: np.ndarray            # float32 radians, natural view order
    i0: float
    n_fg: tuple                   # range of fibroglandular blob counts
    n_calc: int
    phantom_seed: int
    noise_seed: int
    recon: OSTRConfig

    @property
    def n_views(self) -> int:
        return int(self.angles.shape[0])


def arc_angles(n_views: int, half_span_deg: float) -> np.ndarray:
    """n_views angles evenly over [-half_span, +half_span] degrees, radians f32."""
    return np.radians(np.linspace(-half_span_deg, half_span_deg, n_views)).astype(np.float32)


def scan_small() -> Scan:
    """200x100x20 voxels of 0.5x0.5x1 mm, 9 views over +-7.5 deg, 256x128
    detector of 0.56 mm, I0 = 1e4, 3 subsets x 10 iterations."""
    g = Geometry.half_cone(200, 100, 20, 0.5, 0.5, 1.0, 256, 128, 0.56, 0.56)
    return Scan("dbt-small-200x100x20-9v", g, arc_angles(9, 7.5), 1e4, (10, 20), 5, 1, 1001,
                OSTRConfig(n_subsets=3, n_iters=10))


def scan_clinical() -> Scan:
    """1400x600x60 voxels of 0.1x0.1x1 mm, 15 views over +-7.5 deg, 2048x1664
    detector of 0.14 mm, I0 = 5e3, 5 subsets x 20 iterations."""
    g = Geometry.half_cone(1400, 600, 60, 0.1, 0.1, 1.0, 2048, 1664, 0.14, 0.14)
    return Scan("dbt-clinical-1400x600x60-15v", g, arc_angles(15, 7.5), 5e3, (50, 200), 20,
                2, 1002, OSTRConfig(n_subsets=5, n_iters=20))


def scan_tiny(nx=12, ny=10, nz=6, nbx=24, nby=16, n_views=3) -> Scan:
    """Tiny geometry for unit tests (same placement rules, coarse voxels)."""
    g = Geometry.half_cone(nx, ny, nz, 2.0, 2.0, 2.0, nbx, nby, 2.0, 2.0)
    return Scan(f"dbt-tiny-{nx}x{ny}x{nz}-{n_views}v", g, arc_angles(n_views, 7.5), 1e4,
                (2, 3), 1, 7, 1007, OSTRConfig(n_subsets=n_views, n_iters=3))


def scan_train5(seed: int = 100000) -> Scan:
    """A training case of the learned warm start at its stated size: 256x256x48
    voxels of 0.1x0.1x1 mm, 15 views over +-7.5 deg, I0 = 5e3, a 320x224
    detector of 0.14 mm (the volume's shadow with a margin), 10-20
    fibroglandular blobs and 5 calcifications; phantom seed `seed`, noise seed
    seed + 1000; its target is a 50-iteration OSTR (5 subsets)."""
    g = Geometry.half_cone(256, 256, 48, 0.1, 0.1, 1.0, 320, 224, 0.14, 0.14)
    return Scan(f"dbt-train-256x256x48-15v-{seed}", g, arc_angles(15, 7.5), 5e3, (10, 20), 5,
                seed, seed + 1000, OSTRConfig(n_subsets=5, n_iters=50))


BATCH_THICKNESS = (30, 40, 45, 50, 55, 60, 70, 80)


def scan_batch(scale: int = 1) -> List[Scan]:
    """The uneven batch: 8 cases of 30..80 slices of 1 mm, in-plane sizes
    linearly spaced from 800x400 to 1400x700 voxels of 0.1 mm, view counts
    cycling 9/15/25 over +-7.5/15/25 deg with 3/5/5 subsets, 2048x1664
    detector of 0.14 mm, I0 = 5e3; seeds 40..47 (phantom), 1040..1047 (noise).
    `scale` > 1 divides the in-plane sizes, thickness and detector by it (voxel
    and bin sizes times it), a reduced copy of the same table for tests."""
    out = []
    for k, nz in enumerate(BATCH_THICKNESS):
        nx = int(round(800 + 600 * k / 7)) // scale
        ny = int(round(400 + 300 * k / 7)) // scale
        nv, span, m = ((9, 7.5, 3), (15, 15.0, 5), (25, 25.0, 5))[k % 3]
        g = Geometry.half_cone(nx, ny, max(nz // scale, 2), 0.1 * scale, 0.1 * scale,
                               1.0 * scale, 2048 // scale, 1664 // scale, 0.14 * scale,
                               0.14 * scale)
        out.append(Scan(f"dbt-batch{k}-{g.nx}x{g.ny}x{g.nz}-{nv}v", g, arc_angles(nv, span),
                        5e3, (50, 200), 20, 40 + k, 1040 + k,
                        OSTRConfig(n_subsets=m, n_iters=20)))
    return out


SCANS = {"small": scan_small, "clinical": scan_clinical, "tiny": scan_tiny}


def regroup_index(n_views: int, n_subsets: int) -> np.ndarray:
    """Subset-major slot->view table: slot m*Vs+j holds view j*M+m
    (Tomo_OSTR_gold.cpp:628-642)."""
    if n_subsets <= 0 or n_views % n_subsets:
        raise ValueError(f"n_views={n_views} is not a multiple of n_subsets={n_subsets}")
    vs = n_views // n_subsets
    return np.array([j * n_subsets + m for m in range(n_subsets) for j in range(vs)],
                    dtype=np.int32)
