"""The reference's run files: the key/value datafile, raw projections, angles
and reconstruction outputs.

* Datafile (readdata.cu:54-303, keys read at Tomo_OSTR.cu:2578-2650): a text
  file of whitespace-separated tokens; a parameter is the first token equal to
  its label followed by its value(s). A missing label or a numeric value that
  is `#` or starts with a letter is an error (the reference exits; here
  `DatafileError`).
* Projections (Tomo_OSTR_gold.cpp:107-179): one raw float32 file per view,
  `<folder><name>NNNN` with NNNN the 1-based view number zero-padded to four
  digits, stored [BINSy][BINSx] (x fastest) and flipped to [BINSx][BINSy]
  (x_y_flip, :93-105).
* Angles (Tomo_OSTR_gold.cpp:1474-1488): raw float32, radians, one per view.
* Initial guess / outputs (:567-589, :870-885): raw float32 [z][y][x];
  outputs named `<folder><name><iteration>` (0-based iteration number).
"""
from __future__ import annotations

import dataclasses
import re
import os
from typing import Dict, List, Optional

import numpy as np

from .geometry import Geometry, OSTRConfig


class DatafileError(ValueError):
    pass


INT_KEYS = {
    "image_size_x": "nx", "image_size_y": "ny", "image_size_z": "nz",
    "bin_samples_x": "nbx", "bin_samples_y": "nby",
}
FLOAT_KEYS = {
    "Voxel_size_x_mm": "dx", "Voxel_size_y_mm": "dy", "Voxel_size_z_mm": "dz",
    "object_xpos_mm": "x_p0", "object_ypos_mm": "y_p0", "object_zpos_mm": "z_p0",
    "detector_x0_mm": "x_d0", "detector_y0_mm": "y_d0",
    "detector_COR_mm": "detector_r", "source_COR_mm": "source_r",
    "source_Y_shift_mm": "source_y",
    "Bin_size_x_mm": "dbx", "Bin_size_y_mm": "dby",
}
RUN_INT_KEYS = ["angular_samples", "Readout_Every_X_Iters", "Recon_Method",
                "iteration_number", "subset_number"]
RUN_FLOAT_KEYS = ["Beta", "Delta"]
STRING_KEYS = ["ini_guess", "angle_name", "projection_folder", "projection_name",
               "scatter_folder", "scatter_name", "blank_folder", "blank_name",
               "output_recon_folder", "output_recon_name"]
# launch-shape keys of the reference (BACKPRJ_*, PRJ_*, nBatch*): read if
# present and ignored; this build chooses its own work-group shapes.


def _tokens(path: str) -> List[str]:
    try:
        with open(path) as f:
            return f.read().split()
    except OSError as e:
        raise DatafileError(f"{path} not found") from e


def read_param(tokens: List[str], label: str, count: int = 1, numeric: bool = True):
    """Values following the first token equal to `label`."""
    try:
        k = tokens.index(label)
    except ValueError:
        raise DatafileError(f'the parameter label "{label}" was not found') from None
    vals = tokens[k + 1:k + 1 + count]
    if len(vals) < count:
        raise DatafileError(f'too few values after "{label}"')
    if numeric:
        for v in vals:
            if v == "#" or v[:1].isalpha():
                raise DatafileError(f'something wrong with the values of "{label}": {v!r}')
    return vals


_INT_PREFIX = re.compile(r"\s*([+-]?\d+)")
_FLOAT_PREFIX = re.compile(r"\s*([+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))",
                           re.IGNORECASE)


def atoi(v: str) -> int:
    """C atoi (readdata.cu:114): the leading integer of v ("512.0" -> 512),
    0 when there is none."""
    m = _INT_PREFIX.match(v)
    return int(m.group(1)) if m else 0


def atof(v: str) -> float:
    """C atof (readdata.cu:300): the leading decimal number of v ("0.1mm" ->
    0.1), 0.0 when there is none."""
    m = _FLOAT_PREFIX.match(v)
    return float(m.group(1)) if m else 0.0


@dataclasses.dataclass
class RunFile:
    geom: Geometry
    cfg: OSTRConfig
    n_views: int
    readout_every: int
    paths: Dict[str, str]


def load_datafile(path: str) -> RunFile:
    """Parse a reference datafile (Tomo_OSTR.cu:2578-2650)."""
    tok = _tokens(path)
    g = {}
    for key, field in INT_KEYS.items():
        g[field] = atoi(read_param(tok, key)[0])
    for key, field in FLOAT_KEYS.items():
        g[field] = float(np.float32(atof(read_param(tok, key)[0])))
    run = {k: atoi(read_param(tok, k)[0]) for k in RUN_INT_KEYS}
    run.update({k: float(np.float32(atof(read_param(tok, k)[0]))) for k in RUN_FLOAT_KEYS})
    paths = {k: read_param(tok, k, numeric=False)[0] for k in STRING_KEYS}
    geom = Geometry(**g)
    geom.validate()
    if run["subset_number"] <= 0 or run["angular_samples"] % run["subset_number"]:
        raise DatafileError("angular_samples must be a multiple of subset_number")
    method = run["Recon_Method"]
    if method not in (0, 1):
        raise DatafileError(f"Recon_Method {method} is not an OSTR method (0 = Q2, 1 = Lange)")
    cfg = OSTRConfig(n_subsets=run["subset_number"], n_iters=run["iteration_number"],
                     beta=run["Beta"], delta=run["Delta"],
                     prior="q2" if method == 0 else "lange", gamma_mode="ray_length",
                     guard=False)
    return RunFile(geom, cfg, run["angular_samples"], run["Readout_Every_X_Iters"], paths)


def view_file(folder: str, name: str, view: int) -> str:
    """`<folder><name>NNNN`, NNNN = view + 1 zero-padded to 4 digits (:134-161)."""
    return f"{folder}{name}{view + 1:04d}"


def load_projections(folder: str, name: str, n_views: int, nbx: int, nby: int,
                     index: Optional[np.ndarray] = None) -> np.ndarray:
    """[V][nbx][nby] float32 from per-view raw [nby][nbx] files (load_prj + x_y_flip)."""
    index = np.arange(n_views) if index is None else index
    out = np.empty((n_views, nbx, nby), np.float32)
    for k, v in enumerate(index):
        fn = view_file(folder, name, int(v))
        raw = np.fromfile(fn, dtype=np.float32) if os.path.exists(fn) else None
        if raw is None or raw.size < nbx * nby:
            raise DatafileError(f"File {fn} is missing or truncated")
        out[k] = raw[:nbx * nby].reshape(nby, nbx).T
    return out


def save_projections(folder: str, name: str, prj: np.ndarray):
    """Inverse of load_projections (writes [nby][nbx] per view)."""
    for v in range(prj.shape[0]):
        np.ascontiguousarray(prj[v].T, dtype=np.float32).tofile(view_file(folder, name, v))


def load_angles(path: str, n_views: int) -> np.ndarray:
    a = np.fromfile(path, dtype=np.float32)
    if a.size < n_views:
        raise DatafileError(f"error in reading angles from {path}")
    return a[:n_views].copy()


def output_file(folder: str, name: str, iteration: int) -> str:
    return f"{folder}{name}{iteration}"
