"""CPU restatement of ReconDBT's SART and ML loops (test infrastructure only).

ReconDBT/SART_dbt.m:40-119 and ML_dbt.m:61-113 restated in numpy FP64 on the
oracle's projector pair: G = the Siddon walk without cos (oracle.fproj), G^T =
the exact voxel-driven clip transpose without cos (oracle.bproj_clip). The
MATLAB functions call Fessler's IRT system matrices (Gcone), a third-party
package absent here; their loop structure, masks and NaN rules are kept.
Only tests/ use this module.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def _G(g, ang, x, v):
    return O.fproj(g, ang, x, v, v + 1, cos_weight=False, dtype=np.float64)


def _Gt(g, ang, s, v, nthreads):
    return O.bproj_clip(g, ang, s, v, v + 1, cos_weight=False, nthreads=nthreads)


def sart_dbt(g, ang, proj, x0, niter, stepsize, saveiter=False, nthreads=8):
    nv = len(ang)
    x = np.array(x0, np.float64)
    thresh = 5.0 * g.dz                                            # :57
    row = [O.gamma(g, ang, v, v + 1, 1, dtype=np.float64) for v in range(nv)]   # :55
    col = [_Gt(g, ang, np.ones((1,) + g.view_shape), v, nthreads) for v in range(nv)]  # :72
    y = np.zeros((1,) + g.view_shape)                              # :50
    out = []
    for _ in range(niter):
        for v in range(nv):
            pdif = proj[v:v + 1] - _G(g, ang, x, v)                # :65
            m = row[v] > thresh
            y[m] = pdif[m] / row[v][m]                             # :66
            imgi = _Gt(g, ang, y, v, nthreads)                     # :70
            with np.errstate(divide="ignore", invalid="ignore"):
                imgj = imgi / col[v]                               # :73
            imgj[np.isnan(imgj)] = 0                               # :74
            x = x + stepsize * imgj                                # :76
        x[x < 0] = 0                                               # :98
        if saveiter:
            out.append(x.copy())
    return np.stack(out) if saveiter else x


def ml_dbt(g, ang, counts, x0, I0, niter, stepsize, saveiter=False, nthreads=8):
    nv = len(ang)
    x = np.array(x0, np.float64)
    upd = np.zeros_like(x)                                         # :75
    out = []
    for _ in range(niter):
        imgi = np.zeros_like(x)
        denom = np.zeros_like(x)
        for v in range(nv):
            drr = _G(g, ang, x, v)                                 # :88
            i0 = I0 if np.ndim(I0) == 0 else np.asarray(I0, np.float64)[v:v + 1]
            yb = i0 * np.exp(-drr)                                 # :90-92
            ydif = yb - counts[v:v + 1]                            # :94
            imgi += _Gt(g, ang, ydif, v, nthreads)                 # :95
            denom += _Gt(g, ang, drr * yb, v, nthreads)            # :96
        m = denom != 0                                             # :100
        with np.errstate(divide="ignore", invalid="ignore"):
            upd[m] = imgi[m] / denom[m]                            # :101
        upd[np.isnan(upd)] = 0                                     # :102
        x = x + x * upd * stepsize                                 # :103
        x[np.isnan(x)] = 0                                         # :104
        x[x < 0] = 0                                               # :105
        if saveiter:
            out.append(x.copy())
    return np.stack(out) if saveiter else x


def ml_cost(g, ang, counts, x, I0):
    """sum over views of sum(y + proj .* drr) (ML_dbt.m:98), FP64."""
    ll = 0.0
    for v in range(len(ang)):
        drr = _G(g, ang, x, v)
        i0 = I0 if np.ndim(I0) == 0 else np.asarray(I0, np.float64)[v:v + 1]
        ll += float(np.sum(i0 * np.exp(-drr) + counts[v:v + 1] * drr))
    return ll
