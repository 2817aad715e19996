"""ReconDBT's SART and ML reconstructions on this build's projector pair.

ReconDBT (ReconDBT/SART_dbt.m, ML_dbt.m) runs view-by-view SART and the
transmission ML update with Fessler's IRT system matrices. Here G is the Siddon
forward projector without the cos weight (ostr_fproj) and G^T its exact
geometric transpose (ostr_bproj_matched without cos); the loops, masks and
NaN rules are the MATLAB functions'. Both serve as alternative initial images
for OSTR (x0 of newrecon_amd.ostr.OSTR).

proj / counts are [V][nbx][nby] float32 in the natural view order; images are
[nz][ny][nx]. The MATLAB functions also return a cost per iteration; these
return the images only.
"""
from __future__ import annotations

import torch

from .geometry import Geometry
from .ostr import DBTSystem


def _system(geom: Geometry, angles, device) -> DBTSystem:
    return DBTSystem(geom, angles, device=device)


def _f32(a, dev):
    return torch.as_tensor(a).to(device=dev, dtype=torch.float32).contiguous()


def column_sums(S: DBTSystem, v: int, ones_view: torch.Tensor) -> torch.Tensor:
    """sum(G) of view v (SART_dbt.m:72): G_v^T 1 as a volume."""
    return S.bproj_matched(ones_view, v, v + 1, img=S.volume(), cos_weight=False)


def sart_dbt(geom: Geometry, angles, proj, x0, niter: int, stepsize: float,
             saveiter: bool = False, device=None):
    """SART_dbt.m:40-119. proj: line integrals. Returns the final image, or the
    [niter][nz][ny][nx] images with saveiter."""
    S = _system(geom, angles, device)
    dev, nv = S.device, S.n_views
    p = _f32(proj, dev)
    x = _f32(x0, dev).clone()
    thresh = 5.0 * geom.dz                       # lmask = l > 5 dz (:57)
    ones = S.sinogram(1)
    ones.fill_(1.0)
    row = [S.gamma(v, v + 1, "row_sum") for v in range(nv)]        # sum(G') (:55)
    col = [column_sums(S, v, ones) for v in range(nv)]              # sum(G) (:72)
    y = S.sinogram(1)                            # persists across views (:50, :66)
    num = S.volume()
    out = []
    for _ in range(niter):
        for v in range(nv):
            est = S.fproj(x, v, v + 1, cos_weight=False)
            S.rdbt_sart_residual(y, p[v:v + 1], est, row[v], thresh)
            num.zero_()
            S.bproj_matched(y, v, v + 1, img=num, cos_weight=False)
            S.rdbt_sart_update(x, num, col[v], stepsize, clamp=(v == nv - 1))
        if saveiter:
            out.append(x.clone())
    return torch.stack(out) if saveiter else x


def ml_dbt(geom: Geometry, angles, counts, x0, I0, niter: int, stepsize: float,
           saveiter: bool = False, device=None):
    """ML_dbt.m:61-113. counts: measured counts; I0: a scalar or [V][nbx][nby]."""
    S = _system(geom, angles, device)
    dev, nv = S.device, S.n_views
    c = _f32(counts, dev)
    scalar = not torch.is_tensor(I0) and torch.as_tensor(I0).numel() == 1
    i0 = None if scalar else _f32(I0, dev)
    i0s = float(torch.as_tensor(I0).reshape(-1)[0]) if scalar else 0.0
    x = _f32(x0, dev).clone()
    upd = S.volume()                             # persists where denom == 0 (:75, :101)
    num, den = S.volume(), S.volume()
    ydif, wdrr = S.sinogram(1), S.sinogram(1)
    out = []
    for _ in range(niter):
        num.zero_()
        den.zero_()
        for v in range(nv):
            drr = S.fproj(x, v, v + 1, cos_weight=False)
            S.rdbt_ml_terms(ydif, wdrr, drr, c[v:v + 1], None if scalar else i0[v:v + 1], i0s)
            S.bproj_matched(ydif, v, v + 1, img=num, cos_weight=False)
            S.bproj_matched(wdrr, v, v + 1, img=den, cos_weight=False)
        S.rdbt_ml_update(x, upd, num, den, stepsize)
        if saveiter:
            out.append(x.clone())
    return torch.stack(out) if saveiter else x
