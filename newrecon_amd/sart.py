"""OS-SART on the GPU SART launchers.

The reference ships the OS-SART kernels and their launch wrappers
(Tomo_OSTR.cu:1737-2183 kernels, :2249-2560 launchers: fprojectCB_1R_GPU_SART[_cos],
SART_prj_diff, bprojectCB_4B_GPU_SART[_R], prior_GPU_SART) but no loop that
calls them (Tomo_OSTR_gold.cpp's method switch has an empty SART case). This
module runs the standard ordered-subsets SART loop on them, subset by subset:

    est, norm = A_s x (cos-weighted if cos_weight), A_s 1
    diff      = (p_s - est) / norm          (0 where norm == 0)
    [prior    = 0; prior += Huber gradient of x]     when beta != 0
    x         = max(0, x + lam * (SBP(diff)/sens [+ beta prior/sens]))

p holds line integrals (log(blank / counts), or a forward projection) in the
natural view order [V][nbx][nby]; subsets use the OSTR order (subset m holds
views m, m + M, m + 2M, ...). All buffers are float32 on one device.

fused=True (default) runs each subset as two launches with the same
arithmetic: the forward writes diff directly (detector-row-major, p kept in
that layout), and the update evaluates the prior in the same pass, reading x
and writing a second image buffer.
"""
from __future__ import annotations

from typing import Callable, Optional

import torch

from .geometry import Geometry, regroup_index
from .ostr import DBTSystem


class OSSART:
    def __init__(self, geom: Geometry, angles, p, x0=None, n_subsets: int = 3, lam: float = 1.0,
                 beta: float = 0.0, delta: float = 1e-3, cos_weight: bool = False,
                 fused: bool = True, device=None):
        nv = len(angles)
        if n_subsets <= 0 or nv % n_subsets:
            raise ValueError(f"ANGLES % subset_number != 0 ({nv} % {n_subsets})")
        if nv // n_subsets > 64:
            raise ValueError("OS-SART backprojects at most 64 views per subset")
        if beta != 0.0 and delta == 0.0:
            raise ValueError("delta cannot be zero")
        self.geom, self.M, self.vs = geom, n_subsets, nv // n_subsets
        self.lam, self.beta, self.delta, self.cos_weight = lam, beta, delta, cos_weight
        self.sys = DBTSystem(geom, angles, device=device)
        self.dev = self.sys.device
        self.fused = fused
        p_nat = torch.as_tensor(p).to(device=self.dev, dtype=torch.float32).contiguous()
        self.p = self.sys.regroup_prj(p_nat, n_subsets, to_rows=fused)
        del p_nat
        self.sys.set_index(regroup_index(nv, n_subsets))
        if x0 is None:
            self.x = self.sys.volume()
        else:
            self.x = torch.as_tensor(x0).to(device=self.dev, dtype=torch.float32).contiguous()
        if fused:
            self.diff = torch.empty((self.vs, geom.nby, geom.nbx), device=self.dev)
            self.x_next = torch.empty_like(self.x)
        else:
            self.est = self.sys.sinogram(self.vs)
            self.norm = self.sys.sinogram(self.vs)
            self.diff = self.sys.sinogram(self.vs)
            self.prior = self.sys.volume() if beta != 0.0 else None
        self.iteration = 0

    def _subset(self, m: int):
        s = self.sys
        a0, a1 = m * self.vs, (m + 1) * self.vs
        if self.fused:
            s.sart_fproj_diff(self.x, self.p[a0:a1], a0, a1, self.diff, self.cos_weight)
            s.sart_update_fused(self.x, self.x_next, self.diff, a0, a1, self.lam, self.beta,
                                self.delta)
            self.x, self.x_next = self.x_next, self.x
            return
        s.fproj_sart(self.x, a0, a1, self.est, self.norm, self.cos_weight)
        s.sart_diff(self.p[a0:a1], self.est, self.norm, self.diff)
        if self.prior is not None:
            self.prior.zero_()
            s.prior_sart(self.x, self.delta, self.prior)
        s.bproj_sart(self.diff, a0, a1, self.x, self.lam, self.prior, self.beta)

    def iterate(self, n_iters: int = 1, callback: Optional[Callable] = None) -> torch.Tensor:
        """Run n_iters full passes over the subsets; callback(it, x) after each."""
        for _ in range(n_iters):
            for m in range(self.M):
                self._subset(m)
            self.iteration += 1
            if callback is not None:
                callback(self.iteration, self.x)
        return self.x
