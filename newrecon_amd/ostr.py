"""OSTR DBT reconstruction host: ctypes over libostr_hip.so.

This module is the Python host of the path (the role ostr.py / the notebook
driver cell and Tomo_OSTR_gold.cpp play in the reference). PyTorch is used
only for device buffers and the current HIP stream; all arithmetic runs in the
gfx950 kernels of libostr_hip.so. If the library is missing the import of
`lib()` raises: there is no CPU fallback.

Two levels:
* `DBTSystem` - one library handle (geometry + view table) with the
  reference's launcher and host-op names (fprojectCB_1R_GPU_OSTR_cos,
  bprojectCB_GPU_SBP, prior_GPU_OSTR, compute_h, update_est, ...), same
  argument meaning: device tensors in, results written or accumulated in
  place, half-open view ranges [angleStart, angleEnd).
* `OSTR` - the reconstruction driver of Tomo_OSTR_gold.cpp:301-918: the
  precompute of gamma/yry/d, the OS regroup, and per subset forward
  projection -> h -> backprojection -> prior -> update. `fused=False` issues
  the reference's launch sequence one operation at a time; `fused=True` uses
  the fused forward+h and backprojection+prior+update kernels (same
  arithmetic, two launches per subset).
"""
from __future__ import annotations

import ctypes
import os
import time
from typing import Callable, List, Optional

import numpy as np
import torch

from .geometry import (BACKPROJECTORS, GAMMA_MODES, POTENTIALS, Geometry, OSTRConfig,
                       regroup_index)

LIB_PATH = os.environ.get("OSTR_HIP_LIB") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "libostr_hip.so")

F32, BF16, F16 = 0, 1, 2
_DT = {torch.float32: F32, torch.bfloat16: BF16, torch.float16: F16}


class OSTRError(RuntimeError):
    """A libostr_hip call returned an error code."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load libostr_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                           "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    from .geometry import OstrGeom
    sig = {
        "ostr_create": [ctypes.POINTER(OstrGeom), vp, vp, i32, ctypes.POINTER(vp)],
        "ostr_set_index": [vp, vp],
        "ostr_fproj": [vp, vp, i32, vp, i32, i32, i32, i32, vp],
        "ostr_gamma": [vp, vp, i32, i32, i32, vp],
        "ostr_bproj_sbp": [vp, vp, vp, i32, i32, i32, vp],
        "ostr_prior": [vp, vp, vp, vp, i32, i32, f32, f32, vp],
        "ostr_yry": [vp, vp, vp, vp, i64, vp],
        "ostr_mul": [vp, vp, vp, vp, i64, vp],
        "ostr_compute_h": [vp, vp, vp, vp, vp, vp, i64, vp],
        "ostr_update_est": [vp, vp, vp, vp, vp, vp, f32, f32, i32, vp],
        "ostr_regroup": [vp, vp, vp, i32, i32, vp],
        "ostr_to_rows": [vp, vp, vp, i32, i32, vp],
        "ostr_bproj_sbp_rows": [vp, vp, vp, i32, i32, i32, vp],
        "ostr_bproj_matched": [vp, vp, vp, i32, i32, i32, i32, i32, vp],
        "ostr_bp_init": [vp, vp, vp, vp, i32, vp],
        "ostr_precompute_d": [vp, vp, vp, vp, i32, vp],
        "ostr_fproj_h": [vp, vp, i32, vp, vp, vp, vp, i32, i32, i32, vp],
        "ostr_update_fused": [vp, vp, vp, i32, vp, vp, i32, i32, i32, f32, f32, i32, f32, f32,
                              i32, vp],
        "ostr_count_nnz": [vp, i32, i32, ctypes.POINTER(i64), vp],
        "ostr_objective": [vp, vp, i32, vp, vp, vp, i32, ctypes.c_double, ctypes.c_double,
                           ctypes.POINTER(ctypes.c_double), vp],
        "ostr_objective_noalloc": [vp, vp, i32, vp, vp, vp, i32, ctypes.c_double,
                                   ctypes.c_double, ctypes.POINTER(ctypes.c_double), vp],
        "ostr_diff_norms": [vp, vp, vp, i32, i64, ctypes.POINTER(ctypes.c_double), vp],
        "ostr_extrapolate": [vp, vp, vp, vp, f32, vp],
        "ostr_restart_dot": [vp, vp, vp, vp, ctypes.POINTER(ctypes.c_double), vp],
        "ostr_fproj_sart": [vp, vp, vp, vp, i32, i32, i32, vp],
        "ostr_sart_prj_diff": [vp, vp, vp, vp, vp, i64, vp],
        "ostr_bproj_sart": [vp, vp, vp, i32, i32, f32, vp],
        "ostr_bproj_sart_r": [vp, vp, vp, vp, i32, i32, f32, f32, vp],
        "ostr_prior_sart": [vp, vp, vp, f32, vp],
        "ostr_sart_fproj_diff": [vp, vp, vp, vp, i32, i32, i32, vp],
        "ostr_fproj_4r": [vp, vp, vp, vp, i32, i32, vp],
        "ostr_line_integrals": [vp, vp, vp, vp, i64, vp],
        "ostr_fbp_filter": [vp, vp, vp, vp, i32, i32, i32, vp],
        "ostr_momentum_init": [vp, vp, vp],
        "ostr_update_fused_matched": [vp, vp, vp, i32, vp, vp, i32, i32, i32, f32, f32, i32, f32,
                                      f32, i32, vp],
        "ostr_momentum_step": [vp, vp, vp, vp, vp, i32, vp, vp],
        "ostr_accum_f64": [vp, vp, vp, i64, vp],
        "ostr_f64_to_f32": [vp, vp, vp, i64, vp],
        "ostr_precompute_d_f64": [vp, vp, vp, vp, i32, vp],
        "ostr_rdbt_sart_residual": [vp, vp, vp, vp, vp, f32, i64, vp],
        "ostr_rdbt_sart_update": [vp, vp, vp, vp, f32, i32, i64, vp],
        "ostr_rdbt_ml_terms": [vp, vp, vp, vp, vp, vp, f32, i64, vp],
        "ostr_rdbt_ml_update": [vp, vp, vp, vp, vp, f32, i64, vp],
        "ostr_sart_update_fused": [vp, vp, vp, vp, i32, i32, f32, f32, f32, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.ostr_destroy.argtypes = [vp]
    L.ostr_destroy.restype = None
    L.ostr_last_error.argtypes = [vp]
    L.ostr_last_error.restype = ctypes.c_char_p
    L.ostr_n_views.argtypes = [vp]
    L.ostr_n_views.restype = i32
    _lib = L
    return L


def _p(t: Optional[torch.Tensor]):
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensors passed to libostr_hip must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _dt(t: torch.Tensor) -> int:
    try:
        return _DT[t.dtype]
    except KeyError:
        raise TypeError(f"unsupported dtype {t.dtype}") from None


def _f32(t: torch.Tensor, what: str):
    if t.dtype != torch.float32:
        raise TypeError(f"{what} must be float32")


class DBTSystem:
    """One geometry + view table on one device (a libostr_hip handle)."""

    def __init__(self, geom: Geometry, angles, index=None, device=None):
        geom.validate()
        self.geom = geom
        self.angles = np.ascontiguousarray(np.asarray(angles, np.float32))
        self.n_views = int(self.angles.shape[0])
        self.device = torch.device(device if device is not None else "cuda")
        self._L = lib()
        idx = None if index is None else np.ascontiguousarray(np.asarray(index, np.int32))
        self.index = np.arange(self.n_views, dtype=np.int32) if idx is None else idx.copy()
        h = ctypes.c_void_p()
        self._gc = geom.to_c()
        with torch.cuda.device(self.device):
            rc = self._L.ostr_create(ctypes.byref(self._gc), self.angles.ctypes.data_as(ctypes.c_void_p),
                                     None if idx is None else idx.ctypes.data_as(ctypes.c_void_p),
                                     self.n_views, ctypes.byref(h))
        if rc != 0:
            raise OSTRError(rc, self._L.ostr_last_error(None).decode())
        self._h = h

    # -- plumbing ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._L.ostr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            raise OSTRError(rc, self._L.ostr_last_error(self._h).decode())

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def volume(self, dtype=torch.float32, fill=0.0):
        return torch.full(self.geom.vol_shape, fill, dtype=dtype, device=self.device)

    def sinogram(self, n_slots: int, dtype=torch.float32):
        return torch.zeros((n_slots,) + self.geom.view_shape, dtype=dtype, device=self.device)

    def set_index(self, index):
        idx = np.ascontiguousarray(np.asarray(index, np.int32))
        self._check(self._L.ostr_set_index(self._h, idx.ctypes.data_as(ctypes.c_void_p)))
        self.index = idx.copy()

    # -- projector family ---------------------------------------------------
    def fproj(self, img, a0, a1, prj=None, cos_weight=True, prj_dtype=torch.float32):
        prj = self.sinogram(a1 - a0, prj_dtype) if prj is None else prj
        self._check(self._L.ostr_fproj(self._h, _p(img), _dt(img), _p(prj), _dt(prj), a0, a1,
                                       int(cos_weight), self._stream()))
        return prj

    def gamma(self, a0, a1, mode="ray_length", out=None):
        out = self.sinogram(a1 - a0) if out is None else out
        _f32(out, "gamma")
        self._check(self._L.ostr_gamma(self._h, _p(out), a0, a1, GAMMA_MODES[mode],
                                       self._stream()))
        return out

    def bproj_sbp(self, prj, a0, a1, img=None):
        img = self.volume() if img is None else img
        _f32(img, "backprojection output")
        self._check(self._L.ostr_bproj_sbp(self._h, _p(img), _p(prj), _dt(prj), a0, a1,
                                           self._stream()))
        return img

    def bproj_matched(self, prj, a0, a1, img=None, cos_weight=True, rows=False):
        """img += (cos A)^T prj: the exact transpose of fproj (prj [S][nbx][nby],
        or detector-row-major [S][nby][nbx] with rows=True)."""
        img = self.volume() if img is None else img
        _f32(img, "backprojection output")
        self._check(self._L.ostr_bproj_matched(self._h, _p(img), _p(prj), _dt(prj), a0, a1,
                                               int(cos_weight), int(rows), self._stream()))
        return img

    def bp_init(self, p=None, y=None, b=None, img=None):
        """One-step backprojected image (ReconDBT/BP.m:54-82) of line integrals
        p, or of counts y with blank b (p = log(b / max(y, 1))); natural view
        order, reference layout [V][nbx][nby]."""
        img = self.volume() if img is None else img
        counts = p is None
        src = y if counts else p
        if src is None or (counts and b is None):
            raise ValueError("bp_init needs p, or y and b")
        self._check(self._L.ostr_bp_init(self._h, _p(img), _p(src), _p(b) if counts else None,
                                         int(counts), self._stream()))
        return img

    def prior(self, img, potential="hyperbola", delta=5e-4, z_xy_ratio=None, rdd=None,
              rd=None):
        rho = self.geom.dz / self.geom.dx if z_xy_ratio is None else z_xy_ratio
        rdd = self.volume() if rdd is None else rdd
        rd = self.volume() if rd is None else rd
        self._check(self._L.ostr_prior(self._h, _p(rdd), _p(rd), _p(img), _dt(img),
                                       POTENTIALS[potential], delta, rho, self._stream()))
        return rdd, rd

    def count_nnz(self, a0, a1) -> int:
        out = ctypes.c_int64()
        self._check(self._L.ostr_count_nnz(self._h, a0, a1, ctypes.byref(out), self._stream()))
        return int(out.value)

    def objective(self, img, y, b, r=None, potential="hyperbola", delta=5e-4, z_xy_ratio=None):
        """FP64 (data term, penalty R); the objective is data + beta * R.
        y, b, r: natural view order, reference layout [V][nbx][nby]."""
        rho = self.geom.dz / self.geom.dx if z_xy_ratio is None else z_xy_ratio
        out = (ctypes.c_double * 2)()
        self._check(self._L.ostr_objective(self._h, _p(img), _dt(img), _p(y), _p(b), _p(r),
                                           POTENTIALS[potential], delta, rho, out,
                                           self._stream()))
        return float(out[0]), float(out[1])

    def diff_norms(self, a, b):
        """FP64 (sum (a-b)^2, sum b^2)."""
        if a.dtype != b.dtype or a.numel() != b.numel():
            raise ValueError("diff_norms: a and b must have the same dtype and size")
        out = (ctypes.c_double * 2)()
        self._check(self._L.ostr_diff_norms(self._h, _p(a), _p(b), _dt(a), a.numel(), out,
                                            self._stream()))
        return float(out[0]), float(out[1])

    def rel_rmse(self, a, ref) -> float:
        """||a - ref|| / ||ref|| in FP64."""
        dd, bb = self.diff_norms(a, ref)
        return (dd / bb) ** 0.5 if bb > 0 else float("inf")

    # -- fused --------------------------------------------------------------
    def extrapolate(self, x, z, z_prev, c):
        """x = max(0, z + c (z - z_prev)) (momentum step)."""
        self._check(self._L.ostr_extrapolate(self._h, _p(x), _p(z), _p(z_prev), c,
                                             self._stream()))
        return x

    def momentum_init(self, state):
        """state (4 doubles on the device) = {t, c, restarts, test} = {1, 0, 0, 0}."""
        self._check(self._L.ostr_momentum_init(self._h, _p(state), self._stream()))
        return state

    def momentum_step(self, x_next, z, z_prev, x_k, state):
        """Restart test, Nesterov coefficient and extrapolation, all on the
        device (asynchronous): x_next = max(0, z + c (z - z_prev))."""
        dt = x_next.dtype
        if any(t.dtype != dt for t in (z, z_prev, x_k)):
            raise TypeError("momentum_step: the four images must share one dtype")
        self._check(self._L.ostr_momentum_step(self._h, _p(x_next), _p(z), _p(z_prev), _p(x_k),
                                               _dt(x_next), _p(state), self._stream()))
        return x_next

    def restart_dot(self, x, z, z_prev) -> float:
        """sum (x - z)(z - z_prev) in FP64 (synchronous)."""
        out = ctypes.c_double()
        self._check(self._L.ostr_restart_dot(self._h, _p(x), _p(z), _p(z_prev),
                                             ctypes.byref(out), self._stream()))
        return out.value

    def precompute_d(self, y, r=None, gamma_mode="row_sum", d=None, fp64=False):
        """d = SBP_all(gamma yry); fp64: the sum over views in double."""
        d = self.volume() if d is None else d
        fn = self._L.ostr_precompute_d_f64 if fp64 else self._L.ostr_precompute_d
        self._check(fn(self._h, _p(d), _p(y), _p(r), GAMMA_MODES[gamma_mode], self._stream()))
        return d

    def accum_f64(self, acc, x):
        """acc (float64) += x (float32), elementwise on the device."""
        if acc.dtype != torch.float64 or acc.numel() != x.numel():
            raise TypeError("accum_f64: acc must be float64 with x's size")
        _f32(x, "x")
        self._check(self._L.ostr_accum_f64(self._h, _p(acc), _p(x), x.numel(), self._stream()))
        return acc

    def f64_to_f32(self, acc, out=None):
        out = torch.empty(acc.shape, dtype=torch.float32, device=acc.device) if out is None \
            else out
        self._check(self._L.ostr_f64_to_f32(self._h, _p(out), _p(acc), acc.numel(),
                                            self._stream()))
        return out

    def fproj_h(self, img, y, b, r, a0, a1, hsub=None, prj_dtype=torch.float32):
        hsub = self.sinogram(a1 - a0, prj_dtype) if hsub is None else hsub
        self._check(self._L.ostr_fproj_h(self._h, _p(img), _dt(img), _p(y), _p(b), _p(r),
                                         _p(hsub), _dt(hsub), a0, a1, self._stream()))
        return hsub

    def update_fused(self, x_in, x_out, d, hsub, a0, a1, beta, capl_scale, potential,
                     delta, z_xy_ratio=None, guard=True):
        rho = self.geom.dz / self.geom.dx if z_xy_ratio is None else z_xy_ratio
        self._check(self._L.ostr_update_fused(
            self._h, _p(x_in), _p(x_out), _dt(x_in), _p(d), _p(hsub), _dt(hsub), a0, a1, beta,
            capl_scale, POTENTIALS[potential], delta, rho, int(guard), self._stream()))
        return x_out

    def update_fused_matched(self, x_in, x_out, d, hsub, a0, a1, beta, capl_scale, potential,
                             delta, z_xy_ratio=None, guard=True):
        """update_fused with the exact transpose: x_out = x_in - (M G^T h + beta RD) /
        (d + 2 beta RDD); image and h in F32 / BF16 / F16 storage, d FP32, h
        detector-row-major."""
        rho = self.geom.dz / self.geom.dx if z_xy_ratio is None else z_xy_ratio
        _f32(d, "d")
        if x_out.dtype != x_in.dtype:
            raise TypeError("update_fused_matched: x_in and x_out must share one dtype")
        self._check(self._L.ostr_update_fused_matched(
            self._h, _p(x_in), _p(x_out), _dt(x_in), _p(d), _p(hsub), _dt(hsub), a0, a1, beta,
            capl_scale, POTENTIALS[potential], delta, rho, int(guard), self._stream()))
        return x_out

    # -- reference names (Tomo_OSTR.cu launch wrappers, gold.cpp host ops) ----
    def fprojectCB_1R_GPU_OSTR_cos(self, estbuf, prj_est, angleStart, angleEnd):
        """Tomo_OSTR.cu:2334: prj_est[a-angleStart] = cos t * A_a estbuf."""
        return self.fproj(estbuf, angleStart, angleEnd, prj=prj_est, cos_weight=True)

    def fprojectCB_1R_GPU_OSTR_normprj(self, d_normprj, angleStart, angleEnd,
                                       mode="ray_length"):
        """Tomo_OSTR.cu:2361 (mode ray_length) / ostr.ipynb normprj (row_sum)."""
        return self.gamma(angleStart, angleEnd, mode, out=d_normprj)

    def bprojectCB_GPU_SBP(self, d_objbuf, d_prjbuf, angleStart, angleEnd):
        """Tomo_OSTR.cu:2435: d_objbuf += SBP(d_prjbuf) over [angleStart, angleEnd)."""
        return self.bproj_sbp(d_prjbuf, angleStart, angleEnd, img=d_objbuf)

    def prior_GPU_OSTR(self, d_RDD, d_RD, d_est, delta, z_xy_ratio, potential="lange"):
        """Tomo_OSTR.cu:2216 (Lange); potential='hyperbola' for the hyperbola prior."""
        if delta == 0:
            raise OSTRError(-1, "delta cannot be zero")
        return self.prior(d_est, potential, delta, z_xy_ratio, rdd=d_RDD, rd=d_RD)

    def prior_GPU_OSTR_Q2(self, d_RDD, d_RD, d_est, z_xy_ratio):
        """Tomo_OSTR.cu:2189."""
        return self.prior(d_est, "q2", 1.0, z_xy_ratio, rdd=d_RDD, rd=d_RD)

    def compute_yry(self, yry, prj, scat=None):
        """Tomo_OSTR_gold.cpp:183: yry = (y-r)^2/y (y-r<=0 -> y; y==0 -> 0)."""
        self._check(self._L.ostr_yry(self._h, _p(yry), _p(prj), _p(scat), prj.numel(),
                                     self._stream()))
        return yry

    def compute_gamma_yry(self, gamma_yry, yry, gamma):
        """Tomo_OSTR_gold.cpp:208: gamma_yry = yry * gamma."""
        self._check(self._L.ostr_mul(self._h, _p(gamma_yry), _p(yry), _p(gamma), yry.numel(),
                                     self._stream()))
        return gamma_yry

    def compute_h(self, h_sub, prj_sub, blank_sub, line_sub, scat_sub=None):
        """Tomo_OSTR_gold.cpp:245: h = (y/(b e^-l + r) - 1) b e^-l."""
        self._check(self._L.ostr_compute_h(self._h, _p(h_sub), _p(prj_sub), _p(blank_sub),
                                           _p(line_sub), _p(scat_sub), prj_sub.numel(),
                                           self._stream()))
        return h_sub

    def update_est(self, est, capL, RD, d, RDD, beta, capl_scale=1.0, guard=False):
        """Tomo_OSTR_gold.cpp:265 (+ the capL*M scaling of :769)."""
        self._check(self._L.ostr_update_est(self._h, _p(est), _p(capL), _p(RD), _p(d), _p(RDD),
                                            beta, capl_scale, int(guard), self._stream()))
        return est

    def regroup_prj(self, src, n_subsets, dst=None, to_rows=False):
        """Tomo_OSTR_gold.cpp:220: subset-major reorder of [V][nbx][nby]
        (to_rows: output detector-row-major [V][nby][nbx])."""
        if dst is None:
            shape = (src.shape[0], self.geom.nby, self.geom.nbx) if to_rows else src.shape
            dst = torch.empty(shape, dtype=src.dtype, device=src.device)
        self._check(self._L.ostr_regroup(self._h, _p(dst), _p(src), n_subsets, int(to_rows),
                                         self._stream()))
        return dst

    def to_rows(self, src, dst=None):
        """[S][nbx][nby] -> detector-row-major [S][nby][nbx]."""
        dst = (torch.empty((src.shape[0], self.geom.nby, self.geom.nbx), dtype=src.dtype,
                           device=src.device) if dst is None else dst)
        self._check(self._L.ostr_to_rows(self._h, _p(dst), _p(src), src.shape[0], _dt(src),
                                         self._stream()))
        return dst

    def bproj_sbp_rows(self, prj_rows, a0, a1, img=None):
        """SBP of a detector-row-major sinogram [S][nby][nbx], accumulated into img."""
        img = self.volume() if img is None else img
        _f32(img, "backprojection output")
        self._check(self._L.ostr_bproj_sbp_rows(self._h, _p(img), _p(prj_rows), _dt(prj_rows),
                                                a0, a1, self._stream()))
        return img

    # -- OS-SART launchers ----------------------------------------------------
    def fproj_sart(self, img, a0, a1, prj=None, norm=None, cos_weight=False):
        """prj = (cos t) A img (cos_weight) or A img, norm = A 1, over [a0, a1)."""
        prj = self.sinogram(a1 - a0) if prj is None else prj
        norm = self.sinogram(a1 - a0) if norm is None else norm
        for t, what in ((img, "SART image"), (prj, "SART projection"), (norm, "SART norm")):
            _f32(t, what)
        self._check(self._L.ostr_fproj_sart(self._h, _p(img), _p(prj), _p(norm), a0, a1,
                                            int(cos_weight), self._stream()))
        return prj, norm

    def sart_diff(self, p, est, norm, diff=None):
        """diff = (p - est) / norm where norm != 0, else 0."""
        diff = torch.empty_like(est) if diff is None else diff
        for t, what in ((diff, "diff"), (p, "p"), (est, "est"), (norm, "norm")):
            _f32(t, what)
            if t.numel() != est.numel():
                raise ValueError("SART_prj_diff: buffers differ in size")
        self._check(self._L.ostr_sart_prj_diff(self._h, _p(diff), _p(p), _p(est), _p(norm),
                                               est.numel(), self._stream()))
        return diff

    def bproj_sart(self, diff, a0, a1, img, lam, prior=None, beta=0.0):
        """In place: img = max(0, img + lam (SBP(diff)/sens [+ beta prior/sens]))."""
        _f32(img, "SART image")
        _f32(diff, "diff")
        if prior is None:
            self._check(self._L.ostr_bproj_sart(self._h, _p(img), _p(diff), a0, a1, lam,
                                                self._stream()))
        else:
            _f32(prior, "prior")
            self._check(self._L.ostr_bproj_sart_r(self._h, _p(img), _p(diff), _p(prior), a0, a1,
                                                  lam, beta, self._stream()))
        return img

    def prior_sart(self, img, delta, prior=None):
        """prior += sum_n (1/dist_n) d_n / sqrt(1 + (d_n/delta)^2), d_n = img - img_n."""
        prior = self.volume() if prior is None else prior
        _f32(img, "SART image")
        _f32(prior, "prior")
        self._check(self._L.ostr_prior_sart(self._h, _p(prior), _p(img), delta, self._stream()))
        return prior

    def sart_fproj_diff(self, img, p_rows, a0, a1, diff_rows=None, cos_weight=False):
        """Fused forward + SART_prj_diff: diff_rows = (p - A img) / A 1 over
        [a0, a1), p and diff detector-row-major [S][nby][nbx]."""
        diff_rows = (torch.empty((a1 - a0, self.geom.nby, self.geom.nbx), device=self.device)
                     if diff_rows is None else diff_rows)
        for t, what in ((img, "SART image"), (p_rows, "p"), (diff_rows, "diff")):
            _f32(t, what)
        self._check(self._L.ostr_sart_fproj_diff(self._h, _p(img), _p(p_rows), _p(diff_rows),
                                                 a0, a1, int(cos_weight), self._stream()))
        return diff_rows

    def sart_update_fused(self, x_in, x_out, diff_rows, a0, a1, lam, beta=0.0, delta=1e-3):
        """x_out = SART update of x_in from diff_rows (+ Huber prior when beta != 0)."""
        for t, what in ((x_in, "x_in"), (x_out, "x_out"), (diff_rows, "diff")):
            _f32(t, what)
        self._check(self._L.ostr_sart_update_fused(self._h, _p(x_in), _p(x_out), _p(diff_rows),
                                                   a0, a1, lam, beta, delta, self._stream()))
        return x_out

    def fproj_4r(self, img, a0, a1, prj=None, prj_4r=None, keep_4r=False):
        """2x2 supersampled forward: prj [S][nby][nbx] = mean of the 4 sub-bin ray
        sums (no cos); prj_4r [S][2 nby][2 nbx] filled when given or keep_4r."""
        g = self.geom
        prj = torch.empty((a1 - a0, g.nby, g.nbx), device=self.device) if prj is None else prj
        if prj_4r is None and keep_4r:
            prj_4r = torch.empty((a1 - a0, 2 * g.nby, 2 * g.nbx), device=self.device)
        _f32(img, "image")
        _f32(prj, "projection")
        if prj_4r is not None:
            _f32(prj_4r, "sub-bin projection")
        self._check(self._L.ostr_fproj_4r(self._h, _p(img), _p(prj), _p(prj_4r), a0, a1,
                                          self._stream()))
        return (prj, prj_4r) if prj_4r is not None else prj

    def fprojectCB1_4R_GPU_SART(self, d_objbuf, d_prjbuf, d_4R_prjbuf, angleStart, angleEnd):
        """Tomo_OSTR.cu:2398: d_prjbuf [S][nby][nbx] = merged 2x2 sub-bin sums."""
        return self.fproj_4r(d_objbuf, angleStart, angleEnd, d_prjbuf, d_4R_prjbuf)

    def fbp_filter(self, p, taps, a0, a1, q=None):
        """Cosine-weighted views filtered along detector x with odd-length taps."""
        q = self.sinogram(a1 - a0) if q is None else q
        for t, what in ((p, "p"), (taps, "taps"), (q, "q")):
            _f32(t, what)
        if taps.numel() % 2 == 0:
            raise ValueError("FBP filter taps must have odd length")
        self._check(self._L.ostr_fbp_filter(self._h, _p(q), _p(p), _p(taps), taps.numel(), a0,
                                            a1, self._stream()))
        return q

    def line_integrals(self, y, b, p=None):
        """p = log(b / max(y, 1))."""
        p = torch.empty_like(y) if p is None else p
        for t in (p, y, b):
            _f32(t, "line-integral operand")
            if t.numel() != y.numel():
                raise ValueError("line_integrals: buffers differ in size")
        self._check(self._L.ostr_line_integrals(self._h, _p(p), _p(y), _p(b), y.numel(),
                                                self._stream()))
        return p

    # -- ReconDBT SART / ML steps (ReconDBT/SART_dbt.m, ML_dbt.m) -----------------
    def rdbt_sart_residual(self, y, p, est, l, thresh):
        """y = (p - est) / l where l > thresh; elsewhere y is kept."""
        for t in (y, p, est, l):
            _f32(t, "SART_dbt operand")
            if t.numel() != y.numel():
                raise ValueError("SART_dbt residual: buffers differ in size")
        self._check(self._L.ostr_rdbt_sart_residual(self._h, _p(y), _p(p), _p(est), _p(l),
                                                    thresh, y.numel(), self._stream()))
        return y

    def rdbt_sart_update(self, x, num, den, step, clamp=False):
        """x += step * num / den (NaN -> 0); clamp: then negatives -> 0."""
        for t in (x, num, den):
            _f32(t, "SART_dbt operand")
            if t.numel() != x.numel():
                raise ValueError("SART_dbt update: buffers differ in size")
        self._check(self._L.ostr_rdbt_sart_update(self._h, _p(x), _p(num), _p(den), step,
                                                  int(clamp), x.numel(), self._stream()))
        return x

    def rdbt_ml_terms(self, ydif, wdrr, drr, counts, i0=None, i0_scalar=0.0):
        """yb = I0 exp(-drr): ydif = yb - counts, wdrr = drr * yb."""
        for t in (ydif, wdrr, drr, counts) + ((i0,) if i0 is not None else ()):
            _f32(t, "ML_dbt operand")
            if t.numel() != drr.numel():
                raise ValueError("ML_dbt terms: buffers differ in size")
        self._check(self._L.ostr_rdbt_ml_terms(self._h, _p(ydif), _p(wdrr), _p(drr),
                                               _p(counts), _p(i0), i0_scalar, drr.numel(),
                                               self._stream()))
        return ydif, wdrr

    def rdbt_ml_update(self, x, upd, num, den, step):
        """upd = num / den where den != 0; x = max(0, x + x upd step), NaN -> 0."""
        for t in (x, upd, num, den):
            _f32(t, "ML_dbt operand")
            if t.numel() != x.numel():
                raise ValueError("ML_dbt update: buffers differ in size")
        self._check(self._L.ostr_rdbt_ml_update(self._h, _p(x), _p(upd), _p(num), _p(den),
                                                step, x.numel(), self._stream()))
        return x

    # reference names (Tomo_OSTR.cu:2249-2560)
    def fprojectCB_1R_GPU_SART(self, estbuf, prj_est, d_normprj, angleStart, angleEnd):
        """Tomo_OSTR.cu:2278: prj_est = A estbuf, d_normprj = A 1."""
        return self.fproj_sart(estbuf, angleStart, angleEnd, prj_est, d_normprj, False)

    def fprojectCB_1R_GPU_SART_cos(self, estbuf, prj_est, d_normprj, angleStart, angleEnd):
        """Tomo_OSTR.cu:2306: prj_est = (cos t) A estbuf, d_normprj = A 1."""
        return self.fproj_sart(estbuf, angleStart, angleEnd, prj_est, d_normprj, True)

    def SART_prj_diff(self, diff_line, prjbuf, prj_est, normprj, angleStart, angleEnd):
        """Tomo_OSTR.cu:2537 over the subset's bins [angleEnd - angleStart][nbx][nby]."""
        n = (angleEnd - angleStart) * self.geom.nbx * self.geom.nby
        if diff_line.numel() != n:
            raise ValueError("SART_prj_diff: diff_line does not hold the subset's bins")
        return self.sart_diff(prjbuf, prj_est, normprj, diff_line)

    def bprojectCB_4B_GPU_SART(self, d_objbuf, d_prjbuf, angleStart, angleEnd, lambda_):
        """Tomo_OSTR.cu:2460."""
        return self.bproj_sart(d_prjbuf, angleStart, angleEnd, d_objbuf, lambda_)

    def bprojectCB_4B_GPU_SART_R(self, d_objbuf, d_prjbuf, d_prior, angleStart, angleEnd,
                                 lambda_, beta):
        """Tomo_OSTR.cu:2497 (the prior term enters with the reference's + sign)."""
        return self.bproj_sart(d_prjbuf, angleStart, angleEnd, d_objbuf, lambda_, d_prior, beta)

    def prior_GPU_SART(self, d_prior, d_est, delta):
        """Tomo_OSTR.cu:2249 (accumulates into d_prior; delta == 0 is an error)."""
        return self.prior_sart(d_est, delta, d_prior)


class OSTR:
    """OSTR driver (Tomo_OSTR_gold.cpp:301-918) on device tensors.

    y, b, r: measured counts, blank scan, scatter as [V][nbx][nby] float32 in
    natural view order (r may be None = 0). x0: initial image (None = zeros,
    the reference's effective start, Tomo_OSTR_gold.cpp:578).
    """

    def __init__(self, geom: Geometry, angles, y, b, r=None, cfg: Optional[OSTRConfig] = None,
                 x0=None, img_dtype=torch.float32, prj_dtype=torch.float32, device=None):
        self.cfg = cfg or OSTRConfig()
        self.geom = geom
        nv = len(angles)
        M = self.cfg.n_subsets
        if M <= 0 or nv % M:
            raise ValueError(f"ANGLES % subset_number != 0 ({nv} % {M})")
        if geom.dx != geom.dy:
            raise ValueError("voxel size in x must equal voxel size in y (prior weights)")
        if self.cfg.prior != "q2" and self.cfg.delta == 0:
            raise ValueError("delta cannot be zero")
        if self.cfg.backprojector not in BACKPROJECTORS:
            raise ValueError(f"unknown backprojector {self.cfg.backprojector!r}")
        self.matched = self.cfg.backprojector == "matched"
        # fused: two launches per subset (forward + h, then backprojection +
        # prior + update), image and h in F32 / BF16 / F16 storage
        self.fused = self.cfg.fused
        if not self.fused and (img_dtype != torch.float32 or prj_dtype != torch.float32):
            raise ValueError("BF16/FP16 storage is implemented by the fused kernels only")
        self.sys = DBTSystem(geom, angles, device=device)
        self.dev = self.sys.device
        self.M, self.vs = M, nv // M
        self.img_dtype, self.prj_dtype = img_dtype, prj_dtype
        self.rho = geom.dz / geom.dx

        def dev32(a):
            if a is None:
                return None
            t = torch.as_tensor(a)
            return t.to(device=self.dev, dtype=torch.float32).contiguous()

        self.y_nat, self.b_nat, self.r_nat = dev32(y), dev32(b), dev32(r)
        if x0 is None:
            self.x = self.sys.volume(img_dtype)
        elif isinstance(x0, str) and x0 == "bp":
            self.x = self.sys.bp_init(y=self.y_nat, b=self.b_nat).to(img_dtype).contiguous()
        elif isinstance(x0, str) and x0 == "fbp":
            # FDK-style FBP (Hann window, the fbp_dbt.m default), negatives set to 0
            from .fbp import fbp_dbt
            p = self.sys.line_integrals(self.y_nat, self.b_nat)
            img = fbp_dbt(geom, angles, p, "hann", device=self.dev)
            self.x = self.sys.extrapolate(torch.empty_like(img), img, img, 0.0).to(img_dtype)
        elif isinstance(x0, str) and x0 in ("sart", "ml"):
            # ReconDBT's iterative initialisers from the BP image, with the
            # example's settings (dbt_recon_circular_example.m:92, :100)
            from .recondbt import ml_dbt, sart_dbt
            xbp = self.sys.bp_init(y=self.y_nat, b=self.b_nat)
            if x0 == "sart":
                p = self.sys.line_integrals(self.y_nat, self.b_nat)
                xi = sart_dbt(geom, angles, p, xbp, 2, 0.15, device=self.dev)
            else:
                xi = ml_dbt(geom, angles, self.y_nat, xbp, self.b_nat, 3, 2.0, device=self.dev)
            self.x = xi.to(img_dtype).contiguous()
        else:
            # a copy: the iteration works in this buffer, never in the caller's
            self.x = dev32(x0).to(img_dtype).clone(memory_format=torch.contiguous_format)
        self.iteration = 0
        self.z_prev = None  # momentum state (see _momentum_step)
        self._mstate = None
        self._graph = None
        self._precomputed = False
        # optional per-kernel timing: name -> list of (start, end) torch.cuda.Events
        self.events = None

    def reset(self, x0: torch.Tensor):
        """Restart from image x0 (copied, in the image storage type): iteration
        count, momentum state and the recorded graph are cleared; the
        precompute is kept."""
        self.x = x0.to(device=self.dev, dtype=self.img_dtype).clone(
            memory_format=torch.contiguous_format)
        if hasattr(self, "x_next"):
            self.x_next = torch.empty_like(self.x)
        self.iteration = 0
        self.z_prev = None
        self._mstate = None
        self._xk = None
        self._graph = None

    def _timed(self, name, fn, *args, **kw):
        if self.events is None:
            return fn(*args, **kw)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        out = fn(*args, **kw)
        e1.record()
        self.events.setdefault(name, []).append((e0, e1))
        return out

    def kernel_times_ms(self):
        """Average duration (ms) per timed launch name; needs a synchronize first."""
        if not self.events:
            return {}
        return {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in self.events.items()}

    def precompute(self):
        """gamma, yry, d = SBP(gamma*yry) (natural order), then the OS regroup."""
        s, cfg = self.sys, self.cfg
        nv = self.sys.n_views
        if self.matched:  # d = G^T(yry * G 1)
            gam = s.fproj(s.volume(fill=1.0), 0, nv, cos_weight=True)
            yry = s.compute_yry(torch.empty_like(self.y_nat), self.y_nat, self.r_nat)
            s.compute_gamma_yry(gam, yry, gam)
            if cfg.d_fp64:  # each view's term in FP32, the sum over views in FP64
                acc = torch.zeros(self.geom.vol_shape, dtype=torch.float64, device=self.dev)
                tmp = s.volume()
                for v in range(nv):
                    tmp.zero_()
                    s.accum_f64(acc, s.bproj_matched(gam[v:v + 1], v, v + 1, img=tmp,
                                                     cos_weight=True))
                self.d = s.f64_to_f32(acc)
                del acc, tmp
            else:
                self.d = s.bproj_matched(gam, 0, nv, cos_weight=True)
            del gam, yry
        elif self.fused or cfg.d_fp64:
            self.d = s.precompute_d(self.y_nat, self.r_nat, cfg.gamma_mode, fp64=cfg.d_fp64)
        else:
            gam = s.gamma(0, nv, cfg.gamma_mode)
            yry = s.compute_yry(torch.empty_like(self.y_nat), self.y_nat, self.r_nat)
            s.compute_gamma_yry(gam, yry, gam)
            self.d = s.bprojectCB_GPU_SBP(s.volume(), gam, 0, nv)
            del gam, yry
        # the fused kernels read subset sinograms detector-row-major
        rows = self.fused
        self.y = s.regroup_prj(self.y_nat, self.M, to_rows=rows)
        self.b = s.regroup_prj(self.b_nat, self.M, to_rows=rows)
        self.r = None if self.r_nat is None else s.regroup_prj(self.r_nat, self.M, to_rows=rows)
        s.set_index(regroup_index(nv, self.M))
        if self.fused:
            self.h = torch.zeros((self.vs, self.geom.nby, self.geom.nbx), dtype=self.prj_dtype,
                                 device=self.dev)
            self.x_next = torch.empty_like(self.x)
        else:
            self.line = s.sinogram(self.vs)
            self.h = s.sinogram(self.vs)
            self.capL = s.volume()
            self.RD, self.RDD = s.volume(), s.volume()
        self._precomputed = True

    def _subset(self, m: int):
        s, cfg = self.sys, self.cfg
        a0, a1 = m * self.vs, (m + 1) * self.vs
        sl = slice(a0, a1)
        r_sub = None if self.r is None else self.r[sl]
        if self.fused:
            self._timed("fproj_h", s.fproj_h, self.x, self.y[sl], self.b[sl], r_sub, a0, a1,
                        hsub=self.h)
            if self.matched:
                self._timed("update_fused_matched", s.update_fused_matched, self.x, self.x_next,
                            self.d, self.h, a0, a1, cfg.beta, float(self.M), cfg.prior,
                            cfg.delta if cfg.prior != "q2" else 1.0, self.rho, cfg.guard)
            else:
                self._timed("update_fused", s.update_fused, self.x, self.x_next, self.d, self.h,
                            a0, a1, cfg.beta, float(self.M), cfg.prior, cfg.delta, self.rho,
                            cfg.guard)
            self.x, self.x_next = self.x_next, self.x
        else:
            s.fprojectCB_1R_GPU_OSTR_cos(self.x, self.line, a0, a1)
            s.compute_h(self.h, self.y[sl], self.b[sl], self.line, r_sub)
            self.capL.zero_()
            if self.matched:
                s.bproj_matched(self.h, a0, a1, img=self.capL, cos_weight=True)
            else:
                s.bprojectCB_GPU_SBP(self.capL, self.h, a0, a1)
            self.RDD.zero_()
            self.RD.zero_()
            s.prior(self.x, cfg.prior, cfg.delta if cfg.prior != "q2" else 1.0, self.rho,
                    rdd=self.RDD, rd=self.RD)
            s.update_est(self.x, self.capL, self.RD, self.d, self.RDD, cfg.beta, float(self.M),
                         cfg.guard)

    @property
    def estimate(self) -> torch.Tensor:
        """The current image: x, or with momentum the last OS-pass result z
        (x is then the extrapolated point the next pass starts from)."""
        return self.z_prev if getattr(self, "z_prev", None) is not None else self.x

    @property
    def restarts(self) -> int:
        """Momentum restarts so far (reads the device state)."""
        return 0 if self._mstate is None else int(self._mstate[2].item())

    def _momentum_step(self):
        """After an OS pass from x_k (self.x now holds z_{k+1}):
        t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2,
        x_{k+1} = max(0, z_{k+1} + (t_k - 1)/t_{k+1} (z_{k+1} - z_k)),
        with a reset (t = 1, x = z) when <x_k - z_{k+1}, z_{k+1} - z_k> > 0
        (gradient restart, O'Donoghue & Candes 2015). The test and t live on
        the device: no host round trip."""
        z = self.x
        if self.z_prev is None:  # first pass: z_0 = x_0
            self.z_prev, self._spare = self._xk.clone(), torch.empty_like(z)
            self._mstate = torch.empty(4, dtype=torch.float64, device=self.dev)
            self.sys.momentum_init(self._mstate)
        xn = self._spare
        self.sys.momentum_step(xn, z, self.z_prev, self._xk, self._mstate)
        # rotate: z_prev <- z_{k+1}, x <- x_{k+1}, spare <- z_k
        self._spare, self.z_prev, self.x = self.z_prev, z, xn

    def _one_iteration(self):
        # x_k, the point this pass starts from: with one fused subset it is
        # still in the input buffer of the update after the pass
        keep_in_place = self.fused and self.M == 1
        if self.cfg.momentum and not keep_in_place:
            if getattr(self, "_xk", None) is None or self._xk is getattr(self, "x_next", None):
                self._xk = torch.empty_like(self.x)
            self._xk.copy_(self.x)
        for m in range(self.M):
            self._subset(m)
        self.iteration += 1
        if self.cfg.momentum:
            if keep_in_place:
                self._xk = self.x_next
            self._momentum_step()

    def _buffers(self):
        return tuple(t.data_ptr() if torch.is_tensor(t) else None
                     for t in (self.x, getattr(self, "x_next", None), self.z_prev,
                               getattr(self, "_spare", None)))

    def _capture(self, max_iters: int = 12):
        """Record iterations into a HIP graph until the image buffers are back
        where they started (the momentum rotation has period 3, the fused
        double buffer 2 per subset); returns the number recorded."""
        start, it0 = self._buffers(), self.iteration
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for k in range(1, max_iters + 1):
                self._one_iteration()
                if self._buffers() == start:
                    break
            else:
                raise RuntimeError("image buffers do not return to their start")
        self.iteration = it0
        self._graph, self._graph_iters, self._graph_start = g, k, start
        return k

    def iterate(self, n_iters: int = 1, callback: Optional[Callable] = None,
                graph: bool = False):
        """Run n_iters full iterations (all subsets); callback(it, estimate)
        after each. graph=True replays the iterations from a HIP graph recorded
        on first use (one eager iteration first, so every work buffer exists;
        no callback)."""
        if not self._precomputed:
            self.precompute()
        if graph and callback is None and n_iters > 0:
            if self._graph is None:
                self._one_iteration()
                n_iters -= 1
                torch.cuda.synchronize(self.dev)
                self._capture()
            # the graph starts from the buffer assignment it was recorded with:
            # eager iterations first until the rotation is back there
            while n_iters > 0 and self._buffers() != self._graph_start:
                self._one_iteration()
                n_iters -= 1
            k = self._graph_iters
            if self._buffers() == self._graph_start:
                for _ in range(n_iters // k):
                    self._graph.replay()
                    self.iteration += k
                n_iters %= k
        for _ in range(n_iters):
            self._one_iteration()
            if callback is not None:
                callback(self.iteration, self.estimate)
        return self.estimate

    def run(self):
        return self.iterate(self.cfg.n_iters)

    def objective(self, x=None) -> float:
        """FP64 penalized negative log-likelihood of x (default: the current image)."""
        cfg = self.cfg
        data, pen = self.sys.objective(self.estimate if x is None else x, self.y_nat, self.b_nat,
                                       self.r_nat, cfg.prior,
                                       cfg.delta if cfg.prior != "q2" else 1.0, self.rho)
        return data + cfg.beta * pen


def save_raw(path: str, vol: torch.Tensor):
    """Raw float32 [z][y][x] (the reference's output format, Tomo_OSTR_gold.cpp:65-91)."""
    vol.detach().to("cpu", torch.float32).contiguous().numpy().tofile(path)


def load_raw(path: str, geom: Geometry) -> np.ndarray:
    """Inverse of save_raw; also the `ini_guess` reader (Tomo_OSTR_gold.cpp:581)."""
    a = np.fromfile(path, dtype=np.float32)
    if a.size != geom.n_vox:
        raise ValueError(f"{path}: {a.size} floats, expected {geom.n_vox}")
    return a.reshape(geom.vol_shape)
