"""Batched OSTR: several reconstructions of uneven geometry, one launch per kernel.

The reference runs one geometry per process, held in __constant__ state
(Tomo_OSTR.cu:81-115), so it can only run cases one after another. Here each
case keeps its own libostr_hip handle and buffers (an `OSTR` driver in fused
mode), and every batch step runs the current subset of ALL cases with two
launches: one batched forward projection + h (ostr_batch_fproj_h) and one
batched fused update (ostr_batch_update_fused). Cases may differ in volume
size, thickness, view count and subset count; case c advances through its own
subsets (step k runs its subset k mod M_c).

Per case the images are bit-identical to running the same case alone with the
single-case fused kernels: every bin and voxel is computed by the same device
code in the same order, and the tile shape only decides which work-group
computes it. The tile and grid shapes are `TileCfg`, chosen by
`newrecon_amd.autotune`.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import torch

from .geometry import POTENTIALS
from .ostr import _DT, OSTR, OSTRError, lib


class TileCfg(ctypes.Structure):
    """ostr_tile_cfg: tile shapes and grid sizing of the batched launches."""

    _fields_ = [("fwd_rows", ctypes.c_int32), ("fwd_groups_per_cu", ctypes.c_int32),
                ("upd_rows", ctypes.c_int32), ("upd_waves", ctypes.c_int32),
                ("upd_groups_per_cu", ctypes.c_int32)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}

    @classmethod
    def from_dict(cls, d):
        return cls(**{f: int(d[f]) for f, _ in cls._fields_})

    def __repr__(self):
        return f"TileCfg({self.as_dict()})"


# the single-case kernels' shapes: 64x4 bins, 64x(8x4) voxels, one group per tile
DEFAULT_TILE = dict(fwd_rows=4, fwd_groups_per_cu=0, upd_rows=8, upd_waves=4,
                    upd_groups_per_cu=0)


class BatchCaseC(ctypes.Structure):
    """ostr_batch_case."""

    _fields_ = [("h", ctypes.c_void_p), ("a0", ctypes.c_int32), ("a1", ctypes.c_int32),
                ("x_in", ctypes.c_void_p), ("x_out", ctypes.c_void_p),
                ("d", ctypes.c_void_p), ("y", ctypes.c_void_p), ("b", ctypes.c_void_p),
                ("r", ctypes.c_void_p), ("hsub", ctypes.c_void_p),
                ("beta", ctypes.c_float), ("capl_scale", ctypes.c_float),
                ("delta", ctypes.c_float), ("z_xy_ratio", ctypes.c_float),
                ("guard", ctypes.c_int32)]


_batch_sig_done = False


def batch_lib():
    global _batch_sig_done
    L = lib()
    if not _batch_sig_done:
        vp, i32 = ctypes.c_void_p, ctypes.c_int32
        L.ostr_batch_create.argtypes = [i32, i32, ctypes.POINTER(vp)]
        L.ostr_batch_create.restype = ctypes.c_int
        L.ostr_batch_destroy.argtypes = [vp]
        L.ostr_batch_destroy.restype = None
        L.ostr_batch_last_error.argtypes = [vp]
        L.ostr_batch_last_error.restype = ctypes.c_char_p
        for name, extra in (("ostr_batch_fproj_h", [i32]), ("ostr_batch_update_fused", [i32, i32])):
            fn = getattr(L, name)
            fn.argtypes = [vp, ctypes.POINTER(BatchCaseC), i32] + extra + [
                ctypes.POINTER(TileCfg), vp]
            fn.restype = ctypes.c_int
        _batch_sig_done = True
    return L


class BatchOSTR:
    """Runs the fused OSTR subsets of several `OSTR` cases in batched launches.

    All cases must be fused, on one device, with float32 images and one
    sinogram dtype and prior potential. `tile`: a TileCfg, a dict, or None
    for the single-case shapes.
    """

    def __init__(self, cases: Sequence[OSTR], tile=None):
        if not cases:
            raise ValueError("empty batch")
        self.cases: List[OSTR] = list(cases)
        c0 = self.cases[0]
        for c in self.cases:
            if not c.fused:
                raise ValueError("batched cases must use the fused kernels")
            if c.img_dtype != torch.float32:
                raise ValueError("batched cases take float32 images")
            if c.prj_dtype != c0.prj_dtype or c.cfg.prior != c0.cfg.prior:
                raise ValueError("batched cases share one sinogram dtype and prior")
            if c.dev != c0.dev:
                raise ValueError("batched cases must be on one device")
            if c.vs > 64:
                raise ValueError("at most 64 views per subset")
        self.dev = c0.dev
        self.prj_dtype = c0.prj_dtype
        self.potential = POTENTIALS[c0.cfg.prior]
        self._L = batch_lib()
        h = ctypes.c_void_p()
        rc = self._L.ostr_batch_create(self.dev.index if self.dev.index is not None
                                       else torch.cuda.current_device(), len(self.cases),
                                       ctypes.byref(h))
        if rc != 0:
            raise OSTRError(rc, self._L.ostr_batch_last_error(None).decode())
        self._b = h
        self.tile = self._tile(tile)
        self.pos = [0] * len(self.cases)  # next subset of each case
        self.steps = 0
        for c in self.cases:
            if not c._precomputed:
                c.precompute()
        self._arr = (BatchCaseC * len(self.cases))()

    @staticmethod
    def _tile(tile) -> TileCfg:
        if tile is None:
            return TileCfg.from_dict(DEFAULT_TILE)
        if isinstance(tile, dict):
            return TileCfg.from_dict(tile)
        return tile

    def close(self):
        if getattr(self, "_b", None):
            self._L.ostr_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise OSTRError(rc, self._L.ostr_batch_last_error(self._b).decode())

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    # -- one batched subset -------------------------------------------------
    def _fill(self):
        for k, c in enumerate(self.cases):
            m = self.pos[k]
            a0, a1 = m * c.vs, (m + 1) * c.vs
            e = self._arr[k]
            e.h = c.sys._h.value
            e.a0, e.a1 = a0, a1
            e.x_in, e.x_out = c.x.data_ptr(), c.x_next.data_ptr()
            e.d = c.d.data_ptr()
            e.y = c.y[a0:a1].data_ptr()
            e.b = c.b[a0:a1].data_ptr()
            e.r = None if c.r is None else c.r[a0:a1].data_ptr()
            e.hsub = c.h.data_ptr()
            e.beta, e.capl_scale = c.cfg.beta, float(c.M)
            e.delta, e.z_xy_ratio = c.cfg.delta, c.rho
            e.guard = int(c.cfg.guard)

    def fproj_h(self, tile: Optional[TileCfg] = None):
        """Batched forward + h of every case's current subset (no state change)."""
        self._fill()
        self._check(self._L.ostr_batch_fproj_h(self._b, self._arr, len(self.cases),
                                               _dt_code(self.prj_dtype),
                                               ctypes.byref(tile or self.tile), self._stream()))

    def update(self, tile: Optional[TileCfg] = None):
        """Batched fused update into each case's x_next (no swap)."""
        self._fill()
        self._check(self._L.ostr_batch_update_fused(self._b, self._arr, len(self.cases),
                                                    _dt_code(self.prj_dtype), self.potential,
                                                    ctypes.byref(tile or self.tile),
                                                    self._stream()))

    def _advance(self):
        for k, c in enumerate(self.cases):
            c.x, c.x_next = c.x_next, c.x
            self.pos[k] += 1
            if self.pos[k] == c.M:
                self.pos[k] = 0
                c.iteration += 1
        self.steps += 1

    def step(self, n: int = 1):
        """n batch steps: each runs the next subset of every case."""
        for _ in range(n):
            self.fproj_h()
            self.update()
            self._advance()

    def images(self):
        return [c.x for c in self.cases]

    def signature(self) -> dict:
        """What the tile choice depends on (the autotuner's cache key)."""
        props = torch.cuda.get_device_properties(self.dev)
        return {
            "device": getattr(props, "gcnArchName", props.name),
            "cus": props.multi_processor_count,
            "prj_dtype": str(self.prj_dtype).replace("torch.", ""),
            "potential": self.potential,
            "cases": [[c.geom.nx, c.geom.ny, c.geom.nz, c.geom.nbx, c.geom.nby, c.vs]
                      for c in self.cases],
        }


def _dt_code(dtype) -> int:
    return _DT[dtype]
