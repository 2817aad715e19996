"""Learned warm start: a 3-D CNN from the one-step backprojected image to an
estimate of the converged OSTR image, trained and run on the GPU.

The reference has no counterpart on its OSTR path (its learned model is a
Keras network in recon1.ipynb, outside the reconstruction); this is new.

Network (3 resolution levels, 16 / 32 / 64 channels, 3x3x3 convolutions):

    x0 = [img * scale, 0 x 7]                               level 0 (full)
    e1a = relu(conv 8->16),  e1b = relu(conv 16->16)
    e2a = relu(conv 16->32 (pool e1b)),  e2b = relu(conv 32->32)   level 1
    e3a = relu(conv 32->64 (pool e2b)),  e3b = relu(conv 64->64)   level 2
    d2 = relu(conv 96->32 [up e3b, e2b])                      level 1
    d1 = relu(conv 48->16 [up d2, e1b])                       level 0
    out = conv 16->16 (d1);  pred = img + out[..., 0] / scale

Every operation runs in libostr_hip.so (include/ostr_cnn.h): BF16 activations
and weights on MFMA with FP32 accumulation, FP32 master weights, biases and
Adam moments; PyTorch only holds the device buffers. Sizes must be multiples
of 4 (two poolings); `pad4` pads a volume.

As a warm start the network's residual is low-passed in-plane (a Gaussian of
WARM_SMOOTH_SIGMA voxels, ostr_cnn_output_smooth) before it is added to the
backprojection. The warm start is judged by the plain method: `warm_started`
runs one plain OSTR pass from the network's image and one from the
backprojection and accepts the network's image only if the FP64 objective
after its pass is the lower; otherwise the reconstruction starts from the
backprojection (the fallback). Plain OSTR iterations always follow.
"""
from __future__ import annotations

import ctypes
import json
import os
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from .ostr import OSTRError, lib

# name, cin, cout, relu (level: e1*, d1, out 0; e2*, d2 1; e3* 2)
LAYERS: List[Tuple[str, int, int, bool]] = [
    ("e1a", 8, 16, True), ("e1b", 16, 16, True),
    ("e2a", 16, 32, True), ("e2b", 32, 32, True),
    ("e3a", 32, 64, True), ("e3b", 64, 64, True),
    ("d2", 96, 32, True), ("d1", 48, 16, True),
    ("out", 16, 16, False),
]
SCALE = 20.0  # attenuation (1/mm) to O(1) network units
# In-plane Gaussian (sigma in voxels) applied to the network's residual when it
# serves as a warm start: its finest in-plane detail is not what the iteration
# needs (with it the time to 1e-3 is longer than from the backprojection)
WARM_SMOOTH_SIGMA = 1.0


def gaussian_taps(sigma: float):
    """Normalised taps exp(-k^2 / (2 sigma^2)), k = -r..r, r = min(4, ceil(3
    sigma)), as float32; sigma <= 0 gives the identity (r = 0)."""
    if sigma <= 0:
        return np.ones(1, np.float32), 0
    r = int(min(4, np.ceil(3.0 * sigma)))
    k = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * (k / sigma) ** 2)
    return (w / w.sum()).astype(np.float32), r

_sig_done = False


def cnn_lib():
    global _sig_done
    L = lib()
    if not _sig_done:
        vp, i32, i64, f32, sz = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float,
                                 ctypes.c_size_t)
        sig = {
            "ostr_cnn_pack": (ctypes.c_int, [vp, i32, i32, vp, vp, vp]),
            "ostr_cnn_conv3d": (ctypes.c_int, [i32, i32, vp, i32, i32, vp, vp, vp, i32, i32, vp,
                                               i32, i32, i32, i32, i32, i32, vp]),
            "ostr_cnn_wgrad_workspace": (sz, [i32, i32, i32, i32, i32, i32]),
            "ostr_cnn_wgrad": (ctypes.c_int, [i32, i32, vp, i32, i32, vp, i32, i32, i32, i32,
                                              i32, vp, vp, vp, sz, i32, vp]),
            "ostr_cnn_pool2": (ctypes.c_int, [vp, i32, i32, vp, i32, i32, i32, i32, i32, i32,
                                              vp]),
            "ostr_cnn_pool2_bwd": (ctypes.c_int, [vp, i32, i32, vp, i32, i32, vp, i32, i32, i32,
                                                  i32, i32, i32, i32, vp]),
            "ostr_cnn_up2": (ctypes.c_int, [vp, i32, i32, vp, i32, i32, i32, i32, i32, i32, vp]),
            "ostr_cnn_up2_bwd": (ctypes.c_int, [vp, i32, i32, vp, i32, i32, vp, i32, i32, i32,
                                                i32, i32, i32, vp]),
            "ostr_cnn_input": (ctypes.c_int, [vp, i64, f32, vp, vp]),
            "ostr_cnn_output": (ctypes.c_int, [vp, i32, vp, i64, f32, i32, vp, vp]),
            "ostr_cnn_output_smooth": (ctypes.c_int, [vp, i32, vp, i32, i32, i32, f32, vp, i32,
                                                      i32, vp, vp]),
            "ostr_cnn_mse_workspace": (i64, [i64]),
            "ostr_cnn_mse": (ctypes.c_int, [vp, vp, i64, f32, vp, i32, vp, i64,
                                            ctypes.POINTER(ctypes.c_double), vp]),
            "ostr_cnn_adam": (ctypes.c_int, [vp, vp, vp, vp, i64, f32, f32, f32, f32, i32, vp]),
            "ostr_cnn_kpad": (i32, [i32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.argtypes, fn.restype = args, res
        _sig_done = True
    return L


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def pad4(vol: torch.Tensor) -> torch.Tensor:
    """Zero-pad a [nz, ny, nx] volume up to multiples of 4."""
    nz, ny, nx = vol.shape
    pz, py, px = (-nz) % 4, (-ny) % 4, (-nx) % 4
    if pz == py == px == 0:
        return vol.contiguous()
    out = torch.zeros((nz + pz, ny + py, nx + px), dtype=vol.dtype, device=vol.device)
    out[:nz, :ny, :nx] = vol
    return out


class WarmStartNet:
    """Parameters, Adam state and the forward / backward passes on one device."""

    def __init__(self, device=None, seed: int = 0):
        self.dev = torch.device(device if device is not None else "cuda")
        self._L = cnn_lib()
        self.offsets: Dict[str, Tuple[int, int, int]] = {}
        n = 0
        for name, cin, cout, _ in LAYERS:
            self.offsets[name] = (n, n + cout * 27 * cin, n + cout * 27 * cin + cout)
            n += cout * 27 * cin + cout
        self.n_params = n
        self.params = torch.zeros(n, dtype=torch.float32, device=self.dev)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step_count = 0
        self.init_params(seed)
        self.packed = {}
        for name, cin, cout, _ in LAYERS:
            kf, kd = self._L.ostr_cnn_kpad(cin), self._L.ostr_cnn_kpad(cout)
            self.packed[name] = (torch.zeros(cout * kf, dtype=torch.int16, device=self.dev),
                                 torch.zeros(cin * kd, dtype=torch.int16, device=self.dev))
        self._bufs = {}
        self.pack()

    # -- parameters ---------------------------------------------------------
    def weight(self, name, t=None) -> torch.Tensor:
        a, b, _ = self.offsets[name]
        _, cin, cout, _ = next(l for l in LAYERS if l[0] == name)
        return (self.params if t is None else t)[a:b].view(cout, 27, cin)

    def bias(self, name, t=None) -> torch.Tensor:
        _, b, c = self.offsets[name]
        return (self.params if t is None else t)[b:c]

    def init_params(self, seed: int):
        """He-normal weights (output layer x0.1), zero biases; host RNG."""
        rng = np.random.default_rng(seed)
        host = np.zeros(self.n_params, np.float32)
        for name, cin, cout, _ in LAYERS:
            a, b, _ = self.offsets[name]
            std = np.sqrt(2.0 / (27 * cin)) * (0.1 if name == "out" else 1.0)
            w = rng.standard_normal((cout, 27, cin)) * std
            if name == "e1a":
                w[:, :, 1:] = 0.0  # input channels 1..7 are always zero
            host[a:b] = w.ravel()
        self.params.copy_(torch.from_numpy(host))

    def set_params(self, host: Dict[str, Tuple[np.ndarray, np.ndarray]]):
        flat = np.zeros(self.n_params, np.float32)
        for name, _, _, _ in LAYERS:
            a, b, c = self.offsets[name]
            flat[a:b] = np.asarray(host[name][0], np.float32).ravel()
            flat[b:c] = np.asarray(host[name][1], np.float32)
        self.params.copy_(torch.from_numpy(flat))
        self.pack()

    def host_params(self) -> Dict[str, Tuple[np.ndarray, np.ndarray]]:
        p = self.params.cpu().numpy()
        out = {}
        for name, cin, cout, _ in LAYERS:
            a, b, c = self.offsets[name]
            out[name] = (p[a:b].reshape(cout, 27, cin).copy(), p[b:c].copy())
        return out

    def save(self, path: str, meta: Optional[dict] = None):
        np.savez_compressed(path, params=self.params.cpu().numpy(),
                            meta=json.dumps(meta or {}))

    def load(self, path: str) -> dict:
        data = np.load(path, allow_pickle=False)
        p = data["params"]
        if p.shape != (self.n_params,):
            raise ValueError(f"{path}: {p.shape[0]} parameters, network has {self.n_params}")
        self.params.copy_(torch.from_numpy(p))
        self.pack()
        return json.loads(str(data["meta"]))

    def _check(self, rc):
        if rc != 0:
            raise OSTRError(rc, self._L.ostr_last_error(None).decode())

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def pack(self):
        for name, cin, cout, _ in LAYERS:
            wf, wd = self.packed[name]
            self._check(self._L.ostr_cnn_pack(_ptr(self.weight(name)), cin, cout, _ptr(wf),
                                              _ptr(wd), self._stream()))

    # -- buffers --------------------------------------------------------------
    def _buffers(self, shape):
        if shape in self._bufs:
            return self._bufs[shape]
        nz, ny, nx = shape
        if nz % 4 or ny % 4 or nx % 4:
            raise ValueError(f"volume {shape}: sizes must be multiples of 4 (use pad4)")
        lv = [(nz, ny, nx), (nz // 2, ny // 2, nx // 2), (nz // 4, ny // 4, nx // 4)]
        nv = [a * b * c for a, b, c in lv]
        bf = lambda lvl, c: torch.zeros(nv[lvl] * c, dtype=torch.int16, device=self.dev)
        B = {
            "x0": bf(0, 8), "e1a": bf(0, 16), "cat1": bf(0, 48), "p1": bf(1, 16),
            "e2a": bf(1, 32), "cat2": bf(1, 96), "p2": bf(2, 32), "e3a": bf(2, 64),
            "e3b": bf(2, 64), "d2": bf(1, 32), "d1": bf(0, 16), "out": bf(0, 16),
            # gradients
            "dout": bf(0, 16), "dd1": bf(0, 16), "dcat1": bf(0, 48), "dd2": bf(1, 32),
            "dcat2": bf(1, 96), "de3b": bf(2, 64), "de3a": bf(2, 64), "dp2": bf(2, 32),
            "de2a": bf(1, 32), "dp1": bf(1, 16), "de1a": bf(0, 16),
            "pred": torch.zeros(nv[0], dtype=torch.float32, device=self.dev),
        }
        nsplit = 64
        wmax = max(self._L.ostr_cnn_wgrad_workspace(cin, cout, *lv[self._level(name)], nsplit)
                   for name, cin, cout, _ in LAYERS)
        B["work"] = torch.zeros(int(wmax), dtype=torch.uint8, device=self.dev)
        B["nsplit"] = nsplit
        me = int(self._L.ostr_cnn_mse_workspace(nv[0]))
        B["mse"] = torch.zeros(me, dtype=torch.float64, device=self.dev)
        B["levels"] = lv
        self._bufs[shape] = B
        return B

    @staticmethod
    def _level(name):
        return {"e1a": 0, "e1b": 0, "d1": 0, "out": 0, "e2a": 1, "e2b": 1, "d2": 1,
                "e3a": 2, "e3b": 2}[name]

    # -- forward --------------------------------------------------------------
    def _conv(self, name, x, xcs, xco, y, ycs, yco, lvl, B, relu=None, mask=None, mcs=0, mco=0,
              grad=False):
        _, cin, cout, r = next(l for l in LAYERS if l[0] == name)
        wf, wd = self.packed[name]
        if grad:  # data gradient: swap channels, flipped weights, no bias
            cin, cout = cout, cin
            w, bias, relu = wd, None, False
        else:
            w, bias, relu = wf, self.bias(name), (r if relu is None else relu)
        self._check(self._L.ostr_cnn_conv3d(cin, cout, _ptr(x), xcs, xco, _ptr(w), _ptr(bias),
                                            _ptr(y), ycs, yco, _ptr(mask), mcs, mco, int(relu),
                                            *B["levels"][lvl], self._stream()))

    def forward(self, img: torch.Tensor, clamp: bool = False,
                smooth_sigma: float = 0.0) -> torch.Tensor:
        """pred = network(img) for a float32 [nz, ny, nx] volume (multiples of 4);
        smooth_sigma > 0: the residual is low-passed in-plane first."""
        img = img.contiguous()
        B = self._buffers(tuple(img.shape))
        L, st = self._L, self._stream()
        lv = B["levels"]
        n0 = img.numel()
        self._check(L.ostr_cnn_input(_ptr(img), n0, SCALE, _ptr(B["x0"]), st))
        self._conv("e1a", B["x0"], 8, 0, B["e1a"], 16, 0, 0, B)
        self._conv("e1b", B["e1a"], 16, 0, B["cat1"], 48, 32, 0, B)
        self._check(L.ostr_cnn_pool2(_ptr(B["cat1"]), 48, 32, _ptr(B["p1"]), 16, 0, 16,
                                     *lv[0], st))
        self._conv("e2a", B["p1"], 16, 0, B["e2a"], 32, 0, 1, B)
        self._conv("e2b", B["e2a"], 32, 0, B["cat2"], 96, 64, 1, B)
        self._check(L.ostr_cnn_pool2(_ptr(B["cat2"]), 96, 64, _ptr(B["p2"]), 32, 0, 32,
                                     *lv[1], st))
        self._conv("e3a", B["p2"], 32, 0, B["e3a"], 64, 0, 2, B)
        self._conv("e3b", B["e3a"], 64, 0, B["e3b"], 64, 0, 2, B)
        self._check(L.ostr_cnn_up2(_ptr(B["e3b"]), 64, 0, _ptr(B["cat2"]), 96, 0, 64, *lv[1],
                                   st))
        self._conv("d2", B["cat2"], 96, 0, B["d2"], 32, 0, 1, B)
        self._check(L.ostr_cnn_up2(_ptr(B["d2"]), 32, 0, _ptr(B["cat1"]), 48, 0, 32, *lv[0],
                                   st))
        self._conv("d1", B["cat1"], 48, 0, B["d1"], 16, 0, 0, B)
        self._conv("out", B["d1"], 16, 0, B["out"], 16, 0, 0, B)
        if smooth_sigma > 0:
            taps, r = gaussian_taps(smooth_sigma)
            nz, ny, nx = img.shape
            self._check(L.ostr_cnn_output_smooth(_ptr(B["out"]), 16, _ptr(img), nz, ny, nx,
                                                 SCALE, taps.ctypes.data_as(ctypes.c_void_p), r,
                                                 int(clamp), _ptr(B["pred"]), st))
        else:
            self._check(L.ostr_cnn_output(_ptr(B["out"]), 16, _ptr(img), n0, SCALE, int(clamp),
                                          _ptr(B["pred"]), st))
        self._img = img
        return B["pred"].view(img.shape)

    # -- backward -------------------------------------------------------------
    def _wgrad(self, name, x, xcs, xco, dy, dcs, dco, lvl, B):
        _, cin, cout, _ = next(l for l in LAYERS if l[0] == name)
        work = B["work"]
        self._check(self._L.ostr_cnn_wgrad(cin, cout, _ptr(x), xcs, xco, _ptr(dy), dcs, dco,
                                           *B["levels"][lvl], _ptr(self.weight(name, self.grads)),
                                           _ptr(self.bias(name, self.grads)), _ptr(work),
                                           work.numel(), B["nsplit"], self._stream()))

    def loss_backward(self, target: torch.Tensor) -> float:
        """mean((pred - target)^2) of the last forward; fills self.grads."""
        img = self._img
        B = self._buffers(tuple(img.shape))
        L, st, lv = self._L, self._stream(), B["levels"]
        n0 = img.numel()
        loss = ctypes.c_double()
        self._check(L.ostr_cnn_mse(_ptr(B["pred"]), _ptr(target.contiguous()), n0,
                                   2.0 / (n0 * SCALE), _ptr(B["dout"]), 16, _ptr(B["mse"]),
                                   B["mse"].numel(), ctypes.byref(loss), st))
        # out (16->16, linear)
        self._wgrad("out", B["d1"], 16, 0, B["dout"], 16, 0, 0, B)
        self._conv("out", B["dout"], 16, 0, B["dd1"], 16, 0, 0, B, mask=B["d1"], mcs=16,
                   grad=True)
        # d1 (48->16)
        self._wgrad("d1", B["cat1"], 48, 0, B["dd1"], 16, 0, 0, B)
        self._conv("d1", B["dd1"], 16, 0, B["dcat1"], 48, 0, 0, B, mask=B["cat1"], mcs=48,
                   grad=True)
        self._check(L.ostr_cnn_up2_bwd(_ptr(B["dcat1"]), 48, 0, _ptr(B["dd2"]), 32, 0,
                                       _ptr(B["d2"]), 32, 0, 32, *lv[0], st))
        # d2 (96->32)
        self._wgrad("d2", B["cat2"], 96, 0, B["dd2"], 32, 0, 1, B)
        self._conv("d2", B["dd2"], 32, 0, B["dcat2"], 96, 0, 1, B, mask=B["cat2"], mcs=96,
                   grad=True)
        self._check(L.ostr_cnn_up2_bwd(_ptr(B["dcat2"]), 96, 0, _ptr(B["de3b"]), 64, 0,
                                       _ptr(B["e3b"]), 64, 0, 64, *lv[1], st))
        # e3b (64->64), e3a (32->64)
        self._wgrad("e3b", B["e3a"], 64, 0, B["de3b"], 64, 0, 2, B)
        self._conv("e3b", B["de3b"], 64, 0, B["de3a"], 64, 0, 2, B, mask=B["e3a"], mcs=64,
                   grad=True)
        self._wgrad("e3a", B["p2"], 32, 0, B["de3a"], 64, 0, 2, B)
        self._conv("e3a", B["de3a"], 64, 0, B["dp2"], 32, 0, 2, B, grad=True)
        # pool(e2b): add into the skip gradient of e2b (dcat2[64:96]), ReLU mask e2b
        self._check(L.ostr_cnn_pool2_bwd(_ptr(B["dp2"]), 32, 0, _ptr(B["dcat2"]), 96, 64,
                                         _ptr(B["cat2"]), 96, 64, 1, 32, *lv[1], st))
        # e2b (32->32), e2a (16->32)
        self._wgrad("e2b", B["e2a"], 32, 0, B["dcat2"], 96, 64, 1, B)
        self._conv("e2b", B["dcat2"], 96, 64, B["de2a"], 32, 0, 1, B, mask=B["e2a"], mcs=32,
                   grad=True)
        self._wgrad("e2a", B["p1"], 16, 0, B["de2a"], 32, 0, 1, B)
        self._conv("e2a", B["de2a"], 32, 0, B["dp1"], 16, 0, 1, B, grad=True)
        self._check(L.ostr_cnn_pool2_bwd(_ptr(B["dp1"]), 16, 0, _ptr(B["dcat1"]), 48, 32,
                                         _ptr(B["cat1"]), 48, 32, 1, 16, *lv[0], st))
        # e1b (16->16), e1a (8->16)
        self._wgrad("e1b", B["e1a"], 16, 0, B["dcat1"], 48, 32, 0, B)
        self._conv("e1b", B["dcat1"], 48, 32, B["de1a"], 16, 0, 0, B, mask=B["e1a"], mcs=16,
                   grad=True)
        self._wgrad("e1a", B["x0"], 8, 0, B["de1a"], 16, 0, 0, B)
        return loss.value / n0

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        self._check(self._L.ostr_cnn_adam(_ptr(self.params), _ptr(self.m), _ptr(self.v),
                                          _ptr(self.grads), self.n_params, lr, b1, b2, eps,
                                          self.step_count, self._stream()))
        # the input channels 1..7 of e1a stay zero (their gradient is zero)
        self.pack()


# ---- data, training, warm-started reconstruction ------------------------------

TRAIN_SEED0, TEST_SEED0 = 100000, 200000


def make_case(scan, seed: int, device):
    """Seeded phantom of the scan's recipe, its cos-weighted Siddon line
    integrals on the GPU, Poisson counts (noise seed = seed + 1000) and blank.
    Returns (x_true np.float32, y torch, b torch)."""
    from .ostr import DBTSystem
    from .phantom import breast_phantom, poisson_counts
    g = scan.geom
    x = breast_phantom(g, seed, scan.n_fg, scan.n_calc)
    sysm = DBTSystem(g, scan.angles, device=device)
    l = sysm.fproj(torch.from_numpy(x).to(sysm.device), 0, scan.n_views)
    y = torch.from_numpy(poisson_counts(l.cpu().numpy(), scan.i0, seed + 1000)).to(sysm.device)
    sysm.close()
    return x, y, torch.full_like(y, scan.i0)


def recon_config(scan, momentum=True):
    """The reconstruction the warm start serves: one subset, exact transpose,
    the scan's prior, Nesterov momentum (accelerated SQS)."""
    from .geometry import OSTRConfig
    c = scan.recon
    return OSTRConfig(n_subsets=1, beta=c.beta, delta=c.delta, prior=c.prior,
                      backprojector="matched", momentum=momentum)


def training_pair(scan, seed: int, iters: int, device):
    """(backprojected image, accelerated-SQS image after `iters` iterations)."""
    from .ostr import OSTR
    _, y, b = make_case(scan, seed, device)
    rec = OSTR(scan.geom, scan.angles, y, b, cfg=recon_config(scan), x0="bp", device=device)
    bp = rec.x.clone()
    target = rec.iterate(iters).clone()
    rec.sys.close()
    return bp, target


def train(net: WarmStartNet, pairs, epochs: int, lr: float = 1e-3, seed: int = 0,
          log=None, val=None):
    """Batch-1 Adam over `pairs` for `epochs` epochs with cosine decay of the
    learning rate; returns per-epoch (mean train loss, mean val loss)."""
    rng = np.random.default_rng(seed)
    total = epochs * len(pairs)
    hist, step = [], 0
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        tl = 0.0
        for i in order:
            bp, tgt = pairs[i]
            net.forward(bp)
            tl += net.loss_backward(tgt)
            lr_t = 0.5 * lr * (1.0 + np.cos(np.pi * step / total))
            net.adam_step(lr=lr_t)
            step += 1
        vl = float("nan")
        if val:
            vl = float(np.mean([evaluate(net, bp, tgt) for bp, tgt in val]))
        hist.append((tl / len(pairs), vl))
        if log:
            log(f"epoch {ep + 1}/{epochs}: train mse {tl / len(pairs):.4e} val mse {vl:.4e}")
    return hist


def evaluate(net: WarmStartNet, bp, target) -> float:
    """Loss of the network on one pair (no gradient)."""
    net.forward(bp)
    B = net._buffers(tuple(bp.shape))
    loss = ctypes.c_double()
    net._check(net._L.ostr_cnn_mse(_ptr(B["pred"]), _ptr(target.contiguous()), bp.numel(), 0.0,
                                   None, 0, _ptr(B["mse"]), B["mse"].numel(),
                                   ctypes.byref(loss), net._stream()))
    return loss.value / bp.numel()


def network_image(net: WarmStartNet, bp: torch.Tensor,
                  smooth_sigma: float = WARM_SMOOTH_SIGMA) -> torch.Tensor:
    """The network's estimate from a backprojected image (clamped at 0), its
    residual low-passed in-plane with a Gaussian of smooth_sigma voxels."""
    shape = bp.shape
    p = pad4(bp)
    out = net.forward(p, clamp=True, smooth_sigma=smooth_sigma)
    return out[:shape[0], :shape[1], :shape[2]].contiguous().clone()


def warm_started(rec, net: WarmStartNet, judge_iters: int = 1):
    """Start rec (an OSTR from its backprojection, not yet iterated) from the
    network's estimate if the plain method prefers it: judge_iters plain passes
    from each start, then the FP64 objectives compared. Returns (accepted,
    obj_bp, obj_net), the objectives after those passes; rec is left at its
    chosen start with no iteration done (the passes are not kept)."""
    if rec.iteration:
        raise ValueError("warm_started: the reconstruction has already iterated")
    bp = rec.x.clone()
    cand = network_image(net, bp.float()).to(bp.dtype)
    objs = {}
    for name, x0 in (("bp", bp), ("net", cand)):
        rec.reset(x0)
        rec.iterate(judge_iters)
        objs[name] = rec.objective()
    accepted = objs["net"] < objs["bp"]
    rec.reset(cand if accepted else bp)
    return accepted, objs["bp"], objs["net"]


def main(argv=None):
    import argparse
    import sys
    import time
    from .geometry import scan_small
    ap = argparse.ArgumentParser(description="train the warm-start network")
    ap.add_argument("--n", type=int, default=2000)
    ap.add_argument("--iters", type=int, default=1000, help="target iterations per pair")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--val", type=int, default=16)
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    scan = scan_small()
    log = lambda m: print(m, flush=True)
    t0 = time.time()
    pairs = []
    for i in range(args.n):
        pairs.append(training_pair(scan, TRAIN_SEED0 + i, args.iters, dev))
        if (i + 1) % 100 == 0:
            log(f"pairs {i + 1}/{args.n} ({time.time() - t0:.0f} s)")
    val = [training_pair(scan, TEST_SEED0 + 100 + i, args.iters, dev) for i in range(args.val)]
    t1 = time.time()
    net = WarmStartNet(dev, seed=0)
    hist = train(net, pairs, args.epochs, args.lr, log=log, val=val)
    meta = {"scan": scan.name, "pairs": args.n, "target_iters": args.iters,
            "epochs": args.epochs, "lr": args.lr, "train_seeds": [TRAIN_SEED0, TRAIN_SEED0 + args.n],
            "history": hist, "data_s": t1 - t0, "train_s": time.time() - t1}
    net.save(args.out, meta)
    log(f"saved {args.out}: data {t1 - t0:.0f} s, training {time.time() - t1:.0f} s")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
