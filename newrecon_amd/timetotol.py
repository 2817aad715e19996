"""Time to a relative RMSE tolerance against the FP64 converged image.

On the 200x100x20 / 9-view workload (seeds 1 / 1001), x* is the FP64
minimiser of the objective computed by oracle/converged.py (accelerated SQS
with the exact transpose, FP64 throughout) and committed as
tests/golden/xstar_small.npz with the SHA-256 of the counts it was computed
from. Measured here, on the GPU, from data resident in HBM:

* plain: OSTR with the exact transpose, one subset, Nesterov momentum
  (accelerated SQS), FP32, from the one-step backprojected image;
* warm: the same iteration from the network's estimate, its residual
  low-passed in-plane (accepted only if, after one plain pass from each
  start, the FP64 objective is lower from the network's image than from the
  backprojection; else the plain start is used: the fallback);
* warm_mp: the warm path with the image and h stored in 16 bits (rays,
  updates and momentum in FP32, the objective, restart test and distance in
  FP64); when it does not reach the tolerance, the error floor it settles at;
* the reference's own iteration, plain OSTR with the SBP backprojector and
  M = 3 subsets, and OS-SQS with the exact transpose and M = 3: both end in an
  ordered-subsets limit cycle; the error after m3_iters is reported.

Each path first runs with a distance check after every iteration (FP64 GPU
reduction) to find the iteration count; the wall-clock is then taken on a
clean run of that many iterations (start image, precompute, network and
fallback test included; checks excluded), the iterations replayed from a HIP
graph (the momentum restart decision stays on the device, so an iteration
needs no host round trip; the graph run equals the eager one bit for bit).
"""
from __future__ import annotations

import json
import os
import time
from typing import Optional

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
XSTAR = os.path.join(ROOT, "tests", "golden", "xstar_small.npz")
WEIGHTS = os.path.join(HERE, "warmstart_small.npz")


def load_xstar(path: str = XSTAR):
    data = np.load(path, allow_pickle=False)
    return data["xstar"], json.loads(str(data["meta"]))


def _data_checksum(y: torch.Tensor) -> str:
    import hashlib
    return hashlib.sha256(y.float().cpu().numpy().astype(np.float32).tobytes()).hexdigest()


def _rel(sysm, x, xs) -> float:
    d, n = sysm.diff_norms(x, xs)
    return float(np.sqrt(d / n))


def _sync():
    torch.cuda.synchronize()


MP_STORAGE = (torch.float32, torch.float16)  # (image, h) of warm_mp


def measure(tol: float = 1e-3, max_iters: int = 20000, device=None, weights: str = WEIGHTS,
            xstar_path: str = XSTAR, ks=(2, 5, 10), m3_iters: int = 3000,
            mp_storage=MP_STORAGE) -> dict:
    from .geometry import OSTRConfig, scan_small
    from .ostr import OSTR
    from .warmstart import WarmStartNet, make_case, recon_config, warm_started

    dev = torch.device(device if device is not None else "cuda:0")
    scan = scan_small()
    xs_np, meta = load_xstar(xstar_path)
    _, y, b = make_case(scan, scan.phantom_seed, dev)
    if _data_checksum(y) != meta["y_sha256"]:
        raise RuntimeError("the counts differ from those x* was computed from")
    xs = torch.from_numpy(xs_np).to(dev)
    net = WarmStartNet(dev)
    wmeta = net.load(weights)
    cfg = recon_config(scan)

    def run(start: str, img=torch.float32, h=torch.float32, n_check: Optional[int] = None,
            n_time: Optional[int] = None):
        """start 'bp' or 'net'; checks per iteration up to n_check (returns
        iterations to tol, the error after ks and the lowest error reached),
        or times n_time iterations."""
        _sync()
        t0 = time.perf_counter()
        rec = OSTR(scan.geom, scan.angles, y, b, cfg=cfg, x0="bp", device=dev, img_dtype=img,
                   prj_dtype=h)
        info = {}
        if start == "net":
            ok, ob, on = warm_started(rec, net)
            info["accepted"] = ok
            info["obj_bp"], info["obj_net"] = ob, on
        if not rec._precomputed:
            rec.precompute()
        if n_time is not None:
            rec.iterate(n_time, graph=True)
            _sync()
            info["seconds"] = time.perf_counter() - t0
            return info
        info["rel_start"] = _rel(rec.sys, rec.x.float(), xs)
        info["rel_after"] = {}
        hit, best, best_k, r = None, np.inf, 0, np.inf
        for k in range(1, n_check + 1):
            rec.iterate(1)
            r = _rel(rec.sys, rec.estimate.float(), xs)
            if r < best:
                best, best_k = r, k
            if k in ks:
                info["rel_after"][k] = r
            if r < tol and hit is None:
                hit = k
                if k >= max(ks):
                    break
        info["iters"] = hit
        info["rel_end"] = r
        info["floor"] = {"rel": best, "at_iter": best_k}
        return info

    out = {"tol": tol, "xstar": {k: meta[k] for k in ("iterations", "fixed_point_residual",
                                                        "objective") if k in meta},
           "weights": {k: wmeta.get(k) for k in ("pairs", "target_iters", "epochs")}}
    paths = {"plain": ("bp", torch.float32, torch.float32),
             "warm": ("net", torch.float32, torch.float32),
             "warm_mp": ("net",) + tuple(mp_storage)}
    for name, (start, img, h) in paths.items():
        chk = run(start, img, h, n_check=max_iters)
        res = dict(chk)
        res["storage"] = {"image": str(img).replace("torch.", ""),
                          "h": str(h).replace("torch.", "")}
        if chk["iters"] is not None:
            run(start, img, h, n_time=chk["iters"])  # warm-up of code paths
            res["seconds"] = run(start, img, h, n_time=chk["iters"])["seconds"]
        else:
            res["seconds"] = None
        out[name] = res
    p = out["plain"]["seconds"]
    for name in ("warm", "warm_mp"):
        w = out[name]["seconds"]
        out[name]["speedup"] = (p / w) if (p and w) else None
    out["speedup"] = out["warm"]["speedup"]

    # ordered subsets (M = 3), no momentum: the reference's own iteration (SBP
    # backprojector) and OS-SQS with the exact transpose; both end in a limit cycle
    for name, bproj in (("reference_sbp_m3", "sbp"), ("matched_m3", "matched")):
        cfg3 = OSTRConfig(n_subsets=3, beta=cfg.beta, delta=cfg.delta, prior=cfg.prior,
                          backprojector=bproj)
        rec = OSTR(scan.geom, scan.angles, y, b, cfg=cfg3, x0="bp", device=dev)
        rec.precompute()
        _sync()
        t0 = time.perf_counter()
        rec.iterate(m3_iters)
        _sync()
        rel = _rel(rec.sys, rec.x, xs)
        out[name] = {"iters": m3_iters, "seconds": time.perf_counter() - t0, "rel": rel,
                     "reached": bool(rel < tol), "backprojector": bproj}
    return out


if __name__ == "__main__":
    print(json.dumps(measure(), indent=1))
