#!/usr/bin/env python
"""Benchmark of the OSTR DBT hot path on one MI355X.

A step is one full OSTR iteration (all subsets) of plain FP32 OSTR on the
clinical workload: 1400x600x60 voxels of 0.1x0.1x1 mm, 15 views over +-7.5
deg, 2048x1664 detector of 0.14 mm, 5 subsets, hyperbola prior (delta 5e-4,
beta 1e3), gamma = row sum. Inputs (seeded phantom, Poisson counts) are
resident in HBM before the timed region. Prints ONE JSON line.

Also reported: per-kernel average durations from HIP events recorded on the
launch stream, the roofline fraction of the dominant kernel against its
algorithmic bytes, voxel-ray updates/s, and the CPU oracle timed on this
host over one subset of the same workload (scaled to s/iter).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "OSTR s/iter + time to 1e-3 rel-RMSE vs FP64 oracle: warm-start vs plain"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(geom, vs, kernel, s_img=4, s_prj=4):
    """Compulsory HBM bytes of one launch: volume read once, outputs written once."""
    nvox, nbin = geom.n_vox, geom.n_bin
    if kernel == "fproj_h":      # x read; y, b read; h written
        return nvox * s_img + vs * nbin * (4 + 4 + s_prj)
    if kernel == "update_fused":  # h read; x read, d read, x written
        return vs * nbin * s_prj + nvox * (s_img + 4 + s_img)
    raise KeyError(kernel)


PROFILE_JSON = os.path.join(ROOT, "profiles",
                            "fproj_h-update_fused_dbt-clinical-1400x600x60-15v.json")
PROFILED_KERNEL = {"fproj_h": "project_kernel<3, float, float>",
                   "update_fused": "update_fused_kernel<2, float, float>"}


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (clinical workload), or None."""
    try:
        data = json.load(open(PROFILE_JSON))
        return data["kernels"][PROFILED_KERNEL[kernel]]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(scan, x_host, y_host, b_host, threads):
    """Oracle (C++/OpenMP restatement) timed over subset 0 of one iteration."""
    import numpy as np
    from newrecon_amd.geometry import regroup_index
    from oracle import oracle as O

    g = scan.geom
    M = scan.recon.n_subsets
    vs = scan.n_views // M
    idx = regroup_index(scan.n_views, M)
    views = idx[:vs]
    y_sub = np.ascontiguousarray(y_host[views])
    b_sub = np.ascontiguousarray(b_host[views])
    d = np.ones(g.vol_shape, np.float32)  # precompute is not part of an iteration
    O.lib()
    t0 = time.perf_counter()
    l = O.fproj(g, scan.angles, x_host, 0, vs, index=idx, nthreads=threads)
    h = O.compute_h(y_sub, b_sub, l)
    capl = O.bproj_sbp(g, scan.angles, h, 0, vs, index=idx, nthreads=threads)
    rdd, rd = O.prior(g, x_host, 2, scan.recon.delta, nthreads=threads)
    O.update_est(x_host, capl, rd, d, rdd, scan.recon.beta, float(M))
    dt = time.perf_counter() - t0
    return {
        "value": round(dt * M, 3),
        "unit": "s/iter",
        "cores": threads,
        "kind": "port",
        "sample": f"1 of {M} subsets ({vs} views) of one iteration of the same workload, "
                  f"{dt:.2f} s measured, scaled x{M}",
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="clinical", choices=["clinical", "small"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed loop, run 3 plain device copies of 256 MiB (a "
                         "known-traffic kernel for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding of the path is design-only)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from newrecon_amd.geometry import SCANS
    from newrecon_amd.ostr import OSTR
    from newrecon_amd.workload import make_measurement

    dev = torch.device("cuda:0")
    scan = SCANS[args.config]()
    g, cfg = scan.geom, scan.recon
    cfg.fused = True
    print(f"[bench] workload {scan.name}: generating phantom and data", file=sys.stderr,
          flush=True)
    x_true, y, b = make_measurement(scan, device=dev)

    rec = OSTR(g, scan.angles, y, b, cfg=cfg, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec.precompute()
    torch.cuda.synchronize()
    t_pre = time.perf_counter() - t0
    print(f"[bench] precompute {t_pre*1e3:.1f} ms; warmup {args.warmup}", file=sys.stderr,
          flush=True)
    rec.iterate(args.warmup)
    torch.cuda.synchronize()

    # timed region: exactly `steps` iterations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec.iterate(args.steps)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0
    s_per_iter = t_total / args.steps
    print(f"[bench] {args.steps} iterations: {s_per_iter*1e3:.2f} ms/iter", file=sys.stderr,
          flush=True)

    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # per-kernel timing pass (events on the launch stream), same iterations
    rec.events = {}
    rec.iterate(max(2, min(args.steps, 5)))
    torch.cuda.synchronize()
    ktimes = rec.kernel_times_ms()
    rec.events = None

    vs = scan.n_views // cfg.n_subsets
    dominant = max(ktimes, key=lambda k: ktimes[k])
    bytes_dom = algorithmic_bytes(g, vs, dominant)
    achieved = bytes_dom / (ktimes[dominant] * 1e-3) / 1e9
    nnz = rec.sys.count_nnz(0, scan.n_views)
    updates = nnz + g.n_vox * scan.n_views

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"[bench] CPU oracle baseline on {threads} threads", file=sys.stderr, flush=True)
        x_host = rec.x.float().cpu().numpy()
        cpu = cpu_baseline(scan, x_host, y.cpu().numpy(), b.cpu().numpy(), threads)

    out = {
        "metric": METRIC,
        "value": round(s_per_iter, 6),
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_per_iter * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded breast phantom, Poisson counts)",
        "config": {
            "workload": scan.name,
            "volume": [g.nx, g.ny, g.nz], "voxel_mm": [g.dx, g.dy, g.dz],
            "detector": [g.nbx, g.nby], "bin_mm": [g.dbx, g.dby],
            "views": scan.n_views, "subsets": cfg.n_subsets, "prior": cfg.prior,
            "beta": cfg.beta, "delta": cfg.delta, "gamma": cfg.gamma_mode,
            "path": "fused fp32 (fproj_h + update_fused per subset)",
        },
        "voxel_ray_updates_per_s": round(updates / s_per_iter, 1),
        "nnz_forward_per_iter": nnz,
        "precompute_s": round(t_pre, 6),
        "kernel_ms": {k: round(v, 4) for k, v in ktimes.items()},
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (round(profiled_traffic(dominant))
                        if args.config == "clinical" and profiled_traffic(dominant) else None),
            "traffic_source": os.path.relpath(PROFILE_JSON, ROOT),
            "algorithmic_bytes": bytes_dom,
        },
        "cpu_baseline": cpu,
        "time_to_1e-3": None,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
