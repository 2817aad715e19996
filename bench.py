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
host over one full iteration of the same workload (on the CPUs this job may
use, and on 1 thread).
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


VALU_JSON = os.path.join(ROOT, "profiles",
                         "valu_fproj_h-update_fused_dbt-clinical-1400x600x60-15v.json")


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (clinical workload), or None."""
    try:
        data = json.load(open(PROFILE_JSON))
        return data["kernels"][PROFILED_KERNEL[kernel]]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def profiled_valu(kernel):
    """(VALU issue share, VALU busy) of `kernel` from the committed rocprofv3
    SQ-counter summary (clinical workload), or None."""
    try:
        k = json.load(open(VALU_JSON))["kernels"][PROFILED_KERNEL[kernel]]
        return round(k["valu_issue"], 3), round(k["valu_busy"], 3)
    except (OSError, KeyError, ValueError):
        return None


def cpu_share():
    """CPUs this process may use: the affinity mask, capped by a cgroup v2
    quota (cpu.max) when one is set; plus os.cpu_count() for the record."""
    n_aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    share = n_aff or os.cpu_count() or 1
    if quota:
        share = min(share, quota)
    return share, {"nproc": os.cpu_count(), "affinity": n_aff, "cgroup_quota": quota,
                   "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(scan, x_host, y_host, b_host, threads):
    """Oracle (C++/OpenMP restatement of Tomo_OSTR_gold.cpp:724-887 and the
    Tomo_OSTR.cu kernels) timed over one full OSTR iteration (all subsets) of
    the same workload on this host: first the precompute (gamma, yry, d =
    SBP(gamma yry), :450-538, its d used by the iteration), then one iteration
    on `threads` threads and one on 1 thread. steady clock, file I/O none."""
    import numpy as np
    from newrecon_amd.geometry import GAMMA_MODES, regroup_index
    from oracle import oracle as O

    g = scan.geom
    cfg = scan.recon
    M = cfg.n_subsets
    nv = scan.n_views
    vs = nv // M
    O.lib()

    t0 = time.perf_counter()
    gam = O.gamma(g, scan.angles, 0, nv, GAMMA_MODES[cfg.gamma_mode], nthreads=threads)
    c = O.yry(y_host)
    d = O.bproj_sbp(g, scan.angles, (gam * c).astype(np.float32), 0, nv, nthreads=threads)
    t_pre = time.perf_counter() - t0
    del gam, c
    idx = regroup_index(nv, M)
    y_sub = [np.ascontiguousarray(y_host[idx[m * vs:(m + 1) * vs]]) for m in range(M)]
    b_sub = [np.ascontiguousarray(b_host[idx[m * vs:(m + 1) * vs]]) for m in range(M)]

    def iteration(nt):
        x = x_host.copy()
        t0 = time.perf_counter()
        for m in range(M):
            a0, a1 = m * vs, (m + 1) * vs
            l = O.fproj(g, scan.angles, x, a0, a1, index=idx, nthreads=nt)
            h = O.compute_h(y_sub[m], b_sub[m], l)
            capl = O.bproj_sbp(g, scan.angles, h, a0, a1, index=idx, nthreads=nt)
            rdd, rd = O.prior(g, x, 2, cfg.delta, nthreads=nt)
            x = O.update_est(x, capl, rd, d, rdd, cfg.beta, float(M))
        return time.perf_counter() - t0

    dt = iteration(threads)
    _, share = cpu_share()
    out = {
        "value": round(dt, 3),
        "unit": "s/iter",
        "cores": threads,
        "kind": "port",
        "sample": f"one full iteration ({M} subsets, {nv} views) of the same workload, "
                  f"d from the oracle's own precompute ({t_pre:.2f} s, not in the value)",
        "cpu_model": cpu_model(),
        "host": share,
    }
    if threads > 1:
        d1 = iteration(1)
        out["single_thread"] = {"value": round(d1, 3), "unit": "s/iter", "cores": 1,
                                "sample": "the same iteration on 1 thread"}
    return out


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def bench_batch(args, dev):
    """Uneven batch: 8 cases (30..80 slices, 800x400..1400x700 voxels, 9/15/25
    views, 3/5/5 subsets), FP32 images, BF16 subset sinograms, tile shapes from
    the autotuner. A step is one batched subset of every case (two launches).
    Reports aggregate and per-case throughput, and the same cases run one
    after another with the single-case kernels."""
    import torch

    from newrecon_amd.autotune import tune
    from newrecon_amd.batch import BatchOSTR
    from newrecon_amd.geometry import scan_batch
    from newrecon_amd.ostr import OSTR
    from newrecon_amd.workload import make_measurement

    scans = scan_batch()
    recs = []
    for s in scans:
        print(f"[bench] case {s.name}: phantom and data", file=sys.stderr, flush=True)
        _, y, b = make_measurement(s, device=dev)
        cfg = s.recon
        cfg.fused = True
        recs.append(OSTR(s.geom, s.angles, y, b, cfg=cfg, prj_dtype=torch.bfloat16,
                         device=dev))
        del y, b
    bt = BatchOSTR(recs)
    torch.cuda.synchronize()
    tuned = tune(bt, cache_path=args.tune_cache, reps=3)
    print(f"[bench] tiles {bt.tile.as_dict()} (cached={tuned['cached']})", file=sys.stderr,
          flush=True)
    bt.step(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bt.step(args.steps)
    torch.cuda.synchronize()
    s_step = (time.perf_counter() - t0) / args.steps

    # kernel durations (events on the launch stream)
    def timed(fn, n=5):
        e = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
             for _ in range(n)]
        for a, z in e:
            a.record()
            fn()
            z.record()
        torch.cuda.synchronize()
        return sum(a.elapsed_time(z) for a, z in e) / n

    k_fwd = timed(bt.fproj_h)
    k_upd = timed(bt.update)
    # the same subsets, case by case, with the single-case fused kernels
    alone = []
    for k, rec in enumerate(recs):
        m = bt.pos[k]
        alone.append(timed(lambda: rec._subset(m)))
    updates, bytes_fwd, bytes_upd, per_case = 0, 0, 0, []
    for s, rec, t_alone in zip(scans, recs, alone):
        g = s.geom
        nnz = rec.sys.count_nnz(0, s.n_views) / rec.M   # per subset, on average
        upd = nnz + g.n_vox * rec.vs
        updates += upd
        bytes_fwd += algorithmic_bytes(g, rec.vs, "fproj_h", 4, 2)
        bytes_upd += algorithmic_bytes(g, rec.vs, "update_fused", 4, 2)
        per_case.append({"case": s.name, "subsets": rec.M, "views": s.n_views,
                         "voxel_ray_updates_per_subset": int(upd),
                         "alone_ms_per_subset": round(t_alone, 4),
                         "alone_s_per_iter": round(t_alone * rec.M / 1e3, 6)})
    dominant, k_dom, b_dom = (("batch_update_fused", k_upd, bytes_upd) if k_upd >= k_fwd
                              else ("batch_fproj_h", k_fwd, bytes_fwd))
    achieved = b_dom / (k_dom * 1e-3) / 1e9
    out = {
        "metric": "batched OSTR voxel-ray updates/s (8 uneven cases, one launch per kernel)",
        "value": round(updates / s_step, 1),
        "unit": "voxel-ray updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(s_step * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 image, bf16 subset sinogram (h)",
        "data": "synthetic (seeded breast phantoms, Poisson counts)",
        "config": {"workload": "dbt-batch-8-uneven", "cases": [s.name for s in scans],
                   "tiles": bt.tile.as_dict(), "tuner_cached": tuned["cached"]},
        "kernel_ms": {"batch_fproj_h": round(k_fwd, 4), "batch_update_fused": round(k_upd, 4)},
        "sequential_single_case_ms_per_step": round(sum(alone), 4),
        "batch_speedup_vs_sequential": round(sum(alone) / (s_step * 1e3), 3),
        "per_case": per_case,
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes": b_dom},
        "tuner_ms": {k: tuned.get(k) for k in ("fwd_ms", "upd_ms")},
    }
    print(json.dumps(out), flush=True)
    return 0


def calibrate_copy(dev):
    """Three plain device copies of 256 MiB: a kernel of known traffic for
    scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


SART_BETA, SART_DELTA, SART_LAMBDA = -2e-3, 1e-3, 1.0


def sart_bytes(geom, vs, kernel):
    """Compulsory HBM bytes of one fused OS-SART launch (FP32 everywhere)."""
    nvox, nbin = geom.n_vox, geom.n_bin
    return {"sart_fproj_diff": nvox * 4 + 2 * vs * nbin * 4,   # x read; p read, diff written
            "sart_update_fused": vs * nbin * 4 + 2 * nvox * 4,  # diff read; x read, x written
            }[kernel]


BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense BF16


def cnn_flops(shape):
    """Convolution flops of one training step of the warm-start network on a
    [nz, ny, nx] volume: per layer 2 V cin cout 27 for the forward, the same
    for the weight gradient and, except the first layer, for the data
    gradient (cin as the kernels run it, the first layer's 8 padded channels)."""
    from newrecon_amd.warmstart import LAYERS
    nz, ny, nx = shape
    level = {"e1a": 0, "e1b": 0, "d1": 0, "out": 0, "e2a": 1, "e2b": 1, "d2": 1,
             "e3a": 2, "e3b": 2}
    fwd = bwd = 0
    for name, cin, cout, _ in LAYERS:
        v = (nz >> level[name]) * (ny >> level[name]) * (nx >> level[name])
        f = 2 * v * cin * cout * 27
        fwd += f
        bwd += f + (0 if name == "e1a" else f)
    return fwd, bwd


def bench_cnn(args, dev):
    """One training step of the warm-start network (forward, loss and
    backward, Adam) at its stated training size: a seeded 256x256x48 case
    (0.1x0.1x1 mm, 15 views over +-7.5 deg, I0 = 5e3), input its backprojected
    image, target its 50-iteration OSTR image. A step is one forward +
    backward + Adam; roofline: convolution flops / time against the BF16 MFMA
    peak (the convolutions run as BF16 MFMA, FP32 accumulation)."""
    import torch

    from newrecon_amd.geometry import OSTRConfig, scan_train5
    from newrecon_amd.ostr import OSTR
    from newrecon_amd.warmstart import WarmStartNet, make_case

    scan = scan_train5()
    print(f"[bench] workload {scan.name}: phantom, counts, 50-iteration target",
          file=sys.stderr, flush=True)
    _, y, b = make_case(scan, scan.phantom_seed, dev)
    rec = OSTR(scan.geom, scan.angles, y, b, cfg=OSTRConfig(n_subsets=5), x0="bp", device=dev)
    bp = rec.x.clone()
    target = rec.iterate(50).clone()
    del rec
    net = WarmStartNet(dev, seed=0)

    def step():
        net.forward(bp)
        loss = net.loss_backward(target)
        net.adam_step(lr=1e-3)
        return loss

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    s_step = (time.perf_counter() - t0) / args.steps
    fwd, bwd = cnn_flops(tuple(bp.shape))
    # forward alone (HIP events on the launch stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        net.forward(bp)
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / 3
    tf = (fwd + bwd) / s_step / 1e12
    out = {
        "metric": "warm-start network training step (forward + backward + Adam)",
        "value": round(s_step, 6), "unit": "s/step", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(s_step * 1e3, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (MFMA, fp32 accumulation)",
        "data": "synthetic (seeded breast phantom, its 50-iteration OSTR image as target)",
        "config": {"workload": scan.name, "volume": [scan.geom.nx, scan.geom.ny, scan.geom.nz],
                   "channels": [16, 32, 64]},
        "forward_ms": round(fwd_ms, 4),
        "conv_gflop_per_step": round((fwd + bwd) / 1e9, 2),
        "loss_first_last": [losses[0], losses[-1]],
        "roofline": {"kernel": "all convolutions of a step", "bound": "mfma",
                     "achieved": round(tf, 2), "peak": BF16_MFMA_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tf / BF16_MFMA_PEAK_TFLOPS, 4),
                     "forward_tflops": round(fwd / (fwd_ms * 1e-3) / 1e12, 2)},
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_sart(args, dev):
    """OS-SART on the clinical workload (5 subsets of 3 views, Huber prior,
    lambda 1), fused: per subset the SART forward with row sum and
    SART_prj_diff in one launch, the backprojection update with the Huber
    prior in another. A step is one full pass over the subsets. The five
    reference launchers run one after another are timed beside it."""
    import torch

    from newrecon_amd.geometry import GAMMA_MODES, SCANS
    from newrecon_amd.sart import OSSART
    from newrecon_amd.workload import make_measurement

    scan = SCANS["clinical"]()
    g, M = scan.geom, scan.recon.n_subsets
    print(f"[bench] workload {scan.name} (OS-SART): phantom and data", file=sys.stderr,
          flush=True)
    _, y, b = make_measurement(scan, device=dev)
    p = torch.log(b / y.clamp(min=1.0)).contiguous()
    del y, b
    rec = OSSART(g, scan.angles, p, n_subsets=M, lam=SART_LAMBDA, beta=SART_BETA,
                 delta=SART_DELTA, cos_weight=True, device=dev)
    torch.cuda.synchronize()
    rec.iterate(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec.iterate(args.steps)
    torch.cuda.synchronize()
    s_iter = (time.perf_counter() - t0) / args.steps
    if args.calibrate_copy:
        calibrate_copy(dev)

    # per-launch durations (events on the launch stream) over one more pass
    s, vs = rec.sys, rec.vs
    ev = {}

    def rec_ev(name, fn):
        a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        z.record()
        ev.setdefault(name, []).append((a, z))

    for m in range(M):
        a0, a1 = m * vs, (m + 1) * vs
        rec_ev("sart_fproj_diff", lambda: s.sart_fproj_diff(rec.x, rec.p[a0:a1], a0, a1,
                                                            rec.diff, True))
        rec_ev("sart_update_fused", lambda: s.sart_update_fused(
            rec.x, rec.x_next, rec.diff, a0, a1, SART_LAMBDA, SART_BETA, SART_DELTA))
        rec.x, rec.x_next = rec.x_next, rec.x
    torch.cuda.synchronize()
    kms = {k: sum(a.elapsed_time(z) for a, z in v) / len(v) for k, v in ev.items()}
    dominant = max(kms, key=lambda k: kms[k])
    b_dom = sart_bytes(g, vs, dominant)
    achieved = b_dom / (kms[dominant] * 1e-3) / 1e9
    # the 2x2 supersampled forward (fprojectCB1_4R_GPU_SART) of one subset,
    # merged in the kernel and with the sub-bin sinogram kept
    ev4 = {}
    prj4 = torch.empty((vs, 2 * g.nby, 2 * g.nbx), device=dev)
    for _ in range(3):
        for keep in (False, True):
            a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            s.fproj_4r(rec.x, 0, vs, prj_4r=prj4 if keep else None)
            z.record()
            ev4.setdefault(keep, []).append((a, z))
    torch.cuda.synchronize()
    k4 = {k: sum(a.elapsed_time(z) for a, z in v) / len(v) for k, v in ev4.items()}
    del prj4
    # the five reference launchers one after another, for comparison
    ref = OSSART(g, scan.angles, p, x0=rec.x, n_subsets=M, lam=SART_LAMBDA, beta=SART_BETA,
                 delta=SART_DELTA, cos_weight=True, fused=False, device=dev)
    ref.iterate(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref.iterate(2)
    torch.cuda.synchronize()
    s_unfused = (time.perf_counter() - t0) / 2
    del ref, p
    nnz = s.count_nnz(0, scan.n_views)
    updates = 2 * nnz + g.n_vox * scan.n_views   # two walks per ray + one gather per voxel-view

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        # the CPUs this job may use (affinity and cgroup quota), not the
        # machine's count: threads beyond the quota only time-slice
        threads = args.cpu_threads or cpu_share()[0]
        print(f"[bench] CPU oracle baseline (OS-SART subset) on {threads} threads",
              file=sys.stderr, flush=True)
        x_host = rec.x.cpu().numpy()
        idx = rec.sys.index
        p_sub = rec.p[:vs].transpose(1, 2).contiguous().cpu().numpy()
        O.lib()
        t0 = time.perf_counter()
        est = O.fproj(g, scan.angles, x_host, 0, vs, index=idx, nthreads=threads)
        norm = O.gamma(g, scan.angles, 0, vs, GAMMA_MODES["row_sum"], index=idx,
                       nthreads=threads)
        diff = O.sart_prj_diff(p_sub, est, norm)
        pri = O.prior_huber(g, x_host, SART_DELTA, nthreads=threads)
        O.bproj_sart(g, scan.angles, diff, 0, vs, x_host, SART_LAMBDA, index=idx, prior=pri,
                     beta=SART_BETA, nthreads=threads)
        dt = time.perf_counter() - t0
        cpu = {"value": round(dt * M, 3), "unit": "s/iter", "cores": threads, "kind": "port",
               "sample": f"1 of {M} subsets ({vs} views) of one OS-SART pass of the same "
                         f"workload, {dt:.2f} s measured, scaled x{M}"}

    out = {
        "metric": "OS-SART s/iter (clinical DBT workload)",
        "value": round(s_iter, 6), "unit": "s/iter",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(s_iter * 1e3, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded breast phantom, Poisson counts)",
        "config": {"workload": scan.name + "-ossart", "subsets": M, "views": scan.n_views,
                   "lambda": SART_LAMBDA, "beta": SART_BETA, "delta": SART_DELTA,
                   "cos_weight": True},
        "voxel_ray_updates_per_s": round(updates / s_iter, 1),
        "kernel_ms": {k: round(v, 4) for k, v in kms.items()},
        "unfused_s_per_iter": round(s_unfused, 6),
        "fproj_4r_ms": {"merged": round(k4[False], 4), "with_sub_bins": round(k4[True], 4),
                        "rays": 4 * vs * g.n_bin,
                        "algorithmic_bytes": {"merged": g.n_vox * 4 + vs * g.n_bin * 4,
                                              "with_sub_bins": g.n_vox * 4 + 5 * vs * g.n_bin * 4}},
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes": b_dom},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_recondbt(args, dev):
    """ReconDBT's SART and ML (SART_dbt.m, ML_dbt.m) on the clinical workload
    with this build's projector pair (Siddon forward without cos, exact
    transpose): a step is one SART_dbt iteration (15 views, one update per
    view); one ML_dbt iteration is timed beside it."""
    import torch

    from newrecon_amd.geometry import SCANS
    from newrecon_amd.ostr import DBTSystem
    from newrecon_amd.recondbt import ml_dbt, sart_dbt
    from newrecon_amd.workload import make_measurement

    scan = SCANS["clinical"]()
    g = scan.geom
    print(f"[bench] workload {scan.name} (ReconDBT SART / ML): phantom and data",
          file=sys.stderr, flush=True)
    _, y, b = make_measurement(scan, device=dev)
    S = DBTSystem(g, scan.angles, device=dev)
    p = S.line_integrals(y, b)
    x0 = S.bp_init(y=y, b=b)
    torch.cuda.synchronize()
    # the per-view row and column sums are set-up work: time one iteration
    # with and without them by running 1 and 1 + steps iterations, after an
    # untimed call that takes the one-time costs (first launches, allocations)
    sart_dbt(g, scan.angles, p, x0, 1, 0.15, device=dev)
    ml_dbt(g, scan.angles, y, x0, b, 1, 2.0, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sart_dbt(g, scan.angles, p, x0, 1, 0.15, device=dev)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sart_dbt(g, scan.angles, p, x0, 1 + args.steps, 0.15, device=dev)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    s_sart = ((t2 - t1) - (t1 - t0)) / args.steps
    print(f"[bench] SART_dbt {s_sart * 1e3:.1f} ms/iter", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ml_dbt(g, scan.angles, y, x0, b, 1, 2.0, device=dev)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ml_dbt(g, scan.angles, y, x0, b, 3, 2.0, device=dev)
    torch.cuda.synchronize()
    s_ml = ((time.perf_counter() - t1) - (t1 - t0)) / 2
    # the dominant kernel: the exact transpose of one view
    ev = []
    one = S.sinogram(1)
    one.fill_(1.0)
    img = S.volume()
    for v in range(3):
        a, z = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        S.bproj_matched(one, v, v + 1, img=img, cos_weight=False)
        z.record()
        ev.append((a, z))
    torch.cuda.synchronize()
    k_bt = sum(a.elapsed_time(z) for a, z in ev) / len(ev)
    b_bt = g.n_bin * 4 + 2 * g.n_vox * 4       # one view read; volume read and written
    achieved = b_bt / (k_bt * 1e-3) / 1e9
    # the FBP initialiser (cosine weight and FFT row filter, SBP), timed with
    # events over the whole call (handle creation included) after one warm-up
    # call, and its row filter alone over all views on a resident handle
    import numpy as np
    from newrecon_amd.fbp import fbp_dbt, filter_taps
    fbp_dbt(g, scan.angles, p, "hann", device=dev)
    fa, fz = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fa.record()
    fbp_dbt(g, scan.angles, p, "hann", device=dev)
    fz.record()
    torch.cuda.synchronize()
    fbp_ms = fa.elapsed_time(fz)
    taps = torch.from_numpy(filter_taps(g.nbx, g.dbx, "hann").astype(np.float32)).to(dev)
    q = S.fbp_filter(p, taps, 0, scan.n_views)
    fa.record()
    for _ in range(5):
        S.fbp_filter(p, taps, 0, scan.n_views, q=q)
    fz.record()
    torch.cuda.synchronize()
    filt_ms = fa.elapsed_time(fz) / 5
    print(f"[bench] FBP (hann) {fbp_ms:.2f} ms, row filter {filt_ms:.3f} ms", file=sys.stderr,
          flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        import numpy as np
        from oracle import recondbt as R
        # the CPUs this job may use (affinity and cgroup quota), not the
        # machine's count: threads beyond the quota only time-slice
        threads = args.cpu_threads or cpu_share()[0]
        print(f"[bench] CPU oracle (one view of SART_dbt) on {threads} threads",
              file=sys.stderr, flush=True)
        xh = x0.cpu().numpy().astype(np.float64)
        ph = p[:1].cpu().numpy().astype(np.float64)
        t0 = time.perf_counter()
        est = R._G(g, scan.angles, xh, 0)
        rows = R.O.gamma(g, scan.angles, 0, 1, 1, dtype=np.float64, nthreads=threads)
        yv = np.where(rows > 5 * g.dz, (ph - est) / np.where(rows > 0, rows, 1.0), 0.0)
        num = R._Gt(g, scan.angles, yv, 0, threads)
        den = R._Gt(g, scan.angles, np.ones_like(yv), 0, threads)
        with np.errstate(divide="ignore", invalid="ignore"):
            u = np.nan_to_num(num / den)
        xh += 0.15 * u
        dt = time.perf_counter() - t0
        cpu = {"value": round(dt * scan.n_views, 3), "unit": "s/iter", "cores": threads,
               "kind": "port",
               "sample": f"1 of {scan.n_views} views of one SART_dbt iteration (forward, row "
                         f"sums, two FP64 exact transposes), {dt:.2f} s measured, "
                         f"scaled x{scan.n_views}"}

    out = {
        "metric": "ReconDBT SART_dbt s/iter (clinical DBT workload)",
        "value": round(s_sart, 6), "unit": "s/iter",
        "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(s_sart * 1e3, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded breast phantom, Poisson counts)",
        "config": {"workload": scan.name + "-recondbt", "views": scan.n_views,
                   "sart_stepsize": 0.15, "ml_stepsize": 2.0, "x0": "BP image"},
        "ml_dbt_s_per_iter": round(s_ml, 6),
        "fbp_ms": round(fbp_ms, 3),
        "fbp_filter_ms": round(filt_ms, 4),
        "kernel_ms": {"matched_transpose_one_view": round(k_bt, 4)},
        "roofline": {"kernel": "matched_transpose_one_view", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes": b_bt},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="clinical",
                    choices=["clinical", "small", "batch", "sart", "recondbt", "cnn"])
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache (JSON) for --config batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-time-to-tol", action="store_true",
                    help="skip the time-to-1e-3 measurement (200x100x20 workload)")
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
    if args.config == "batch":
        return bench_batch(args, dev)
    if args.config == "sart":
        return bench_sart(args, dev)
    if args.config == "recondbt":
        return bench_recondbt(args, dev)
    if args.config == "cnn":
        return bench_cnn(args, dev)
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
        calibrate_copy(dev)

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
        # the CPUs this job may use (affinity and cgroup quota), not the
        # machine's count: threads beyond the quota only time-slice
        threads = args.cpu_threads or cpu_share()[0]
        print(f"[bench] CPU oracle baseline on {threads} threads", file=sys.stderr, flush=True)
        x_host = rec.x.float().cpu().numpy()
        cpu = cpu_baseline(scan, x_host, y.cpu().numpy(), b.cpu().numpy(), threads)

    ttt = None
    if not args.no_time_to_tol:
        # time to 1e-3 rel-RMSE against the FP64 converged image (200x100x20
        # workload): accelerated OSTR from the backprojection vs from the
        # network's estimate, and the M = 3 OSTR floor
        from newrecon_amd.timetotol import measure
        print("[bench] time to 1e-3 vs the FP64 converged image", file=sys.stderr, flush=True)
        del rec
        torch.cuda.empty_cache()
        m = measure(device=dev)
        ttt = {
            "workload": "dbt-small-200x100x20-9v", "tol": m["tol"],
            "iteration": "OSTR, 1 subset, exact transpose, Nesterov momentum",
            "plain_s": m["plain"]["seconds"], "plain_iters": m["plain"]["iters"],
            "warm_s": m["warm"]["seconds"], "warm_iters": m["warm"]["iters"],
            "warm_accepted": m["warm"]["accepted"],
            "speedup": m["speedup"],
            "warm_mp": {k: m["warm_mp"].get(k) for k in ("storage", "seconds", "iters",
                                                          "speedup", "floor", "accepted")},
            "rel_start": {"plain": m["plain"]["rel_start"], "warm": m["warm"]["rel_start"],
                          "warm_mp": m["warm_mp"]["rel_start"]},
            "rel_after_k": {"plain": m["plain"]["rel_after"], "warm": m["warm"]["rel_after"]},
            "reference_sbp_m3_floor": m["reference_sbp_m3"],
            "matched_m3_floor": m["matched_m3"],
            "xstar": m["xstar"],
        }

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
            # what bounds the kernel instead: the share of the VALU issue
            # slots it uses and its VALU busy fraction (SQ counters)
            "valu_issue": (profiled_valu(dominant) or (None, None))[0]
                          if args.config == "clinical" else None,
            "valu_busy": (profiled_valu(dominant) or (None, None))[1]
                         if args.config == "clinical" else None,
            "valu_source": os.path.relpath(VALU_JSON, ROOT),
        },
        "cpu_baseline": cpu,
        "time_to_1e-3": ttt,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
