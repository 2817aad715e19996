"""Summaries of rocprofv3 CSV output for the OSTR kernels.

    python -m newrecon_amd.profiling stats  <dir-with-*_kernel_stats.csv>
    python -m newrecon_amd.profiling hbm    <fetch-run-dir> <write-run-dir> [--json OUT]

    python -m newrecon_amd.profiling valu   <sq-run-dir> <grbm-run-dir> [--json OUT]

`stats` tabulates the kernel-trace statistics. `hbm` turns two counter runs
(`rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`, KiB per dispatch) into
HBM bytes per launch. Each counter is scaled by what it reads for a plain copy
of known size in the same run (the `__amd_rocclr_copyBuffer` dispatches of
256 MiB made by `bench.py --calibrate-copy`): scale = 262144 KiB / counter.
`valu` reads one run with SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU, SQ_INSTS_SALU,
SQ_INSTS_VMEM_RD and SQ_WAVES and one with GRBM_GUI_ACTIVE (summed over the 8
XCDs: / 8 = cycles of the dispatch) and gives, per kernel, the share of the
256 CUs' VALU issue slots used (wave64 VALU instructions / (256 cycles): each
CU's 4 SIMDs issue one wave64 instruction per 4 cycles) and the VALU busy
fraction (SQ_ACTIVE_INST_VALU, which also counts the extra cycles of
multi-cycle instructions such as rsq and rcp, / (256 cycles)).
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re
import sys

CALIB_KIB = 262144.0  # 256 MiB


def short_name(name: str) -> str:
    m = re.search(r"ostr::\(anonymous namespace\)::(\w+)(<[^(]*>)?", name)
    if m:
        return m.group(1) + (m.group(2) or "")
    return name.split("(")[0][:60]


def _one(pattern: str) -> str:
    files = glob.glob(pattern, recursive=True)
    if not files:
        raise FileNotFoundError(pattern)
    return files[0]


def stats_table(run_dir: str) -> str:
    path = _one(os.path.join(run_dir, "**", "*kernel_stats.csv"))
    rows = list(csv.DictReader(open(path)))
    out = ["| kernel | calls | avg us | min us | max us | % time |", "|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| `{short_name(r['Name'])}` | {r['Calls']} | {float(r['AverageNs'])/1e3:.1f} "
                   f"| {float(r['MinNs'])/1e3:.1f} | {float(r['MaxNs'])/1e3:.1f} "
                   f"| {float(r['Percentage']):.1f} |")
    return "\n".join(out)


N_CU = 256
N_XCD = 8


def counter_means(run_dir: str, counter: str):
    path = _one(os.path.join(run_dir, "**", "*counter_collection.csv"))
    per = collections.defaultdict(list)
    calib = []
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        v = float(r["Counter_Value"])
        if "copyBuffer" in r["Kernel_Name"] and int(r["Grid_Size"]) == 131072:
            calib.append(v)  # 256 MiB copy: 64Mi floats, 512 per work-item x 131072
        else:
            per[short_name(r["Kernel_Name"])].append(v)
    means = {k: sum(v) / len(v) for k, v in per.items()}
    scale = CALIB_KIB / (sum(calib) / len(calib)) if calib else None
    return means, scale


def hbm_table(fetch_dir: str, write_dir: str):
    f, fs = counter_means(fetch_dir, "FETCH_SIZE")
    w, ws = counter_means(write_dir, "WRITE_SIZE")
    if fs is None or ws is None:
        raise RuntimeError("no 256 MiB calibration copy found (run bench.py --calibrate-copy)")
    out = [f"calibration (256 MiB copy): FETCH_SIZE scale {fs:.4f}, WRITE_SIZE scale {ws:.4f}", "",
           "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM MB per launch (scaled) |",
           "|---|---|---|---|"]
    res = {}
    for k in sorted(set(f) | set(w)):
        fk, wk = f.get(k, 0.0), w.get(k, 0.0)
        mb = (fk * fs + wk * ws) * 1024 / 1e6
        res[k] = {"fetch_kib": fk, "write_kib": wk, "traffic_bytes": mb * 1e6}
        out.append(f"| `{k}` | {fk:.0f} | {wk:.0f} | {mb:.1f} |")
    return "\n".join(out), {"fetch_scale": fs, "write_scale": ws, "kernels": res}


def valu_table(sq_dir: str, grbm_dir: str):
    cyc, _ = counter_means(grbm_dir, "GRBM_GUI_ACTIVE")
    cols = {c: counter_means(sq_dir, c)[0] for c in
            ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD",
             "SQ_ACTIVE_INST_VALU")}
    out = ["| kernel | cycles | waves | VALU instr (wave) | SALU instr | VMEM read instr "
           "| VALU issue used | VALU busy |", "|---|---|---|---|---|---|---|---|"]
    res = {}
    for k in sorted(set(cyc) & set(cols["SQ_INSTS_VALU"])):
        c = cyc[k] / N_XCD
        if c <= 0:
            continue
        issue = cols["SQ_INSTS_VALU"][k] / (N_CU * c)
        busy = cols["SQ_ACTIVE_INST_VALU"].get(k, 0.0) / (N_CU * c)
        res[k] = {"cycles": c, "waves": cols["SQ_WAVES"].get(k, 0.0),
                  "valu_instr": cols["SQ_INSTS_VALU"][k],
                  "salu_instr": cols["SQ_INSTS_SALU"].get(k, 0.0),
                  "vmem_rd_instr": cols["SQ_INSTS_VMEM_RD"].get(k, 0.0),
                  "valu_issue": issue, "valu_busy": busy}
        out.append(f"| `{k}` | {c:,.0f} | {res[k]['waves']:,.0f} | {res[k]['valu_instr']:.3g} "
                   f"| {res[k]['salu_instr']:.3g} | {res[k]['vmem_rd_instr']:.3g} "
                   f"| {issue:.2f} | {busy:.2f} |")
    return "\n".join(out), {"kernels": res}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    b = sub.add_parser("hbm")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--json")
    v = sub.add_parser("valu")
    v.add_argument("sq_dir")
    v.add_argument("grbm_dir")
    v.add_argument("--json")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(stats_table(args.dir))
    else:
        if args.cmd == "hbm":
            text, data = hbm_table(args.fetch_dir, args.write_dir)
        else:
            text, data = valu_table(args.sq_dir, args.grbm_dir)
        print(text)
        if args.json:
            with open(args.json, "w") as fh:
                json.dump(data, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
