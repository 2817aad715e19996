"""Command-line OSTR run from a reference datafile.

    python -m newrecon_amd.ostr_cli <datafile> [--fused | --unfused] [--gamma row_sum]

Does what `Tomo_OSTR <datafile>` does (main, Tomo_OSTR.cu:2561-2702, and
Reconstruction, Tomo_OSTR_gold.cpp:1493-1528): reads the datafile, the angle
file, the projection / scatter / blank series, runs OSTR with the configured
prior (Recon_Method 0 = Q2, 1 = Lange), starting from `ini_guess` when that
file exists (zeros otherwise, the reference's effective start), and writes the
estimate every Readout_Every_X_Iters iterations. The defaults reproduce the
reference's gamma (ray length) and its unguarded update; --gamma row_sum and
--guard select the alternatives. Errors are reported with a non-zero exit
status, never a partial run.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("datafile")
    ap.add_argument("--unfused", action="store_true", help="launch-by-launch reference order")
    ap.add_argument("--gamma", choices=["ray_length", "row_sum"], default=None)
    ap.add_argument("--guard", action="store_true", help="skip voxels with a zero denominator")
    args = ap.parse_args(argv)

    from .datafile import (DatafileError, load_angles, load_datafile, load_projections,
                           output_file)
    from .ostr import OSTR, save_raw, load_raw

    try:
        run = load_datafile(args.datafile)
        g, cfg, p = run.geom, run.cfg, run.paths
        if args.gamma:
            cfg.gamma_mode = args.gamma
        cfg.guard = args.guard
        cfg.fused = not args.unfused
        angles = load_angles(p["angle_name"], run.n_views)
        y = load_projections(p["projection_folder"], p["projection_name"], run.n_views, g.nbx,
                             g.nby)
        r = load_projections(p["scatter_folder"], p["scatter_name"], run.n_views, g.nbx, g.nby)
        b = load_projections(p["blank_folder"], p["blank_name"], run.n_views, g.nbx, g.nby)
        x0 = load_raw(p["ini_guess"], g) if os.path.exists(p["ini_guess"]) else None
    except (DatafileError, ValueError) as e:
        print(f"ostr_cli: {e}", file=sys.stderr)
        return 1

    rec = OSTR(g, angles, y, b, r=r, cfg=cfg, x0=x0)
    t0 = time.perf_counter()

    def readout(it, x):
        print(f"Finished iteration {it} ({time.perf_counter() - t0:.3f} s)", flush=True)
        if run.readout_every > 0 and it % run.readout_every == 0:
            save_raw(output_file(p["output_recon_folder"], p["output_recon_name"], it - 1), x)

    rec.iterate(cfg.n_iters, callback=readout)
    return 0


if __name__ == "__main__":
    sys.exit(main())
