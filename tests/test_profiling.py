"""The rocprofv3 summaries bench.py reads (HBM bytes per launch, VALU issue
share and busy fraction), from small hand-made counter CSVs."""
import csv
import json

import pytest

from newrecon_amd import profiling

FIELDS = ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"]
K = "void ostr::(anonymous namespace)::update_fused_kernel<2, float, float>(ostr::DevGeom)"
COPY = "__amd_rocclr_copyBuffer"


def write_run(d, rows):
    d.mkdir()
    with open(d / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for name, grid, counter, value in rows:
            w.writerow({"Kernel_Name": name, "Grid_Size": grid, "Counter_Name": counter,
                        "Counter_Value": value})


def test_hbm_scales_each_counter_by_the_copy(tmp_path):
    # the 256 MiB copy reads 131072 KiB and writes 262144 KiB: fetch scale 2, write scale 1
    write_run(tmp_path / "f", [(COPY, 131072, "FETCH_SIZE", 131072.0),
                               (K, 1000, "FETCH_SIZE", 100.0), (K, 1000, "FETCH_SIZE", 300.0)])
    write_run(tmp_path / "w", [(COPY, 131072, "WRITE_SIZE", 262144.0),
                               (K, 1000, "WRITE_SIZE", 50.0)])
    text, data = profiling.hbm_table(str(tmp_path / "f"), str(tmp_path / "w"))
    assert data["fetch_scale"] == pytest.approx(2.0) and data["write_scale"] == pytest.approx(1.0)
    k = data["kernels"]["update_fused_kernel<2, float, float>"]
    assert k["fetch_kib"] == pytest.approx(200.0)  # mean over dispatches
    assert k["traffic_bytes"] == pytest.approx((200.0 * 2 + 50.0) * 1024)
    assert "update_fused_kernel<2, float, float>" in text


def test_hbm_requires_the_calibration_copy(tmp_path):
    write_run(tmp_path / "f", [(K, 1000, "FETCH_SIZE", 100.0)])
    write_run(tmp_path / "w", [(K, 1000, "WRITE_SIZE", 50.0)])
    with pytest.raises(RuntimeError):
        profiling.hbm_table(str(tmp_path / "f"), str(tmp_path / "w"))


def test_valu_shares(tmp_path):
    cycles = 1000.0
    # GRBM_GUI_ACTIVE is summed over the 8 XCDs
    write_run(tmp_path / "g", [(K, 1000, "GRBM_GUI_ACTIVE", 8 * cycles)])
    write_run(tmp_path / "s", [(K, 1000, "SQ_INSTS_VALU", 0.5 * 256 * cycles),
                               (K, 1000, "SQ_ACTIVE_INST_VALU", 0.75 * 256 * cycles),
                               (K, 1000, "SQ_WAVES", 64.0), (K, 1000, "SQ_INSTS_SALU", 7.0),
                               (K, 1000, "SQ_INSTS_VMEM_RD", 3.0)])
    text, data = profiling.valu_table(str(tmp_path / "s"), str(tmp_path / "g"))
    k = data["kernels"]["update_fused_kernel<2, float, float>"]
    assert k["cycles"] == pytest.approx(cycles)
    assert k["valu_issue"] == pytest.approx(0.5) and k["valu_busy"] == pytest.approx(0.75)
    assert k["waves"] == 64.0 and k["salu_instr"] == 7.0 and k["vmem_rd_instr"] == 3.0
    assert "| 0.50 | 0.75 |" in text


def test_committed_summaries_name_the_bench_kernels():
    """bench.py looks the dominant kernel up in the committed JSON sidecars."""
    import os
    import bench
    for path in (bench.PROFILE_JSON, bench.VALU_JSON):
        kernels = json.load(open(path))["kernels"]
        for name in bench.PROFILED_KERNEL.values():
            assert name in kernels, (os.path.basename(path), name)
    issue, busy = bench.profiled_valu("update_fused")
    # busy counts the second issue slot of rsq / rcp and the division helpers,
    # so it can pass 1 (never 2: every instruction takes at most two slots)
    assert 0.0 < issue <= busy < 2.0
