"""The rocprofv3 summary tool (stencil_solver_amd/profile.py) on synthetic
per-dispatch files: one row per kernel and launch grid, rates only on the
grid they price, counters matched to the profiled kernel's grid."""
from __future__ import annotations

import csv

import pytest

from stencil_solver_amd import profile as P


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_trace_rows_split_by_grid(tmp_path):
    hdr = ["Kind", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "Workgroup_Size_X", "Workgroup_Size_Y",
           "Workgroup_Size_Z", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]
    rows = []
    t = 0
    for _ in range(10):  # full grid: 252 workgroups of 512, 340 us
        rows.append(["K", "zq<64,8>", t, t + 340_000, 512, 1, 1, 252 * 512, 1, 1])
        t += 400_000
    for _ in range(30):  # slab: 126 workgroups, 48 us
        rows.append(["K", "zq<64,8>", t, t + 48_000, 512, 1, 1, 126 * 512, 1, 1])
        t += 60_000
    rows.append(["K", "copy", t, t + 80_000, 256, 1, 1, 1 << 20, 1, 1])
    _write(tmp_path / "run_kernel_trace.csv", hdr, rows)
    got = P._trace_rows(tmp_path)
    zq = sorted((r for r in got if r["name"] == "zq<64,8>"), key=lambda r: r["workgroups"])
    assert [(r["workgroups"], r["calls"]) for r in zq] == [(126, 30), (252, 10)]
    assert zq[0]["avg_ns"] == pytest.approx(48_000) and zq[1]["avg_ns"] == pytest.approx(340_000)
    # an unqualified rate prices the full grid only; @126 prices the slab
    rates = P._parse_rates([r"zq<64=2.048e9", r"zq<64=0.28e9@126"])
    assigned = P._assign_rates(got, rates)
    full = got.index(zq[1])
    slab = got.index(zq[0])
    assert assigned[full] == pytest.approx(2.048e9) and assigned[slab] == pytest.approx(0.28e9)
    assert all(i in (full, slab) for i in assigned)


def test_counters_are_matched_to_a_grid(tmp_path):
    hdr = ["Kernel_Name", "Grid_Size", "Workgroup_Size", "VGPR_Count", "Counter_Name", "Counter_Value"]
    rows = [["zq<64,8>", 252 * 512, 512, 120, "FETCH_SIZE", 1000.0],
            ["zq<64,8>", 252 * 512, 512, 120, "FETCH_SIZE", 1200.0],
            ["zq<64,8>", 126 * 512, 512, 120, "FETCH_SIZE", 300.0],
            ["copy", 1 << 20, 256, 16, "FETCH_SIZE", 500.0]]
    _write(tmp_path / "run_counter_collection.csv", hdr, rows)
    c, vgpr = P._counter_avgs(tmp_path)
    assert c["FETCH_SIZE"][("zq<64,8>", 252 * 512)] == pytest.approx(1100.0)
    assert c["FETCH_SIZE"][("zq<64,8>", 126 * 512)] == pytest.approx(300.0)
    assert P._copy_value({("x::(anonymous namespace)::copy_kernel(f)", 10): 5.0,
                          ("x::(anonymous namespace)::copy_kernel(f)", 99): 7.0}) == 7.0
    assert vgpr[("zq<64,8>", 252 * 512)] == 120


def test_same_grid_different_work_is_split(tmp_path):
    """A 72-plane slab launched with the full grid's workgroup count: its
    launches form their own row, and the rate goes to the longer one."""
    hdr = ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Workgroup_Size", "Grid_Size"]
    rows, t = [], 0
    for d in [340_000, 352_000, 345_000, 48_000, 49_500, 47_000, 50_000]:
        rows.append(["zq", t, t + d, 512, 252 * 512])
        t += 400_000
    _write(tmp_path / "run_kernel_trace.csv", hdr, rows)
    got = P._trace_rows(tmp_path)
    assert sorted(r["calls"] for r in got) == [3, 4]
    assigned = P._assign_rates(got, P._parse_rates(["zq=2.048e9"]))
    (i,) = assigned
    assert got[i]["calls"] == 3 and got[i]["avg_ns"] > 300_000
