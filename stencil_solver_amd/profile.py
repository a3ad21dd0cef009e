"""rocprofv3 wrapper for the bench workloads (the role of the reference's
scripts/profile_nsight.sh): one kernel-trace pass and counter passes over
``bench.py``, summarised per kernel and launch grid.

    python -m stencil_solver_amd.profile --out DIR --name NAME [--kernel REGEX] -- <bench.py arguments>

* trace pass: ``rocprofv3 --kernel-trace --stats`` -> per kernel AND launch
  grid the calls and the average / min / max duration, from the per-dispatch
  trace (one row per distinct grid; launches of one kernel on one grid whose
  durations jump by more than 1.6x -- different work with the same
  workgroup count -- are split into rows too);
* traffic passes: ``rocprofv3 --pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE``, each
  in a run of its own, with ``--calibrate`` appended to the bench arguments
  so the same run holds a STREAM copy of known size (n^3 floats read and
  written); the counters of the profiled kernel (its busiest grid) are scaled
  by bytes / counter of that copy;
* ``--counters A,B,..`` adds one more pass per set (e.g. the VALU counters
  SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU, GRBM_GUI_ACTIVE), reported per dispatch
  of the profiled kernel, with derived figures where their counters are there
  (GRBM_GUI_ACTIVE sums the busy cycles of the XCDs): GPU clock =
  GRBM_GUI_ACTIVE / XCDs / traced duration; VALU issue share = 4 *
  SQ_INSTS_VALU / (GRBM_GUI_ACTIVE / XCDs * 4 SIMDs * CUs) (a wave64 FP32 VALU
  instruction holds a 16-lane SIMD for 4 cycles; packed FP32 is not used);
  VALU instructions per point = 64 * SQ_INSTS_VALU / --points.

``--rate REGEX=BYTES`` prices the rows of the kernels matching REGEX at
BYTES algorithmic bytes per launch (GB/s and fraction of 8 TB/s).  Only the
row of each such kernel with the most workgroups gets it (the full grid;
among rows of equal grids the longest), unless a workgroup count is given:
``REGEX=BYTES@WG``.

Writes ``DIR/NAME.md`` (commands, kernel table, traffic, counters) and
``DIR/traffic_NAME.json`` (the bytes per launch ``bench.py`` reports as
``roofline.traffic``), and deletes the raw per-dispatch files unless --keep.
Every pass runs as a child process under its own time limit.
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import os
import re
import shlex
import shutil
import subprocess
import sys
import time
from collections import defaultdict
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
HBM_PEAK_GBS = 8000.0


def _run(cmd, cwd, timeout, log):
    print(f"[profile] {' '.join(cmd)}", flush=True)
    env = dict(os.environ, TMPDIR="/tmp")
    t0 = time.time()
    with open(log, "w") as f:
        try:
            r = subprocess.run(cmd, cwd=cwd, env=env, stdout=f, stderr=subprocess.STDOUT, timeout=timeout)
        except subprocess.TimeoutExpired:
            raise SystemExit(f"[profile] timed out after {timeout}s: see {log}")
    print(f"[profile] rc={r.returncode} in {time.time() - t0:.1f}s", flush=True)
    return r.returncode


def _find(d: Path, suffix: str):
    hits = sorted(d.rglob(f"*{suffix}"))
    return hits[0] if hits else None


def _col(row: dict, *names):
    """The value of the first column present, matched without case."""
    low = {k.lower(): v for k, v in row.items()}
    for n in names:
        if n.lower() in low and low[n.lower()] not in (None, ""):
            return low[n.lower()]
    return None


def _work_items(row: dict) -> int:
    """Total work-items of a dispatch: Grid_Size, or Grid_Size_X*Y*Z."""
    g = _col(row, "Grid_Size")
    if g is not None:
        return int(float(g))
    prod = 1
    for a in ("X", "Y", "Z"):
        v = _col(row, f"Grid_Size_{a}")
        prod *= int(float(v)) if v is not None else 1
    return prod


def _wg_size(row: dict) -> int:
    w = _col(row, "Workgroup_Size")
    if w is not None:
        return max(1, int(float(w)))
    prod = 1
    for a in ("X", "Y", "Z"):
        v = _col(row, f"Workgroup_Size_{a}")
        prod *= int(float(v)) if v is not None else 1
    return max(1, prod)


SPLIT_RATIO = 1.6


def _duration_clusters(durs):
    """Durations of one kernel on one grid, split where the sorted values jump
    by more than SPLIT_RATIO: the same grid running different work (a slab
    launched with the full grid's workgroup count) lands in separate rows."""
    v = sorted(durs)
    out, cur = [], [v[0]]
    for x in v[1:]:
        if x > SPLIT_RATIO * cur[-1]:
            out.append(cur)
            cur = []
        cur.append(x)
    out.append(cur)
    return out


def _trace_rows(d: Path):
    """Rows per (kernel, grid): calls and durations from the per-dispatch
    trace; the name-only stats table when no trace file is there."""
    tr = _find(d, "kernel_trace.csv")
    groups = defaultdict(list)
    if tr is not None:
        with open(tr) as f:
            for r in csv.DictReader(f):
                name = _col(r, "Kernel_Name")
                t0, t1 = _col(r, "Start_Timestamp"), _col(r, "End_Timestamp")
                if name is None or t0 is None or t1 is None:
                    continue
                items, wg = _work_items(r), _wg_size(r)
                groups[(name, items)].append((float(t1) - float(t0), math.ceil(items / wg)))
    rows = []
    if groups:
        for (name, items), v in groups.items():
            for part, durs in enumerate(_duration_clusters([x[0] for x in v])):
                rows.append({"name": name, "items": items, "workgroups": v[0][1], "part": part,
                             "calls": len(durs), "avg_ns": sum(durs) / len(durs), "min_ns": min(durs),
                             "max_ns": max(durs)})
    else:
        st = _find(d, "kernel_stats.csv")
        if st is None:
            raise SystemExit(f"[profile] no kernel trace or stats under {d}")
        with open(st) as f:
            for r in csv.DictReader(f):
                rows.append({"name": r["Name"], "items": None, "workgroups": None, "calls": int(r["Calls"]),
                             "avg_ns": float(r["AverageNs"]), "min_ns": float(r["MinNs"]),
                             "max_ns": float(r["MaxNs"])})
    rows.sort(key=lambda r: -r["avg_ns"] * r["calls"])
    return rows


def _counter_avgs(d: Path):
    """{counter: {(kernel, work-items): average per dispatch}} and the VGPR
    count per (kernel, work-items)."""
    f = _find(d, "counter_collection.csv")
    if f is None:
        raise SystemExit(f"[profile] no counter_collection.csv under {d}")
    per = defaultdict(lambda: defaultdict(list))
    vgpr = {}
    with open(f) as fh:
        for r in csv.DictReader(fh):
            key = (_col(r, "Kernel_Name"), _work_items(r))
            per[_col(r, "Counter_Name")][key].append(float(_col(r, "Counter_Value")))
            v = _col(r, "VGPR_Count", "Arch_VGPR_Count")
            if v is not None:
                vgpr[key] = int(float(v))
    return {c: {k: sum(v) / len(v) for k, v in t.items()} for c, t in per.items()}, vgpr


def _parse_rates(specs):
    out = []
    for spec in specs:
        pat, _, rest = spec.rpartition("=")
        nb, _, wg = rest.partition("@")
        out.append((re.compile(pat), float(nb), int(wg) if wg else None))
    return out


def _assign_rates(rows, rates):
    """Algorithmic bytes per row: a rate with @WG goes to the rows of that
    workgroup count; one without goes to the row of each matching kernel with
    the most workgroups."""
    assigned = {}
    for pat, nb, wg in rates:
        by_name = defaultdict(list)
        for i, r in enumerate(rows):
            if pat.search(r["name"]):
                by_name[r["name"]].append(i)
        for name, idx in by_name.items():
            if wg is not None:
                targets = [i for i in idx if rows[i]["workgroups"] == wg]
            else:
                top = max(idx, key=lambda i: (rows[i]["workgroups"] or 0, rows[i]["avg_ns"]))
                targets = [top]
            for i in targets:
                assigned.setdefault(i, nb)
    return assigned


def _copy_value(table: dict):
    hits = [(k, v) for k, v in table.items() if k[0] and re.search(r"namespace\)::copy_kernel\(", k[0])]
    if not hits:
        return None
    # the calibration copy is the largest copy in the run
    return max(hits, key=lambda kv: kv[0][1])[1]


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if "--" in argv:
        i = argv.index("--")
        own, bench_args = argv[:i], argv[i + 1:]
    else:
        own, bench_args = argv, []
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True, help="output directory (raw rocprof files go below it)")
    ap.add_argument("--name", required=True, help="summary name, e.g. wave_tiled_512")
    ap.add_argument("--kernel", default=r"zq_kernel", help="regex of the kernel whose traffic is reported")
    ap.add_argument("--algorithmic-bytes", type=float, default=None,
                    help="algorithmic bytes per launch of that kernel (for the traffic ratio)")
    ap.add_argument("--points", type=float, default=None, help="interior points per launch of that kernel")
    ap.add_argument("--cus", type=int, default=256, help="compute units of the device")
    ap.add_argument("--xcds", type=int, default=8, help="XCDs of the device (GRBM_GUI_ACTIVE sums them)")
    ap.add_argument("--no-pmc", action="store_true", help="no FETCH_SIZE / WRITE_SIZE passes")
    ap.add_argument("--counters", action="append", default=[], metavar="A,B,...",
                    help="one more counter pass with these counters (repeatable)")
    ap.add_argument("--rate", action="append", default=[], metavar="REGEX=BYTES[@WG]",
                    help="algorithmic bytes per launch of the kernels matching REGEX (see the module doc)")
    ap.add_argument("--copy-n", type=int, default=None, help="cells per axis of the calibration copy (--n of bench)")
    ap.add_argument("--timeout", type=int, default=600)
    ap.add_argument("--keep", action="store_true", help="keep raw per-dispatch files")
    a = ap.parse_args(own)
    out = Path(a.out).resolve()
    out.mkdir(parents=True, exist_ok=True)
    bench = [sys.executable, "-u", str(ROOT / "bench.py")] + bench_args
    cmds = {}

    tdir = out / f"{a.name}_trace"
    cmd = ["rocprofv3", "--kernel-trace", "--stats", "--output-format", "csv", "-d", str(tdir), "-o", "run", "--"] + bench
    cmds["trace"] = cmd
    if _run(cmd, "/tmp", a.timeout, out / f"{a.name}_trace.log") != 0:
        raise SystemExit("[profile] trace pass failed")
    rows = _trace_rows(tdir)

    passes = [] if a.no_pmc else [("FETCH_SIZE",), ("WRITE_SIZE",)]
    passes += [tuple(c for c in s.split(",") if c) for s in a.counters]
    counters, vgpr, raw_dirs = {}, {}, []
    for pi, cs in enumerate(passes):
        tag = cs[0] if len(cs) == 1 else f"set{pi}"
        cdir = out / f"{a.name}_{tag}"
        raw_dirs.append(cdir)
        cmd = ["rocprofv3", "--pmc", *cs, "--output-format", "csv", "-d", str(cdir), "-o", "run", "--"] + bench + \
            ["--calibrate"]
        cmds[" ".join(cs)] = cmd
        if _run(cmd, "/tmp", a.timeout, out / f"{a.name}_{tag}.log") != 0:
            raise SystemExit(f"[profile] counter pass {cs} failed")
        got, vg = _counter_avgs(cdir)
        counters.update(got)
        vgpr.update(vg)

    # the profiled kernel: the (kernel, grid) row of the regex with the most time
    kcand = [r for r in rows if re.search(a.kernel, r["name"])]
    krow = max(kcand, key=lambda r: r["avg_ns"] * r["calls"]) if kcand else None
    kkey = (krow["name"], krow["items"]) if krow else None

    envs = " ".join(f"{k}={shlex.quote(v)}" for k, v in sorted(os.environ.items()) if k.startswith("STENCIL_"))
    wrapper = (envs + " " if envs else "") + "python -m stencil_solver_amd.profile " + " ".join(
        shlex.quote(x) for x in argv)
    lines = [f"# rocprofv3 summary: {a.name}", "",
             "Made on one MI355X from the repository root by", "", "```", wrapper, "```", "",
             "which ran these passes:", "", "```"]
    for c in cmds.values():
        lines.append(" ".join(c).replace(str(ROOT) + "/", ""))
    lines += ["```", "", "## Kernels (trace pass)", "",
              "One row per kernel and launch grid (workgroups per dispatch); launches of one kernel and grid whose "
              f"durations jump by more than {SPLIT_RATIO}x are separate rows."]
    rates = _parse_rates(a.rate)
    assigned = _assign_rates(rows, rates)
    if rates:
        lines += ["Algorithmic bytes per launch from `--rate` (the rate goes to the kernel's largest grid unless "
                  "a workgroup count is given); GB/s = those bytes / average duration; fraction of the 8 TB/s "
                  "HBM peak.", "",
                  "| Kernel | Workgroups | Calls | Average us | Min us | Max us | Algorithmic GB | GB/s | of 8 TB/s |",
                  "|---|---|---|---|---|---|---|---|---|"]
    else:
        lines += ["", "| Kernel | Workgroups | Calls | Average us | Min us | Max us |", "|---|---|---|---|---|---|"]
    for i, r in enumerate(rows[:28 if rates else 14]):
        cell = (f"| `{r['name'][:110]}` | {r['workgroups'] if r['workgroups'] is not None else '-'} | {r['calls']} | "
                f"{r['avg_ns'] / 1e3:.1f} | {r['min_ns'] / 1e3:.1f} | {r['max_ns'] / 1e3:.1f} |")
        if rates:
            nb = assigned.get(i)
            if nb is not None:
                gbs = nb / r["avg_ns"]
                cell += f" {nb / 1e9:.4f} | {gbs:.0f} | {gbs / HBM_PEAK_GBS:.3f} |"
            else:
                cell += " | | |"
        lines.append(cell)

    traffic = {"kernel": kkey[0] if kkey else None, "work_items": kkey[1] if kkey else None,
               "avg_kernel_ns": krow["avg_ns"] if krow else None}
    n = a.copy_n
    if n is None:
        n = int(bench_args[bench_args.index("--n") + 1]) if "--n" in bench_args else 512
    copy_bytes = 4.0 * n ** 3
    if kkey and "FETCH_SIZE" in counters and "WRITE_SIZE" in counters:
        lines += ["", "## HBM traffic (one counter per run, KiB per dispatch)", "",
                  f"Profiled kernel: `{kkey[0][:110]}`, {krow['workgroups']} workgroups.", "",
                  "| Counter | profiled kernel | STREAM copy of %.2f MB | scale (bytes / KiB read) |" % (copy_bytes / 1e6),
                  "|---|---|---|---|"]
        total = 0.0
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            kval = counters[c].get(kkey)
            cval = _copy_value(counters[c])
            if not kval or not cval:
                lines.append(f"| {c} | {kval} | {cval} | n/a |")
                continue
            scale = copy_bytes / (cval * 1024.0)
            b = kval * 1024.0 * scale
            total += b
            traffic[c.lower()] = {"kib_per_dispatch": kval, "copy_kib": cval, "scale": round(scale, 4), "bytes": b}
            lines.append(f"| {c} | {kval:.1f} | {cval:.1f} | x{scale:.3f} |")
        traffic["bytes_per_launch"] = total
        lines += ["", f"Scaled HBM bytes per launch of the profiled kernel: **{total / 1e9:.4f} GB**."]
        if a.algorithmic_bytes:
            traffic["algorithmic_bytes"] = a.algorithmic_bytes
            traffic["traffic_over_algorithmic"] = round(total / a.algorithmic_bytes, 4)
            lines.append(f"Algorithmic {a.algorithmic_bytes / 1e9:.4f} GB: ratio **{total / a.algorithmic_bytes:.3f}**.")
        if total:
            lines.append(f"At the traced average {krow['avg_ns'] / 1e3:.1f} us: {total / krow['avg_ns']:.0f} GB/s of "
                         f"measured traffic.")

    extra = [c for c in counters if c not in ("FETCH_SIZE", "WRITE_SIZE")]
    if kkey and extra:
        lines += ["", "## Counters of the profiled kernel (average per dispatch)", "",
                  f"`{kkey[0][:110]}`, {krow['workgroups']} workgroups, traced average {krow['avg_ns'] / 1e3:.1f} us"
                  + (f", {vgpr[kkey]} VGPRs" if kkey in vgpr else "") + ".", "",
                  "| Counter | per dispatch |", "|---|---|"]
        vals = {}
        for c in sorted(extra):
            v = counters[c].get(kkey)
            vals[c] = v
            lines.append(f"| {c} | {v:.4g} |" if v is not None else f"| {c} | n/a |")
        traffic["counters"] = vals
        derived = []
        gui, valu = vals.get("GRBM_GUI_ACTIVE"), vals.get("SQ_INSTS_VALU")
        if gui:
            cyc = gui / a.xcds
            clk = cyc / krow["avg_ns"]
            derived.append(f"GPU clock during the kernel: GRBM_GUI_ACTIVE / {a.xcds} XCDs = {cyc:.4g} cycles in "
                           f"{krow['avg_ns'] / 1e3:.1f} us = **{clk:.2f} GHz**")
            traffic["clock_ghz"] = round(clk, 3)
        if gui and valu:
            share = 4.0 * valu / (gui / a.xcds * 4.0 * a.cus)
            derived.append(f"VALU issue share: 4 cycles x SQ_INSTS_VALU / (those cycles x 4 SIMDs x {a.cus} CUs) = "
                           f"**{share:.3f}**")
            traffic["valu_issue_share"] = round(share, 4)
        if valu and a.points:
            per = 64.0 * valu / a.points
            derived.append(f"VALU instructions per interior point: 64 x SQ_INSTS_VALU / {a.points:.6g} = **{per:.1f}**")
            traffic["valu_per_point"] = round(per, 2)
        if derived:
            lines += [""] + [f"* {d}" for d in derived]
    (out / f"{a.name}.md").write_text("\n".join(lines) + "\n")
    (out / f"traffic_{a.name}.json").write_text(json.dumps(traffic, indent=1) + "\n")
    print("\n".join(lines), flush=True)
    if not a.keep:
        for d in [tdir] + raw_dirs:
            shutil.rmtree(d, ignore_errors=True)


if __name__ == "__main__":
    main()
