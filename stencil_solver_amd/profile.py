"""rocprofv3 wrapper for the bench workloads (the role of the reference's
scripts/profile_nsight.sh): one kernel-trace pass and one counter pass per
HBM counter over ``bench.py``, summarised per kernel.

    python -m stencil_solver_amd.profile --out DIR --name NAME [--kernel REGEX] -- <bench.py arguments>

* trace pass: ``rocprofv3 --kernel-trace --stats`` -> calls and average
  duration of every kernel;
* counter passes: ``rocprofv3 --pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE``, each
  in a run of its own, with ``--calibrate`` appended to the bench arguments
  so the same run holds a STREAM copy of known size (n^3 floats read and
  written); the counters of the profiled kernel are scaled by
  bytes / counter of that copy.

Writes ``DIR/NAME.md`` (commands, kernel table, traffic) and
``DIR/traffic_NAME.json`` (the bytes per launch ``bench.py`` reports as
``roofline.traffic``), and deletes the raw per-dispatch files unless --keep.
Every pass runs as a child process under its own time limit.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import re
import shutil
import subprocess
import sys
import time
from collections import defaultdict
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


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


def _find(d: Path, suffix: str) -> Path:
    hits = sorted(d.rglob(f"*{suffix}"))
    if not hits:
        raise SystemExit(f"[profile] no *{suffix} under {d}")
    return hits[0]


def _kernel_stats(d: Path):
    rows = []
    with open(_find(d, "kernel_stats.csv")) as f:
        for r in csv.DictReader(f):
            rows.append({"name": r["Name"], "calls": int(r["Calls"]), "avg_ns": float(r["AverageNs"]),
                         "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"]),
                         "pct": float(r.get("Percentage", 0.0))})
    rows.sort(key=lambda r: -r["avg_ns"] * r["calls"])
    return rows


def _counter_avgs(d: Path, counter: str):
    per = defaultdict(list)
    with open(_find(d, "counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                per[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in per.items()}


def _pick(table: dict, pattern: str):
    hits = {k: v for k, v in table.items() if re.search(pattern, k)}
    if not hits:
        return None, None
    k = max(hits, key=lambda k: hits[k])
    return k, hits[k]


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
    ap.add_argument("--no-pmc", action="store_true", help="trace pass only")
    ap.add_argument("--rate", action="append", default=[], metavar="REGEX=BYTES",
                    help="algorithmic bytes per launch of the kernels matching REGEX: adds their "
                         "GB/s and fraction of the 8 TB/s peak to the kernel table (repeatable)")
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
    stats = _kernel_stats(tdir)

    counters = {}
    if not a.no_pmc:
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            cdir = out / f"{a.name}_{c}"
            cmd = ["rocprofv3", "--pmc", c, "--output-format", "csv", "-d", str(cdir), "-o", "run", "--"] + bench + \
                ["--calibrate"]
            cmds[c] = cmd
            if _run(cmd, "/tmp", a.timeout, out / f"{a.name}_{c}.log") != 0:
                raise SystemExit(f"[profile] {c} pass failed")
            counters[c] = _counter_avgs(cdir, c)

    kname, _ = _pick({r["name"]: r["avg_ns"] * r["calls"] for r in stats}, a.kernel)
    krow = next((r for r in stats if r["name"] == kname), None)
    import shlex

    # the library's own environment switches are part of the command
    envs = " ".join(f"{k}={shlex.quote(v)}" for k, v in sorted(os.environ.items()) if k.startswith("STENCIL_"))
    wrapper = (envs + " " if envs else "") + "python -m stencil_solver_amd.profile " + " ".join(
        shlex.quote(x) for x in argv)
    lines = [f"# rocprofv3 summary: {a.name}", "",
             "Made on one MI355X from the repository root by", "", "```", wrapper, "```", "",
             "which ran these passes:", "", "```"]
    for k, c in cmds.items():
        lines.append(" ".join(c).replace(str(ROOT) + "/", ""))
    rates = []
    for spec in a.rate:
        pat, _, nb = spec.rpartition("=")
        rates.append((re.compile(pat), float(nb)))

    def rate_cells(name, avg_ns):
        for pat, nb in rates:
            if pat.search(name):
                gbs = nb / avg_ns
                return f" {nb / 1e9:.4f} | {gbs:.0f} | {gbs / 8000.0:.3f} |"
        return " | | |"

    lines += ["```", "", "## Kernels (trace pass)", ""]
    if rates:
        lines += ["Algorithmic bytes per launch as given by `--rate`; GB/s = those bytes / average duration; "
                  "fraction of the 8 TB/s HBM peak.", "",
                  "| Kernel | Calls | Average us | Min us | Max us | Algorithmic GB | GB/s | of 8 TB/s |",
                  "|---|---|---|---|---|---|---|---|"]
    else:
        lines += ["| Kernel | Calls | Average us | Min us | Max us |", "|---|---|---|---|---|"]
    for r in stats[:24 if rates else 12]:
        row = (f"| `{r['name'][:110]}` | {r['calls']} | {r['avg_ns'] / 1e3:.1f} | {r['min_ns'] / 1e3:.1f} | "
               f"{r['max_ns'] / 1e3:.1f} |")
        lines.append(row + (rate_cells(r["name"], r["avg_ns"]) if rates else ""))
    traffic = {"kernel": kname, "avg_kernel_ns": krow["avg_ns"] if krow else None}
    if counters:
        n = a.copy_n
        if n is None:
            n = 512
            if "--n" in bench_args:
                n = int(bench_args[bench_args.index("--n") + 1])
        copy_bytes = 4.0 * n ** 3
        lines += ["", "## HBM traffic (one counter per run, KiB per dispatch)", "",
                  "| Counter | profiled kernel | STREAM copy of %.2f MB | scale (bytes / KiB read) |" % (copy_bytes / 1e6),
                  "|---|---|---|---|"]
        total = 0.0
        for c, table in counters.items():
            _, kval = _pick({k: v for k, v in table.items() if re.search(a.kernel, k)} or {"": 0.0}, ".")
            _, cval = _pick(table, r"namespace\)::copy_kernel\(")
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
        if krow and total:
            lines.append(f"At the traced average {krow['avg_ns'] / 1e3:.1f} us: {total / krow['avg_ns']:.0f} GB/s of "
                         f"measured traffic.")
    (out / f"{a.name}.md").write_text("\n".join(lines) + "\n")
    (out / f"traffic_{a.name}.json").write_text(json.dumps(traffic, indent=1) + "\n")
    print("\n".join(lines), flush=True)
    if not a.keep:
        for d in [tdir] + [out / f"{a.name}_{c}" for c in counters]:
            shutil.rmtree(d, ignore_errors=True)


if __name__ == "__main__":
    main()
