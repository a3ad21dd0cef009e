"""Benchmark of the 8th-order acoustic time step on one MI355X.

Workload (one "step" = one launch of the time-step kernel over the grid):
acoustic wave, 512^3 cells (504^3 interior), h = 10 m, dt = 1e-3 s, FP32,
seeded 16-layer velocity model v ~ U[1500, 4000] m/s per layer with +-5 %
per-cell noise, centred Gaussian pulse (sigma = 3 cells) in p and p_prev.
Inputs are resident in HBM before the timed region; K steps are timed with
HIP events on the stream the kernels are launched on.

Prints ONE JSON line.  ``value`` is interior points updated per second (Gpts/s)
of the tiled kernel; ``roofline`` prices it against 8 TB/s with the algorithmic
16 B / point (read p, p_prev, vel2; write p_next); ``cpu_baseline`` times the
reference's own CPU path (oracle/_ref, OpenMP on the host cores) on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Gpts/s @512³: FP32 plain vs BF16-store (rel-L2≤1e-3 vs FP64); grad time"
HBM_PEAK_GBS = 8000.0
REF_GPTS_RTX4050 = 11.17  # reference naive FP32 at 512^3: 11.463 ms/step (docs/CASE_STUDY.md:86)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512, help="cells per axis (halo included)")
    ap.add_argument("--variant", default="tiled", choices=["tiled", "naive", "tiled_fused"])
    ap.add_argument("--storage", default="fp32", choices=["fp32", "bf16", "fp16", "bfp16"],
                    help="field storage of the timed loop: FP32, or 16-bit storage with FP32 "
                         "arithmetic (tiled kernel, fused-weight form)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the side measurements")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target CPU-baseline sample length per row (two rows: no FTZ/DAZ, FTZ/DAZ)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library default shape)")
    ap.add_argument("--shape", default=None, metavar="TX,TY,CHUNKS,XCD",
                    help="FP32 tiled step at this launch shape (tile_x cells, tile_y rows, z-chunks, XCD "
                         "remap 0/1) instead of the tuned one; implies --no-tune for that step")
    ap.add_argument("--no-precision", action="store_true", help="skip the 16-bit storage / FP64 study")
    ap.add_argument("--precision-n", type=int, nargs="*", default=None,
                    help="grid sizes of the 16-bit storage study (default: --n, and 1024 when --n is 512)")
    ap.add_argument("--precision-steps", type=int, default=1000,
                    help="steps of the 16-bit storage accuracy runs (the error bar is stated after 1000 steps; "
                         "independent of --steps)")
    ap.add_argument("--no-fwi", action="store_true", help="skip the adjoint FWI gradient timing")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-shot timing")
    ap.add_argument("--batch-shots", type=int, default=64, help="shots of the batched modelling")
    ap.add_argument("--batch-steps", type=int, default=1500, help="steps of the batched modelling")
    ap.add_argument("--batch-plan", default=None, metavar="TILE_X,TILE_Y,CHUNK,XCD_RUNS",
                    help="fixed ragged-batch plan (cells, rows, z-chunk planes, 0/1) instead of the batch tuner's, "
                         "for both the per-step and the persistent run (profiles fix it so every pass runs one plan)")
    ap.add_argument("--fwi-n", type=int, default=256, help="FWI grid cells per axis")
    ap.add_argument("--fwi-steps", type=int, default=2000, help="FWI time steps")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run 2 STREAM copies of one field (known bytes) for counter scaling")
    ap.add_argument("--tune-csv", default=None, help="write the autotuner sweep to this CSV")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the step runs one grid on one GPU, with no halo exchange).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import stencil_solver_amd as ss

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    n = args.n
    shape = (n, n, n)
    h, dt = np.float32(10.0), np.float32(1e-3)
    grid = ss.GridDims.cube(n, h, dt)
    interior = (n - 8) ** 3

    t0 = time.time()
    p0_h = ss.host_gaussian(shape, 0.5, 0.5, 0.5, 3.0)
    vel2_h = ss.host_layered_noise_vel2(shape, 16, 1500.0, 4000.0, 0.05, 2)
    p0 = torch.from_numpy(p0_h).to(dev)
    vel2 = torch.from_numpy(vel2_h).to(dev)
    print(f"[bench] inputs {n}^3 ready in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    def launcher(variant):
        if variant == "naive":
            return lambda c, p, v, o: ss.launch_naive(c, p, v, o, grid)
        if variant == "tiled_fused":
            return lambda c, p, v, o: ss.launch_tiled_ex(c, p, v, o, grid, exact=False)
        return lambda c, p, v, o: ss.launch_tiled(c, p, v, o, grid)

    def timed_run16(fmt, steps, warmup):
        vs = ss.pow2_scale_for(float(vel2.max())) if fmt == "fp16" else 1.0
        gv = ss.encode16(fmt, vel2, ss.Field16(fmt, shape, dev, vs))
        bufs = [ss.encode16(fmt, x, ss.Field16(fmt, shape, dev, 1.0)) for x in (p0, p0, torch.zeros_like(p0))]
        if not args.no_tune:
            tb = ss.Field16(fmt, shape, dev, 1.0)
            print(f"[bench] tuned {fmt} shape {ss.tune_tiled_16(bufs[1], bufs[0], gv, tb, grid, reps=5)}",
                  file=sys.stderr, flush=True)
            del tb
        i = 0
        for _ in range(warmup):
            ss.launch_tiled_16(bufs[(i + 1) % 3], bufs[i], gv, bufs[(i + 2) % 3], grid)
            i = (i + 1) % 3
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(steps):
            ss.launch_tiled_16(bufs[(i + 1) % 3], bufs[i], gv, bufs[(i + 2) % 3], grid)
            i = (i + 1) % 3
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / steps
        out = ss.decode16(bufs[(i + 1) % 3])
        del bufs, gv
        return ms, out

    def timed_run(variant, steps, warmup):
        fn = launcher(variant)
        bufs = [p0.clone(), p0.clone(), torch.zeros_like(p0)]
        i = 0
        for _ in range(warmup):
            fn(bufs[(i + 1) % 3], bufs[i], vel2, bufs[(i + 2) % 3])
            i = (i + 1) % 3
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()  # the current stream: the one the launchers use
        for _ in range(steps):
            fn(bufs[(i + 1) % 3], bufs[i], vel2, bufs[(i + 2) % 3])
            i = (i + 1) % 3
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / steps
        out = bufs[(i + 1) % 3]
        del bufs
        return ms, out

    tuned = None
    if args.shape:
        if args.variant != "tiled" or args.storage != "fp32":
            sys.exit("--shape applies to the FP32 tiled step")
        tx, ty, nc, xr = (int(v) for v in args.shape.split(","))
        ss.set_tuned_shape(grid, "wave", {"tile_x": tx, "tile_y": ty, "n_chunks": nc, "xcd_remap": xr,
                                          "nontemporal": 1})
        print(f"[bench] fixed shape {ss.tuned_shape(grid, 'wave')[0]}", file=sys.stderr, flush=True)
    elif args.variant == "tiled" and not args.no_tune:
        # One-time autotune of the launch shape for this grid size (outside the
        # timed region; the launcher caches the result).
        tb = [p0.clone(), p0.clone(), torch.zeros_like(p0)]  # distinct buffers, as in the run
        tuned = ss.tune_tiled(tb[1], tb[0], vel2, tb[2], grid, reps=5, csv_path=args.tune_csv)
        del tb
        print(f"[bench] tuned shape {tuned}", file=sys.stderr, flush=True)
    shape_used, _ = ss.tuned_shape(grid, "wave" if args.storage == "fp32" else args.storage)

    if args.storage == "fp32":
        ms, out = timed_run(args.variant, args.steps, args.warmup)
        bytes_pt = 16.0
    else:
        if args.variant != "tiled":
            sys.exit("--storage applies to the tiled kernel only")
        ms, out = timed_run16(args.storage, args.steps, args.warmup)
        bytes_pt = 8.25 if args.storage == "bfp16" else 8.0  # + 1 exponent byte / 16 values
        args.no_extras = True
    gpts = interior / (ms * 1e-3) / 1e9
    alg_bytes = bytes_pt * interior
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    print(f"[bench] {args.variant} ({args.storage}): {ms:.4f} ms/step, {gpts:.1f} Gpts/s, {achieved:.0f} GB/s",
          file=sys.stderr, flush=True)

    result = {
        "metric": METRIC,
        "value": round(gpts, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(gpts / REF_GPTS_RTX4050, 2),
        "baseline": {"value": REF_GPTS_RTX4050, "unit": "Gpts/s",
                     "what": "reference naive FP32 kernel, 512^3, RTX 4050 Laptop (11.463 ms/step)"},
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"acoustic {n}^3 FP32 leapfrog, 8th-order Laplacian, 16-layer +-5% velocity, "
                               f"Gaussian sigma=3, h=10 m, dt=1e-3 s",
                   "grid": n, "interior_points": interior, "kernel": args.variant, "storage": args.storage,
                   "tile_shape": shape_used if args.variant == "tiled" else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": alg_bytes},
    }

    # Measured HBM bytes per launch (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE,
    # scaled by a known-size copy), committed under profiles/ for this kernel.
    tag = f"wave_{args.variant}_{n}" if args.storage == "fp32" else f"wave16_{args.storage}_{n}"
    tfile = ROOT / "profiles" / f"traffic_{tag}.json"
    if tfile.exists():
        t = json.loads(tfile.read_text())
        result["roofline"]["traffic"] = t.get("bytes_per_launch")

    if not args.no_extras:
        extras = {}
        # Same steps through the naive (reference-association) kernel: the
        # measured run's result must agree with it (tolerance 1e-5 rel-L2).
        ms_n, out_n = timed_run("naive", args.steps, args.warmup)
        rel = (torch.linalg.vector_norm((out - out_n).double()) /
               torch.linalg.vector_norm(out_n.double())).item()
        extras["naive_gpts"] = round(interior / (ms_n * 1e-3) / 1e9, 2)
        extras["rel_l2_vs_naive_after_steps"] = rel
        del out_n
        ms_e, _ = timed_run("tiled_fused", min(args.steps, 200), 5)
        extras["tiled_fused_gpts"] = round(interior / (ms_e * 1e-3) / 1e9, 2)
        # Heat step (dedicated single-level kernel): 8 B / point, on the heat
        # workload's own field (U[0,1] noise + a hot cube of side 64).  The
        # reference association costs ~55 VALU per point, so this step is
        # VALU-bound and its rate depends on the data (the mostly-zero pulse
        # field runs ~20 % faster); the fused-weight form is timed beside it.
        a = torch.from_numpy(ss.host_noise_cube((n, n, n), 64, 10.0, 2)).to(dev)
        b = torch.zeros_like(a)
        hgrid = ss.GridDims.cube(n, h, np.float32(np.sqrt(np.float64(5.0))))
        if not args.no_tune:
            extras["heat_shape"] = ss.tune_heat(a, b, hgrid, reps=3)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for key, exact in (("heat_tiled", None), ("heat_tiled_fused", False)):
            for _ in range(5):
                ss.heat_step(a, b, 1.0, hgrid, exact=exact)
                a, b = b, a
            torch.cuda.synchronize()
            e0.record()
            for _ in range(200):
                ss.heat_step(a, b, 1.0, hgrid, exact=exact)
                a, b = b, a
            e1.record()
            torch.cuda.synchronize()
            hms = e0.elapsed_time(e1) / 200
            extras[f"{key}_gbs"] = round(8.0 * interior / (hms * 1e-3) / 1e9, 1)
            extras[f"{key}_ms"] = round(hms, 5)
        del a, b
        # STREAM copy of one field-sized buffer: achievable HBM bandwidth.
        src = torch.ones(n * n * n, device=dev)
        dst = torch.empty_like(src)
        L = ss.lib()
        stream = torch.cuda.current_stream().cuda_stream
        for _ in range(3):
            L.stencil_stream_f32(0, dst.data_ptr(), src.data_ptr(), src.numel(), stream)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(20):
            L.stencil_stream_f32(0, dst.data_ptr(), src.data_ptr(), src.numel(), stream)
        e1.record()
        torch.cuda.synchronize()
        cms = e0.elapsed_time(e1) / 20
        extras["stream_copy_gbs"] = round(8.0 * src.numel() / (cms * 1e-3) / 1e9, 1)
        del src, dst
        # 3-read / 1-write streaming kernel (the stencil's byte mix, no reuse):
        # the practical ceiling this HBM delivers for 16 B / point.  Run with
        # the three read arrays back to back at n^3 floats (2^29-byte strides)
        # and at n^3 + 1024 floats (a 4 KB stagger, the best of the strides
        # measured); the faster of the two is the ceiling quoted.
        for key, pad in (("stream_3r1w_gbs", 0), ("stream_3r1w_staggered_gbs", 1024)):
            cnt = n * n * n + pad
            src3 = torch.ones(3 * cnt, device=dev)
            dst = torch.empty(cnt, device=dev)
            for _ in range(3):
                L.stencil_stream_f32(3, dst.data_ptr(), src3.data_ptr(), cnt, stream)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(20):
                L.stencil_stream_f32(3, dst.data_ptr(), src3.data_ptr(), cnt, stream)
            e1.record()
            torch.cuda.synchronize()
            tms = e0.elapsed_time(e1) / 20
            extras[key] = round(16.0 * cnt / (tms * 1e-3) / 1e9, 1)
            del src3, dst
        best_3r1w = max(extras["stream_3r1w_gbs"], extras["stream_3r1w_staggered_gbs"])
        # The FP32 step's measured HBM traffic rate (counter bytes per launch
        # from profiles/, this run's time) against that ceiling.
        if result["roofline"]["traffic"]:
            traffic_gbs = result["roofline"]["traffic"] / (ms * 1e-3) / 1e9
            extras["step_traffic_gbs"] = round(traffic_gbs, 1)
            extras["step_traffic_vs_best_3r1w"] = round(traffic_gbs / best_3r1w, 3)
        # One rank's share of an 8-way Z-slab split of this grid (the design
        # for more GPUs): 64 owned planes plus the 4 exchanged planes on each
        # face, the same tiled kernel at that shape (no exchange here).
        z0 = n // 2 - 36
        sgrid = ss.GridDims(n, n, 72, h, h, h, dt)
        sp = p0[z0:z0 + 72].contiguous()
        sv = vel2[z0:z0 + 72].contiguous()
        sb = [sp.clone(), sp.clone(), torch.zeros_like(sp)]
        if not args.no_tune:
            extras["slab_shape"] = ss.tune_tiled(sb[1], sb[0], sv, sb[2], sgrid, reps=5)
        i = 0
        for it in range(520):
            if it == 20:
                torch.cuda.synchronize()
                e0.record()
            ss.launch_tiled(sb[(i + 1) % 3], sb[i], sv, sb[(i + 2) % 3], sgrid)
            i = (i + 1) % 3
        e1.record()
        torch.cuda.synchronize()
        sms = e0.elapsed_time(e1) / 500
        spts = (n - 8) * (n - 8) * 64
        extras["slab_8way_ms"] = round(sms, 5)
        extras["slab_8way_gpts"] = round(spts / (sms * 1e-3) / 1e9, 1)
        del sp, sv, sb
        result["extras"] = extras
        print(f"[bench] extras {extras}", file=sys.stderr, flush=True)

    if not args.no_precision:
        sizes = args.precision_n if args.precision_n else ([n, 1024] if n == 512 else [n])
        result["precision"] = {}
        for pn in sizes:
            if pn == n:
                pp0, pv2 = p0, vel2
            else:
                pshape = (pn, pn, pn)
                pp0 = torch.from_numpy(ss.host_gaussian(pshape, 0.5, 0.5, 0.5, 3.0)).to(dev)
                pv2 = torch.from_numpy(ss.host_layered_noise_vel2(pshape, 16, 1500.0, 4000.0, 0.05, 2)).to(dev)
            result["precision"][str(pn)] = precision_study(ss, pp0, pv2, pn, args.precision_steps,
                                                           tune=not args.no_tune)
            del pp0, pv2
        torch.cuda.empty_cache()
        p512 = result["precision"].get(str(n))
        if p512:
            ok = {k: v for k, v in p512.items() if k in ("bf16", "fp16", "bfp16") and v["rel_l2_vs_fp64"] <= 1e-3}
            best = max(ok, key=lambda k: ok[k]["gpts"]) if ok else None
            result["bf16_store"] = {
                "accuracy_steps": args.precision_steps, "bar": "rel-L2 <= 1e-3 vs the FP64 GPU run",
                "bf16_gpts": p512["bf16"]["gpts"], "bf16_rel_l2_vs_fp64": p512["bf16"]["rel_l2_vs_fp64"],
                "best_16bit_within_1e-3": best,
                "best_16bit_gpts": ok[best]["gpts"] if best else None,
                "speedup_vs_fp32": round(ok[best]["gpts"] / gpts, 3) if best else None}

    if not args.no_fwi:
        result["fwi"] = fwi_study(ss, args.fwi_n, args.fwi_steps, dev, tune=not args.no_tune)
    if not args.no_batch:
        plan = tuple(int(v) for v in args.batch_plan.split(",")) if args.batch_plan else None
        result["batch"] = batch_study(ss, args.batch_shots, args.batch_steps, dev, tune=not args.no_tune, plan=plan)

    if args.calibrate:
        # Known-size copy: n^3 floats read + n^3 floats written per launch, the
        # reference the FETCH_SIZE / WRITE_SIZE counters are scaled against.
        src = torch.ones(n * n * n, device=dev)
        dst = torch.empty_like(src)
        stream = torch.cuda.current_stream().cuda_stream
        for _ in range(2):
            ss.lib().stencil_stream_f32(0, dst.data_ptr(), src.data_ptr(), src.numel(), stream)
        torch.cuda.synchronize()
        result["calibration_copy_bytes"] = 4 * src.numel()
        del src, dst

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(p0_h, vel2_h, h, dt, interior, args.cpu_seconds)

    print(json.dumps(result), flush=True)


def precision_study(ss, p0, vel2, n, steps, tune=True):
    """16-bit field storage against an FP64 run of the same seeded inputs:
    `steps` steps of the acoustic workload (h = 10 m, dt = 1e-3 s); for FP32 and
    each 16-bit format: Gpts/s, relative L2 and max-abs error of the final p
    against FP64 over the interior, and the drift of the leapfrog energy
    (FP64 reductions) between the first and the last step."""
    import numpy as np
    import torch

    dev = p0.device
    shape = (n, n, n)
    interior = (n - 8) ** 3
    grid = ss.GridDims.cube(n, np.float32(10.0), np.float32(1e-3))
    red = ss.Reducer(dev)
    out = {"steps": steps}

    def energy(a, b, v):
        k, p = red.energy(a, b, v, grid)
        return k + p

    def ring(step_fn, init, zero, decode=lambda x: x, v_for_energy=None):
        b = [init(), init(), zero()]
        i = 0
        step_fn(b[(i + 1) % 3], b[i], b[(i + 2) % 3])
        i = 1
        # levels after the first step: p^0 = b[1], p^1 = b[2]
        e_first = energy(decode(b[1]), decode(b[2]), v_for_energy) if v_for_energy is not None else None
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(steps - 1):
            step_fn(b[(i + 1) % 3], b[i], b[(i + 2) % 3])
            i = (i + 1) % 3
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / max(steps - 1, 1)
        last, prev = b[(i + 1) % 3], b[i]
        e_last = energy(decode(prev), decode(last), v_for_energy) if v_for_energy is not None else None
        return last, ms, e_first, e_last

    # FP64 reference (naive kernel, exact association), same inputs widened.
    v64 = vel2.double()
    ref, ms64, e0, e1 = ring(lambda c, p, o: ss.launch_naive(c, p, v64, o, grid),
                             lambda: p0.double(), lambda: torch.zeros(shape, dtype=torch.float64, device=dev),
                             v_for_energy=v64)
    out["fp64_ref"] = {"ms": round(ms64, 4), "energy_drift": (e1 - e0) / e0}

    def report(name, last32, ms, ea, eb, bytes_pt):
        diff = (last32.double() - ref)[4:-4, 4:-4, 4:-4]
        out[name] = {"gpts": round(interior / (ms * 1e-3) / 1e9, 2), "ms": round(ms, 5),
                     "bytes_per_point": bytes_pt,
                     "rel_l2_vs_fp64": red.rel_l2(last32, ref, grid),
                     "max_abs_vs_fp64": float(diff.abs().max()),
                     "energy_drift": (eb - ea) / ea}
        del diff

    if tune:  # FP32 gets the same tuning as the 16-bit formats (cached on repeat sizes)
        tb = [p0.clone(), p0.clone(), torch.zeros_like(p0)]
        ss.tune_tiled(tb[1], tb[0], vel2, tb[2], grid, reps=3)
        del tb
    last, ms, ea, eb = ring(lambda c, p, o: ss.launch_tiled(c, p, vel2, o, grid), lambda: p0.clone(),
                            lambda: torch.zeros_like(p0), v_for_energy=vel2)
    report("fp32", last, ms, ea, eb, 16.0)
    out["fp32"]["tile_shape"] = ss.tuned_shape(grid, "wave")[0]
    del last
    for fmt in ("bf16", "fp16", "bfp16"):
        vs = ss.pow2_scale_for(float(vel2.max())) if fmt == "fp16" else 1.0
        gv = ss.encode16(fmt, vel2, ss.Field16(fmt, shape, dev, vs))
        vdec = ss.decode16(gv)
        if tune:
            tb = [ss.encode16(fmt, p0, ss.Field16(fmt, shape, dev, 1.0)) for _ in range(3)]
            ss.tune_tiled_16(tb[1], tb[0], gv, tb[2], grid, reps=3)
            del tb
        last, ms, ea, eb = ring(lambda c, p, o: ss.launch_tiled_16(c, p, gv, o, grid),
                                lambda: ss.encode16(fmt, p0, ss.Field16(fmt, shape, dev, 1.0)),
                                lambda: ss.encode16(fmt, torch.zeros_like(p0), ss.Field16(fmt, shape, dev, 1.0)),
                                decode=ss.decode16, v_for_energy=vdec)
        report(fmt, ss.decode16(last), ms, ea, eb, 8.25 if fmt == "bfp16" else 8.0)
        out[fmt]["form"] = "fused"
        out[fmt]["tile_shape"] = ss.tuned_shape(grid, fmt)[0]
        out[fmt]["speedup_vs_fp32"] = round(out["fp32"]["ms"] / ms, 3)
        del last
        # the bit-reproducible reference-association form, timed on a short run
        b = [ss.encode16(fmt, p0, ss.Field16(fmt, shape, dev, 1.0)) for _ in range(3)]
        for _ in range(3):
            ss.launch_tiled_16(b[1], b[0], gv, b[2], grid, exact=True)
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        i = 0
        for _ in range(50):
            ss.launch_tiled_16(b[(i + 1) % 3], b[i], gv, b[(i + 2) % 3], grid, exact=True)
            i = (i + 1) % 3
        t1.record()
        torch.cuda.synchronize()
        out[fmt]["gpts_exact_form"] = round(interior / (t0.elapsed_time(t1) / 50 * 1e-3) / 1e9, 2)
        del gv, vdec, b
    del ref, v64
    print(f"[bench] precision {n}^3 x {steps}: " +
          ", ".join(f"{k}: {v.get('gpts', '-')} Gpts/s rel {v.get('rel_l2_vs_fp64', 0):.2e}" for k, v in out.items()
                    if k not in ("fp64_ref", "steps")), file=sys.stderr, flush=True)
    return out


def batch_study(ss, n_shots, steps, dev, tune=True, plan=None):
    """Batched multi-shot modelling: n_shots cubic grids with n drawn uniformly
    from {96, 128, 160, 192, 256, 320} (seed 4); each shot: seeded smooth
    velocity (Gaussian-filtered noise, sigma 8 cells) mapped to [1500, 4500]
    m/s, h = 10 m, dt = 8e-4 s, 15 Hz Ricker at a random interior cell; FP32.
    Launch shapes autotuned per grid size; the run is one hipGraph launch.
    Then the same run as one persistent launch (STENCIL_BATCH_MODE=persistent),
    timed the same way and compared shot by shot for bit-equality.  plan:
    (tile_x, tile_y, chunk_planes, xcd_runs) fixed for both runs instead of
    tuning."""
    import numpy as np
    import torch
    from stencil_solver_amd import synth

    rng = np.random.default_rng(4)
    sizes = rng.choice([96, 128, 160, 192, 256, 320], size=n_shots)
    w = ss.ricker(steps, 8e-4, 15.0)
    shots = []
    t0 = time.time()
    for i, n in enumerate(sizes):
        n = int(n)
        v = synth.smooth_velocity((n, n, n), 8.0, 1500.0, 4500.0, seed=1000 + i, device=dev)
        src = tuple(int(c) for c in rng.integers(4, n - 4, size=3))
        shots.append({"grid": ss.GridDims.cube(n, 10.0, 8e-4), "vel2": (v * v).contiguous(), "wavelet": w,
                      "src": src})
        del v
    t_inputs = time.time() - t0
    print(f"[bench] batch inputs: {n_shots} shots in {t_inputs:.1f}s", file=sys.stderr, flush=True)
    t0 = time.time()
    b = ss.ShotBatch(shots, steps, dev, tune=tune and plan is None)
    if plan is not None:
        b.set_plan(plan[0], plan[1], plan[2], bool(plan[3]))
    torch.cuda.synchronize()
    t_build = time.time() - t0
    print(f"[bench] batch graph built in {t_build:.1f}s", file=sys.stderr, flush=True)
    b.run()  # warm-up replay
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    b.run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    pts = sum((int(n) - 8) ** 3 for n in sizes)
    gpts = pts * steps / (ms * 1e-3) / 1e9
    gbs = 16.0 * pts * steps / (ms * 1e-3) / 1e9
    sizes_hist = {int(k): int(v) for k, v in zip(*np.unique(sizes, return_counts=True))}
    out = {"shots": n_shots, "steps": steps, "sizes": sizes_hist, "interior_points_per_step": pts,
           "run_ms": round(ms, 1), "ms_per_step": round(ms / steps, 4), "gpts": round(gpts, 1),
           "achieved_gbs": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4),
           "graph_build_s": round(t_build, 2), "inputs_s": round(t_inputs, 2),
           "mode": "ragged: one launch per step over all shots" if b.ragged else "grouped: one launch per grid per step",
           "groups": b.groups()}
    print(f"[bench] batch {n_shots} shots x {steps}: {ms:.0f} ms, {gpts:.1f} Gpts/s, {gbs:.0f} GB/s "
          f"(graph build {t_build:.1f}s)", file=sys.stderr, flush=True)
    if b.ragged and os.environ.get("STENCIL_BATCH_MODE") is None:
        # the same run as ONE persistent launch (queues + grid barrier), for comparison
        newest = [b.newest(i).clone() for i in range(n_shots)]
        b.close()
        os.environ["STENCIL_BATCH_MODE"] = "persistent"
        try:
            bp = ss.ShotBatch(shots, steps, dev, tune=tune and plan is None)
            if plan is not None:
                bp.set_plan(plan[0], plan[1], plan[2], bool(plan[3]))
        finally:
            del os.environ["STENCIL_BATCH_MODE"]
        bp.run()
        bp.check()
        e0.record()
        bp.run()
        e1.record()
        torch.cuda.synchronize()
        bp.check()
        pms = e0.elapsed_time(e1)
        same = all(bool((bp.newest(i) == newest[i]).all()) for i in range(n_shots))
        out["persistent"] = {"run_ms": round(pms, 1), "workgroups": bp.persistent_grid,
                             "roofline_frac": round(16.0 * pts * steps / (pms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "vs_per_step_launches": round(ms / pms, 3), "bit_equal": same,
                             "groups": bp.groups()}
        print(f"[bench] batch as one persistent launch: {pms:.0f} ms ({ms / pms:.3f}x), bit-equal {same}",
              file=sys.stderr, flush=True)
        bp.close()
        del bp, newest
    else:
        b.close()
    del b, shots
    torch.cuda.empty_cache()
    return out


def fwi_study(ss, n, steps, dev, tune=True):
    """Adjoint FWI velocity gradient: n^3 FP32, h = 10 m, dt = 8e-4 s.  True
    model: seeded smooth noise (sigma 10 cells) mapped to [1800, 4000] m/s plus
    a +10 % sphere of radius n/12.8 cells at the centre; starting model without
    it.  12 Hz Ricker at (z, y, x) = (12, n/2, n/2); receivers: the whole plane
    z = 8.  grad time = checkpointed forward (misfit) + adjoint sweep."""
    import math

    import torch
    from stencil_solver_amd import synth

    shape = (n, n, n)
    h, dt = 10.0, 8e-4
    grid = ss.GridDims.cube(n, h, dt)
    v0 = synth.smooth_velocity(shape, 10.0, 1800.0, 4000.0, seed=5, device=dev)
    sphere = synth.sphere_mask(shape, (n // 2, n // 2, n // 2), n / 12.8, device=dev)
    v_true = torch.where(sphere, v0 * 1.1, v0)
    m0, m_true = (v0 * v0).contiguous(), (v_true * v_true).contiguous()
    w = ss.ricker(steps, dt, 12.0)
    pr = ss.FwiProblem(grid, steps, (12, n // 2, n // 2), 8, w, dev, torch.float32, tune=tune)
    d_obs = pr.traces(m_true)
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    pr.misfit(m0, d_obs)  # warm-up (first-launch costs)
    pr.gradient(m0)
    torch.cuda.synchronize()
    e[0].record()
    J = pr.misfit(m0, d_obs)
    e[1].record()
    g = pr.gradient(m0)
    e[2].record()
    torch.cuda.synchronize()
    fwd_ms, bwd_ms = e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])
    interior = (n - 8) ** 3
    # per interior point and step: forward 16 B, recompute 16 + 4 (L p);
    # adjoint 16 (psi1, psi2, m in; psi out) plus the gradient pass every
    # other step (L p^k, L p^{k+1}, grad in; grad out: 16 B), 24 B on average
    alg = 60.0 * interior * steps
    grad_ms = fwd_ms + bwd_ms
    out = {"grid": n, "steps": steps, "ckpt_every": pr.ckpt_every, "dtype": "f32",
           "grad_ms": round(grad_ms, 2), "forward_ms": round(fwd_ms, 2), "backward_ms": round(bwd_ms, 2),
           "misfit": float(J), "grad_l2": float(g.double().norm()),
           "workspace_gb": round(pr.workspace_bytes / 1e9, 3),
           "algorithmic_bytes": alg, "achieved_gbs": round(alg / (grad_ms * 1e-3) / 1e9, 1),
           "grad_floor_ms_at_8tbs": round(alg / 8e12 * 1e3, 1), "shape": pr.shape}
    print(f"[bench] fwi {n}^3 x {steps}: grad {grad_ms:.1f} ms (forward {fwd_ms:.1f}, backward {bwd_ms:.1f}), "
          f"{out['achieved_gbs']} GB/s algorithmic", file=sys.stderr, flush=True)
    del pr
    torch.cuda.empty_cache()
    # One segment (ckpt_every = steps): every L p^k stays in HBM, so the
    # backward pass is the adjoint steps alone (44 B/pt/step instead of 60).
    cfg_all = dict(ckpt_every=steps)
    need = 1.05 * 4.0 * n ** 3 * (steps + 10)
    if torch.cuda.mem_get_info()[0] > need:
        pa = ss.FwiProblem(grid, steps, (12, n // 2, n // 2), 8, w, dev, torch.float32, tune=False, **cfg_all)
        pa.misfit(m0, d_obs)
        pa.gradient(m0)
        torch.cuda.synchronize()
        e[0].record()
        pa.misfit(m0, d_obs)
        e[1].record()
        ga = pa.gradient(m0)
        e[2].record()
        torch.cuda.synchronize()
        fa, ba = e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])
        out["store_all"] = {"grad_ms": round(fa + ba, 2), "forward_ms": round(fa, 2), "backward_ms": round(ba, 2),
                            "workspace_gb": round(pa.workspace_bytes / 1e9, 2),
                            "grad_bit_identical": bool(torch.equal(ga, g))}
        print(f"[bench] fwi store-all L p ({out['store_all']['workspace_gb']} GB): grad {fa + ba:.1f} ms "
              f"(forward {fa:.1f}, backward {ba:.1f}), identical {out['store_all']['grad_bit_identical']}",
              file=sys.stderr, flush=True)
        del pa, ga
    del d_obs, g, m0, m_true, v0, v_true
    torch.cuda.empty_cache()
    return out


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p0_h, vel2_h, h, dt, interior, seconds):
    """The reference CPU path (src/kernels/stencil_naive.cu:136-170, OpenMP)
    built from its sources into oracle/_ref/libref_v3.so, timed like
    SolverBase::run in CPU mode on a bounded number of steps of the same
    workload, in two rows: as the reference builds it (denormals kept) and
    with flush-to-zero / denormals-are-zero set in every OpenMP thread.  The
    oracle's C restatement (one row) if that build is absent."""
    from oracle.lib import Oracle, RefLib, ref_available

    omp_env = {k: os.environ[k] for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES", "OMP_SCHEDULE")
               if k in os.environ}
    rows = []
    if ref_available("v3"):
        ref = RefLib("v3")
        kind = "reference"
        cores = ref.omp_threads()
        for mode, ftz in (("reference build flags (-O3 -march=x86-64-v3, FMA contraction), denormals kept", False),
                          ("same build, FTZ/DAZ set in every OpenMP thread", True)):
            if ref.set_ftz_daz(ftz) != cores:
                raise RuntimeError("could not set the FTZ/DAZ state in every OpenMP thread")
            one = ref.time_wave_steps(p0_h, vel2_h, h, dt, 0, 1)
            n_steps = int(max(2, min(60, seconds * 1e3 / max(one, 1e-3))))
            ms = ref.time_wave_steps(p0_h, vel2_h, h, dt, 1, n_steps)
            rows.append({"mode": mode, "ftz_daz": ftz, "ms_per_step": round(ms, 2), "timed_steps": n_steps,
                         "gpts": round(interior / (ms * 1e-3) / 1e9, 4)})
        ref.set_ftz_daz(False)
    else:
        import numpy as np
        o = Oracle()
        kind = "port"
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        bufs = [p0_h.copy(), p0_h.copy(), np.zeros_like(p0_h)]
        t = time.perf_counter()
        o.wave_step(bufs[1], bufs[0], vel2_h, bufs[2], h, dt)
        one = (time.perf_counter() - t) * 1e3
        n_steps = int(max(2, min(60, seconds * 1e3 / max(one, 1e-3))))
        t = time.perf_counter()
        i = 1
        for _ in range(n_steps):
            o.wave_step(bufs[(i + 1) % 3], bufs[i], vel2_h, bufs[(i + 2) % 3], h, dt)
            i = (i + 1) % 3
        ms = (time.perf_counter() - t) * 1e3 / n_steps
        rows.append({"mode": "oracle C restatement (-O2, no contraction), denormals kept", "ftz_daz": False,
                     "ms_per_step": round(ms, 2), "timed_steps": n_steps,
                     "gpts": round(interior / (ms * 1e-3) / 1e9, 4)})
    main = rows[0]
    for r in rows:
        print(f"[bench] cpu baseline ({kind}, {cores} threads, {r['mode']}): {r['ms_per_step']:.1f} ms/step, "
              f"{r['gpts']:.3f} Gpts/s", file=sys.stderr, flush=True)
    return {"value": main["gpts"], "unit": "Gpts/s", "cores": cores, "kind": kind,
            "ms_per_step": main["ms_per_step"],
            "sample": f"{main['timed_steps']} timed steps (after 1 warmup) of the same 512^3 workload through the "
                      f"reference launch_naive CPU path (OpenMP, {cores} threads, denormals kept as the reference "
                      f"builds it); FTZ/DAZ row beside it",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(), "omp_env": omp_env,
            "omp_schedule": "static, collapse(3) (stencil_naive.cu:151)", "rows": rows}


if __name__ == "__main__":
    main()
