"""Write the golden fixtures in tests/golden/ from the reference's own CPU path.

Run in a container that has the reference tree (the recipe in oracle/Makefile
compiles its CPU path from the sources where they lie into oracle/_ref/):

    make -C oracle all ref && python tests/golden/make_golden.py

Every array here is an OUTPUT of the reference code (launch_naive,
make_solver/step/download, init_gaussian, Config::make_grid) computed through
oracle/_ref/libref_*.so; the inputs are the reference's own initial conditions
and are regenerated by the tests from the recorded parameters.  Two builds:
``nofma`` (-ffp-contract=off, the bit-exact target) and ``native`` (the
reference's Release flags -O3 -march=native, FMA-contracted, host-ISA
dependent).  ``f64`` is the STENCIL_FP64 build.
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle.lib import RefLib  # noqa: E402

OUT = Path(__file__).resolve().parent


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def summary(a: np.ndarray, rng_seed: int = 1234, n_samples: int = 4096) -> dict:
    rng = np.random.default_rng(rng_seed)
    idx = np.sort(rng.choice(a.size, size=n_samples, replace=False)).astype(np.int64)
    a64 = a.astype(np.float64)
    return {
        "sample_idx": idx,
        "sample_val": a.reshape(-1)[idx].copy(),
        "mid_z": a[a.shape[0] // 2].copy(),
        "sum": np.float64(a64.sum()),
        "sum_sq": np.float64((a64 * a64).sum()),
        "max": np.float64(a64.max()),
        "min": np.float64(a64.min()),
    }


def ring_run(ref: RefLib, p0, vel2, dx, dt, n_steps):
    """n_steps of the reference launch_naive over a 3-buffer ring, as
    WavePropSolver::step_impl does; returns the newest level (cur)."""
    import ctypes as C
    bufs = [p0.copy(), p0.copy(), np.zeros_like(p0)]
    nz, ny, nx = p0.shape
    ref.lib.ref_time_wave_steps(*[C.c_void_p(b.ctypes.data) for b in bufs], C.c_void_p(vel2.ctypes.data),
                                nx, ny, nz, dx, dt, 0, n_steps)
    return bufs[(n_steps + 1) % 3]


# stencil_slice runs recorded from the reference program (flags as passed).
SLICE_RUNS = {
    "slice_wave_32": ["--nx", "32", "--steps", "20", "--every", "5", "--kernel", "naive"],
    "slice_heat_24": ["--nx", "24", "--steps", "12", "--every", "4", "--physics", "heat", "--kernel", "tiled",
                      "--diffusivity", "0.02"],
}


def main():
    nofma, native, f64 = RefLib("nofma"), RefLib("native"), RefLib("f64")
    meta: dict = {"builds": {"nofma": "-O3 -march=x86-64-v3 -ffp-contract=off",
                             "native": "-O3 -march=native (reference Release flags)",
                             "f64": "-DSTENCIL_FP64 -O3 -march=x86-64-v3 -ffp-contract=off"}}

    # 1. One launch_naive call, 24^3 (tests/unit/test_kernels.cpp:43-50 dims:
    #    h = 10, dt = 1e-3), Gaussian sigma 3 into cur and prev, v = 1500.
    n = 24
    g = nofma.gaussian(n, 0.5, 0.5, 0.5, 3.0)
    vel2 = np.full_like(g, np.float32(1500.0) * np.float32(1500.0))
    out = {}
    for name, ref in (("nofma", nofma), ("native", native)):
        nxt = np.zeros_like(g)
        ref.wave_step(g, g, vel2, nxt, np.float32(10.0), np.float32(1e-3))
        out[name] = nxt
    np.savez_compressed(OUT / "kernel_24.npz", init=g, next_nofma=out["nofma"], next_native=out["native"])
    meta["kernel_24"] = {"n": n, "dx": 10.0, "dt": 1e-3, "sigma": 3.0, "velocity": 1500.0,
                         "sha_init": sha(g), "sha_next_nofma": sha(out["nofma"])}

    # 2. make_solver(wave) 32^3, v = 1500, sigma 4 (test_kernel_correctness.cpp:25-38),
    #    1 and 20 steps, naive and tiled variants (CPU tiled == naive).
    d = {}
    for steps in (1, 20):
        for variant, vname in ((0, "naive"), (1, "tiled")):
            a, _ = nofma.solver(0, variant, 32, 10.0, 1500.0, 1e-4, 0.5, 0.5, 0.5, 4.0, steps)
            d[f"{vname}{steps}_nofma"] = a
        a, _ = native.solver(0, 0, 32, 10.0, 1500.0, 1e-4, 0.5, 0.5, 0.5, 4.0, steps)
        d[f"naive{steps}_native"] = a
    # run(n, warmup): warmup steps also advance the state (solver_base.cpp:71-73).
    a, _ = nofma.solver(0, 0, 32, 10.0, 1500.0, 1e-4, 0.5, 0.5, 0.5, 4.0, 5, 2, True)
    d["run5_warm2_nofma"] = a
    np.savez_compressed(OUT / "solver_wave_32.npz", **d)
    meta["solver_wave_32"] = {"n": 32, "dx": 10.0, "velocity": 1500.0, "sigma": 4.0,
                              "dt_make_grid": nofma.make_grid_dt(10.0, 1500.0)}

    # 3. make_solver(heat) 48^3, dx = 1, v = 500 (dt only), alpha = 1e-3,
    #    sigma 5 (tests/integration/test_heat_3d.cpp:29-43), 20 steps.
    h, _ = nofma.solver(1, 0, 48, 1.0, 500.0, 1e-3, 0.5, 0.5, 0.5, 5.0, 20)
    hn, _ = native.solver(1, 0, 48, 1.0, 500.0, 1e-3, 0.5, 0.5, 0.5, 5.0, 20)
    np.savez_compressed(OUT / "solver_heat_48.npz", heat20_nofma=h, heat20_native=hn)
    meta["solver_heat_48"] = {"n": 48, "dx": 1.0, "velocity": 500.0, "diffusivity": 1e-3,
                              "sigma": 5.0, "steps": 20, "sha_nofma": sha(h)}

    # 4. Workload "acoustic 128^3": h = 10, v = 2000, dt = 1e-3, Gaussian
    #    sigma 3 at the centre, 200 steps, through the reference launch_naive.
    n = 128
    res = {}
    for name, ref in (("nofma", nofma), ("native", native), ("f64", f64)):
        T = ref.dtype
        p0 = ref.gaussian(n, 0.5, 0.5, 0.5, 3.0)
        v2 = np.full_like(p0, T(2000.0) * T(2000.0))
        fin = ring_run(ref, p0, v2, T(10.0), T(1e-3), 200)
        s = summary(fin)
        res[name] = s
        meta[f"acoustic128_{name}"] = {"sha_final": sha(fin), "sha_init": sha(p0),
                                       "sum": float(s["sum"]), "sum_sq": float(s["sum_sq"]),
                                       "max": float(s["max"]), "min": float(s["min"])}
        print(name, "||p||2 =", float(np.sqrt(s["sum_sq"])))
    np.savez_compressed(OUT / "acoustic_128.npz",
                        **{f"{k}_{name}": v for name, s in res.items() for k, v in s.items()})
    meta["acoustic_128"] = {"n": 128, "dx": 10.0, "velocity": 2000.0, "dt": 1e-3, "sigma": 3.0,
                            "steps": 200}

    # 5. Scalars: Config::make_grid dt and init_gaussian at a few settings.
    meta["make_grid_dt"] = [{"dx": dx, "velocity": v, "dt": nofma.make_grid_dt(dx, v)}
                            for dx, v in ((10.0, 1500.0), (1.0, 500.0), (5.0, 2000.0), (10.0, 4000.0))]
    meta["gaussian"] = []
    for n, sg, c in ((32, 3.0, 0.5), (33, 5.0, 0.25), (17, 2.5, 0.7)):
        gg = nofma.gaussian(n, c, c, c, sg)
        meta["gaussian"].append({"n": n, "sigma": sg, "c": c, "sha": sha(gg),
                                 "centre_val": float(gg[n // 2, n // 2, n // 2])})

    # 6. init_layered_velocity (grid.hpp:91-99): even and odd nz (the split is
    #    at nz / 2 in integer division), every cell including the halo.
    meta["layered_velocity"] = []
    for shape, vt, vb in (((16, 8, 8), 1500.0, 3000.0), ((17, 6, 5), 2000.0, 4500.0), ((9, 4, 3), 1234.5, 987.25)):
        lv = nofma.layered_velocity(shape, vt, vb)
        meta["layered_velocity"].append({"shape": list(shape), "v_top": vt, "v_bot": vb, "sha": sha(lv),
                                         "first": float(lv.reshape(-1)[0]), "last": float(lv.reshape(-1)[-1])})

    # 7. l2_error / max_abs_error (grid.cpp:86-115) on seeded inputs that the
    #    tests regenerate (numpy default_rng(seed), U[-1, 1), float32).
    meta["norms"] = []
    for seed, shape in ((31, (9, 10, 11)), (32, (12, 12, 12))):
        rng = np.random.default_rng(seed)
        a = rng.uniform(-1.0, 1.0, shape).astype(np.float32)
        b = rng.uniform(-1.0, 1.0, shape).astype(np.float32)
        meta["norms"].append({"seed": seed, "shape": list(shape), "l2_error": nofma.l2_error(a, b),
                              "max_abs_error": nofma.max_abs_error(a, b)})

    # 8. The reference's stencil_slice program (src/main_slice.cpp, built by
    #    oracle/Makefile as oracle/_ref/ref_slice): its .f32 frames and .json
    #    metadata, kept as written.
    import subprocess
    import tempfile
    prog = ROOT / "oracle" / "_ref" / "ref_slice"
    meta["slices"] = {}
    for name, args in SLICE_RUNS.items():
        with tempfile.TemporaryDirectory() as tmp:
            prefix = Path(tmp) / name
            subprocess.run([str(prog), *args, "--slice-out", str(prefix)], check=True, capture_output=True)
            (OUT / f"{name}.f32").write_bytes(prefix.with_suffix(".f32").read_bytes())
            (OUT / f"{name}.json").write_text(prefix.with_suffix(".json").read_text())
        meta["slices"][name] = {"args": args}

    (OUT / "golden_meta.json").write_text(json.dumps(meta, indent=1, sort_keys=True))
    print("wrote", sorted(p.name for p in OUT.glob("*.npz")))


if __name__ == "__main__":
    main()
