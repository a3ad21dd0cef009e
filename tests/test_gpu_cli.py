"""The command-line programs on the GPU: results JSON in the reference's
schema (tiled + the _naive sibling); the slice writer's files equal the files
the reference's own stencil_slice writes for the same flags (fixtures in
tests/golden/, made by tests/golden/make_golden.py from oracle/_ref/ref_slice):
the frame file byte for byte, the metadata key for key; and its frames equal
the solver API's mid-plane after the same steps, bit for bit."""
from __future__ import annotations

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BIN = Path(__file__).resolve().parent.parent / "stencil_solver_amd" / "bin"
KEYS = {"gpu_name", "cuda_version", "stencil_radius", "precision", "tile_x", "tile_y", "pencil_z", "steps",
        "warmup", "results"}
RKEYS = {"label", "grid_size", "time_ms", "bandwidth_gbs", "model_bandwidth_gbs", "gflops"}


def _run(args, timeout=120):
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    return r.stdout


def test_tiled_program_writes_both_json_files(dev, tmp_path):
    out = tmp_path / "res" / "tiled.json"
    out.parent.mkdir()
    text = _run([BIN / "stencil_tiled", "--nx", 64, "--ny", 64, "--nz", 64, "--steps", 10, "--warmup", 2,
                 "--output", out])
    assert "as fast as" in text  # the naive -> tiled comparison line
    for path, label in ((out, "tiled"), (out.parent / "tiled_naive.json", "naive")):
        d = json.loads(path.read_text())
        assert KEYS <= set(d) and d["stencil_radius"] == 4 and d["precision"] == "fp32"
        assert d["steps"] == 10 and d["warmup"] == 2 and d["gpu_name"]
        (r,) = d["results"]
        assert RKEYS <= set(r) and r["label"] == label and r["grid_size"] == 64 and r["time_ms"] > 0
    assert json.loads(out.read_text())["tile_x"] > 0


def test_naive_program(dev, tmp_path):
    out = tmp_path / "naive.json"
    _run([BIN / "stencil_naive", "--nx", 48, "--ny", 48, "--nz", 48, "--steps", 5, "--warmup", 1, "--output", out])
    d = json.loads(out.read_text())
    assert d["results"][0]["grid_size"] == 48 and d["results"][0]["bandwidth_gbs"] > 0


GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", ["slice_wave_32", "slice_heat_24"])
@pytest.mark.parametrize("kernel", ["naive", "tiled"])
def test_slice_files_equal_reference_program(dev, tmp_path, name, kernel):
    """Same flags as the reference run (only --kernel varies: every variant is
    bit-identical to the reference CPU path), same bytes in the .f32 file and
    the same metadata (numbers to 1e-6: the reference prints 6 digits)."""
    meta = json.loads((GOLDEN / "golden_meta.json").read_text())["slices"][name]
    args = list(meta["args"])
    args[args.index("--kernel") + 1] = kernel
    prefix = tmp_path / name
    _run([BIN / "stencil_slice", *args, "--slice-out", prefix])
    assert prefix.with_suffix(".f32").read_bytes() == (GOLDEN / f"{name}.f32").read_bytes()
    want = json.loads((GOLDEN / f"{name}.json").read_text())
    got = json.loads(prefix.with_suffix(".json").read_text())
    assert set(got) == set(want)
    for k, v in want.items():
        if k == "kernel":
            assert got[k] == kernel
        elif isinstance(v, float) or isinstance(got[k], float):
            assert got[k] == pytest.approx(v, rel=1e-6), k
        else:
            assert got[k] == v, k


@pytest.mark.parametrize("physics,kernel", [("wave", "tiled"), ("heat", "naive")])
def test_slice_frames_match_solver(dev, tmp_path, physics, kernel):
    import stencil_solver_amd as ss

    n, steps, every = 48, 12, 4
    prefix = tmp_path / "slice"
    _run([BIN / "stencil_slice", "--nx", n, "--steps", steps, "--every", every, "--slice-out", prefix,
          "--physics", physics, "--kernel", kernel])
    meta = json.loads(prefix.with_suffix(".json").read_text())
    assert meta["frames"] == 1 + steps // every and meta["plane_index"] == n // 2
    assert meta["physics"] == physics and meta["kernel"] == kernel and meta["dtype"] == "float32"
    frames = np.fromfile(prefix.with_suffix(".f32"), dtype=np.float32).reshape(meta["frames"], n, n)
    s = ss.make_solver(ss.Config(nx=n, ny=n, nz=n, physics=physics, variant=kernel))
    assert np.array_equal(frames[0], s.download()[n // 2])
    for f in range(1, meta["frames"]):
        for _ in range(every):
            s.step()
        assert np.array_equal(frames[f], s.download()[n // 2]), f
    assert np.abs(frames[-1]).max() > 0
