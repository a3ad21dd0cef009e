"""The command-line programs (stencil_naive, stencil_tiled, stencil_slice)
exist after the build and reject bad arguments with the usage text before
touching a GPU, like the reference's mains."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent.parent / "stencil_solver_amd" / "bin"


@pytest.mark.parametrize("prog", ["stencil_naive", "stencil_tiled", "stencil_slice"])
def test_programs_reject_unknown_flags(prog):
    exe = BIN / prog
    assert exe.exists(), f"{exe} not built"
    r = subprocess.run([str(exe), "--no-such-flag"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "Usage" in r.stderr


def test_slice_rejects_bad_every():
    r = subprocess.run([str(BIN / "stencil_slice"), "--every", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "--every" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits non-zero with a message
    before touching a device."""
    import sys

    root = Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300, cwd=root)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
