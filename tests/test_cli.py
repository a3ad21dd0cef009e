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
