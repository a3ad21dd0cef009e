"""Runs the C++ suites in tests/cpp (built by __graft_entry__.build()):
test_cpu_units on the host, test_gpu_solver on the device."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "_bin"


def _run(name: str, timeout: int):
    exe = BIN / name
    assert exe.exists(), f"{exe} not built (run __graft_entry__.build())"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=timeout)
    tail = "\n".join(r.stdout.splitlines()[-5:])
    assert r.returncode == 0, f"{name} failed:\n{r.stderr[-4000:]}\n{tail}"
    return r.stdout


def test_cpp_host_units():
    out = _run("test_cpu_units", 120)
    assert "0 failed" in out


@pytest.mark.gpu
def test_cpp_device_suites(dev):
    out = _run("test_gpu_solver", 300)
    assert "0 failed" in out
