"""The host code under AddressSanitizer + UBSan (tests/cpp/Makefile, target
`sanitize`, built by __graft_entry__.build()): the library's host sources and
the host unit suites, and every oracle function on ragged grids.  Any
out-of-bounds access, use after free or undefined behaviour ends the run with
a report and a non-zero status."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

import pytest

SAN = Path(__file__).resolve().parent / "cpp" / "_bin" / "san"
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.mark.parametrize("prog,marker", [("test_cpu_units", "0 failed"), ("oracle_check", "0 failed")])
def test_sanitized_host_code(prog, marker):
    exe = SAN / prog
    assert exe.exists(), f"{exe} not built (run __graft_entry__.build())"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=ENV)
    report = (r.stderr or "")[-4000:]
    assert r.returncode == 0, f"{prog} failed under the sanitizers:\n{report}"
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, report
    assert marker in r.stdout
