"""The HBM calibration kernels behind the roofline figures
(stencil_stream_f32; the reference's tests/perf/bench_stream.cu:53-101):
each operation computes what it claims on a length that is not a multiple of
the workgroup, and bad arguments are rejected."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_stream_operations(dev):
    import torch

    import stencil_solver_amd as ss

    L = ss.lib()
    n = 1 << 20
    n += 4 * 37  # not a multiple of 256 threads x 4 floats
    rng = np.random.default_rng(3)
    src_h = rng.random(3 * n, dtype=np.float32)
    src = torch.from_numpy(src_h).to(dev)
    dst = torch.zeros(n, dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    # copy
    assert L.stencil_stream_f32(0, dst.data_ptr(), src.data_ptr(), n, st) == 0
    assert np.array_equal(dst.cpu().numpy(), src_h[:n])
    # write: dst := 1
    assert L.stencil_stream_f32(2, dst.data_ptr(), None, n, st) == 0
    assert bool((dst == 1.0).all())
    # 3 reads, 1 write: dst = (a + b) + c
    assert L.stencil_stream_f32(3, dst.data_ptr(), src.data_ptr(), n, st) == 0
    want = (src_h[:n] + src_h[n:2 * n]) + src_h[2 * n:]
    assert np.array_equal(dst.cpu().numpy(), want)
    # read: one partial sum per workgroup (at most 2048) whose total is sum(src[:n])
    sink = torch.zeros(2048, dtype=torch.float32, device=dev)
    assert L.stencil_stream_f32(1, sink.data_ptr(), src.data_ptr(), n, st) == 0
    groups = min(2048, -(-(n // 4) // 256))
    parts = sink.cpu().numpy()
    assert np.all(parts[groups:] == 0) and np.all(parts[:groups] > 0)
    assert parts.astype(np.float64).sum() == pytest.approx(src_h[:n].astype(np.float64).sum(), rel=1e-5)
    # bad arguments
    assert L.stencil_stream_f32(0, dst.data_ptr(), src.data_ptr(), n - 2, st) == 1  # n % 4
    assert L.stencil_stream_f32(0, dst.data_ptr(), dst.data_ptr(), n, st) == 1  # copy onto itself
    assert L.stencil_stream_f32(0, dst.data_ptr() + 4, src.data_ptr(), n - 4, st) == 1  # unaligned
    assert L.stencil_stream_f32(7, dst.data_ptr(), src.data_ptr(), n, st) == 1
