"""The allocation-free energy reduction (stencil_energy_f32_noalloc) against
stencil_energy_f32 with a caller workspace, through the C-ABI: the same bits
on two calls in a row, the second on a smaller grid, on the default stream
and on a second stream; invalid grids and null pointers are rejected."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(ss, dev, n, seed):
    """(p^n, p^{n+1}, vel2, grid) after two wave steps of a pulse on a seeded
    layered model, n^3 cells (nx a little wider: ragged x)."""
    import torch

    shape = (n, n, n + 4)
    nz, ny, nx = shape
    grid = ss.GridDims(nx, ny, nz, np.float32(10.0), np.float32(10.0), np.float32(10.0), np.float32(1e-3))
    p0 = torch.from_numpy(ss.host_gaussian(shape, 0.5, 0.5, 0.5, 3.0)).to(dev)
    vel2 = torch.from_numpy(ss.host_layered_noise_vel2(shape, 16, 1500.0, 4000.0, 0.05, seed)).to(dev)
    b = [p0.clone(), p0.clone(), torch.zeros_like(p0)]
    ss.launch_naive(b[1], b[0], vel2, b[2], grid)
    ss.launch_naive(b[2], b[1], vel2, b[0], grid)
    return b[2], b[0], vel2, grid


def _energy_ws(ss, L, p_n, p_np1, vel2, grid, stream):
    import torch

    work = torch.empty(int(L.stencil_reduce_workspace_bytes()) // 8, dtype=torch.float64, device=p_n.device)
    out = torch.full((3,), float("nan"), dtype=torch.float64, device=p_n.device)
    g = grid.to_c()
    rc = L.stencil_energy_f32(p_n.data_ptr(), p_np1.data_ptr(), vel2.data_ptr(), C.byref(g), out.data_ptr(),
                              work.data_ptr(), stream)
    return rc, out


def _energy_noalloc(L, p_n, p_np1, vel2, grid, stream):
    import torch

    out = torch.full((3,), float("nan"), dtype=torch.float64, device=p_n.device)
    g = grid.to_c()
    rc = L.stencil_energy_f32_noalloc(p_n.data_ptr(), p_np1.data_ptr(), vel2.data_ptr(), C.byref(g),
                                      out.data_ptr(), stream)
    return rc, out


def test_noalloc_energy_equals_workspace_energy(dev):
    import torch

    import stencil_solver_amd as ss

    L = ss.lib()
    side = torch.cuda.Stream()
    for n, seed, stream in ((160, 3, None), (40, 4, None), (72, 5, side), (160, 3, side)):
        p_n, p_np1, vel2, grid = _pair(ss, dev, n, seed)
        torch.cuda.synchronize()
        st = (stream or torch.cuda.current_stream()).cuda_stream
        print(f"[no-alloc] grid {grid.nx}x{grid.ny}x{grid.nz}, stream {'side' if stream else 'current'}",
              flush=True)
        rc_a, want = _energy_ws(ss, L, p_n, p_np1, vel2, grid, st)
        rc_b, got = _energy_noalloc(L, p_n, p_np1, vel2, grid, st)
        torch.cuda.synchronize()
        assert rc_a == 0 and rc_b == 0
        assert torch.equal(got, want), (got, want)
        k, pot = got[0].item(), got[1].item()
        assert k > 0 and pot > 0 and np.isfinite(k + pot)


def test_noalloc_energy_rejects_bad_arguments(dev):
    import torch

    import stencil_solver_amd as ss

    L = ss.lib()
    p_n, p_np1, vel2, grid = _pair(ss, dev, 24, 1)
    st = torch.cuda.current_stream().cuda_stream
    out = torch.zeros(3, dtype=torch.float64, device=dev)
    small = ss.GridDims(8, 24, 24, np.float32(10.0), np.float32(10.0), np.float32(10.0), np.float32(1e-3)).to_c()
    huge = ss.GridDims(1 << 20, 24, 24, np.float32(10.0), np.float32(10.0), np.float32(10.0),
                       np.float32(1e-3)).to_c()
    g = grid.to_c()
    assert L.stencil_energy_f32_noalloc(p_n.data_ptr(), p_np1.data_ptr(), vel2.data_ptr(), C.byref(small),
                                        out.data_ptr(), st) == 1
    assert L.stencil_energy_f32_noalloc(p_n.data_ptr(), p_np1.data_ptr(), vel2.data_ptr(), C.byref(huge),
                                        out.data_ptr(), st) == 1
    assert L.stencil_energy_f32_noalloc(None, p_np1.data_ptr(), vel2.data_ptr(), C.byref(g), out.data_ptr(),
                                        st) == 1
    assert L.stencil_energy_f32_noalloc(p_n.data_ptr(), p_np1.data_ptr(), vel2.data_ptr(), C.byref(g), None,
                                        st) == 1
    torch.cuda.synchronize()
    assert not out.any()
