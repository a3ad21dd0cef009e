"""C-ABI library checks that need no device: it loads, exports every symbol the
header declares, rejects invalid arguments before touching a device, and its
host-side input generators equal the oracle's bit for bit."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "stencil_capi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(stencil_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import stencil_solver_amd as ss
    from stencil_solver_amd._lib import SIGNATURES

    L = ss.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    # the Python binding types every declared entry point
    assert set(names) == set(SIGNATURES)


def test_library_is_built_for_gfx950():
    import stencil_solver_amd as ss

    blob = Path(ss.LIB_PATH).read_bytes()
    assert b"gfx950" in blob


def test_invalid_arguments_rejected_without_device():
    import stencil_solver_amd as ss
    from stencil_solver_amd._lib import stencil_grid

    L = ss.lib()
    g = stencil_grid(32, 32, 32, 10.0, 10.0, 10.0, 1e-3)
    a = C.c_void_p(0x1000)
    b = C.c_void_p(0x2000)
    null = C.c_void_p(0)
    assert L.stencil_launch_naive_f32(null, a, a, b, C.byref(g), null) == 1
    assert L.stencil_launch_tiled_f32(a, a, a, a, C.byref(g), null) == 1  # p_next aliases p_cur
    assert L.stencil_launch_f32(7, a, a, a, b, C.byref(g), null) == 1
    assert L.stencil_launch_naive_f32(a, a, a, b, None, null) == 1
    g0 = stencil_grid(0, 32, 32, 10.0, 10.0, 10.0, 1e-3)
    assert L.stencil_launch_naive_f64(a, a, a, b, C.byref(g0), null) == 1
    assert L.stencil_heat_step_f32(a, a, 1.0, C.byref(g), 0, null) == 1
    bad_shape = ss.tile_shape(16, 7)
    assert L.stencil_launch_tiled_ex_f32(a, a, a, b, C.byref(g), 0, C.byref(bad_shape), null) == 1
    ok_shape = ss.tile_shape(16, 32)
    # unaligned base -> the tiled kernel declines (the public launcher then runs naive)
    u = C.c_void_p(0x1004)
    assert L.stencil_launch_tiled_ex_f32(u, a, a, b, C.byref(g), 0, C.byref(ok_shape), null) == 3
    g_odd = stencil_grid(30, 32, 32, 10.0, 10.0, 10.0, 1e-3)
    assert L.stencil_launch_tiled_ex_f32(a, a, a, b, C.byref(g_odd), 0, None, null) == 3
    assert L.stencil_strerror(1) == b"invalid argument"
    from stencil_solver_amd._lib import stencil_config
    cfg = stencil_config()
    L.stencil_config_default(C.byref(cfg))
    cfg.nx = 8  # no interior
    h = C.c_void_p()
    assert L.stencil_solver_create(C.byref(cfg), C.byref(h)) == 1


def test_config_defaults_match_reference():
    """Reference include/stencil/config.hpp:34-70."""
    import stencil_solver_amd as ss

    c = ss.Config().to_c()
    assert (c.nx, c.ny, c.nz) == (256, 256, 256)
    assert (c.dx, c.steps, c.warmup_steps) == (10.0, 100, 10)
    assert c.velocity == 1500.0 and c.diffusivity == float(np.float32(1e-4))
    assert c.pulse_sigma == 5.0 and c.physics == 0 and c.variant == 0


def test_byte_models():
    import stencil_solver_amd as ss

    g = ss.GridDims.cube(512)
    assert ss.wave_bytes_per_step(g) == 16.0 * 504**3
    assert ss.heat_bytes_per_step(g) == 8.0 * 504**3
    assert ss.wave_bytes_per_step(g, 2) == 8.0 * 504**3
    assert ss.wave_bytes_per_step(ss.GridDims.cube(8)) == 0.0


@pytest.mark.parametrize("shape", [(24, 24, 24), (17, 33, 20), (40, 12, 36)])
def test_host_gaussian_equals_oracle(oracle, shape):
    import stencil_solver_amd as ss

    g = ss.host_gaussian(shape, 0.5, 0.4, 0.6, 3.0)
    o = oracle.gaussian(0, 0.5, 0.4, 0.6, 3.0, shape=shape)
    assert np.array_equal(g, o)


def test_host_generators_equal_oracle(oracle):
    import stencil_solver_amd as ss

    shape = (48, 40, 36)
    assert np.array_equal(ss.host_layered_noise_vel2(shape, 16, 1500.0, 4000.0, 0.05, 2),
                          oracle.layered_noise_vel2(shape, 16, 1500.0, 4000.0, 0.05, 2))
    assert np.array_equal(ss.host_noise_cube(shape, 8, 10.0, 5), oracle.heat_noise_cube(shape, 8, 10.0, 5))


@pytest.mark.parametrize("n,ckpt,esz", [(24, 0, 4), (24, 24, 4), (2000, 0, 4), (37, 1, 8), (10, 7, 8)])
def test_fwi_workspace_layout(n, ckpt, esz):
    """Workspace bytes: 3 wavefield and 3 adjoint levels, 2 * S checkpoint
    fields (S = ceil(n / C) segments), two banks of C Laplacian fields when
    S > 1 (the next segment's recompute overlaps this one's adjoint steps),
    one bank when S = 1, then n receiver planes and the reduction scratch,
    each region 256-byte aligned."""
    import math

    import stencil_solver_amd as ss
    from stencil_solver_amd._lib import stencil_fwi_config, stencil_grid

    nx, ny, nz = 44, 36, 30
    g = stencil_grid(nx, ny, nz, 10.0, 10.0, 10.0, 1e-3)
    cfg = stencil_fwi_config(n, 20, 18, 15, 6, ckpt, 1)
    up = lambda b: (b + 255) // 256 * 256
    field = up(nx * ny * nz * esz)
    C_ = min(ckpt or math.ceil(math.sqrt(n)), n)
    S = (n + C_ - 1) // C_
    banks = 2 if S > 1 else 1
    want = 6 * field + 2 * S * field + banks * C_ * field + up(n * nx * ny * esz) + up(1024 * 8) + up(8)
    assert ss.lib().stencil_fwi_workspace_bytes(C.byref(g), C.byref(cfg), esz) == want
    # invalid: receiver plane in the halo, value size 2
    bad = stencil_fwi_config(n, 20, 18, 15, 2, ckpt, 1)
    assert ss.lib().stencil_fwi_workspace_bytes(C.byref(g), C.byref(bad), esz) == 0
    assert ss.lib().stencil_fwi_workspace_bytes(C.byref(g), C.byref(cfg), 2) == 0


def test_fwi_tuner_rejects_invalid_arguments_without_device():
    import stencil_solver_amd as ss
    from stencil_solver_amd._lib import stencil_fwi_config, stencil_grid

    L = ss.lib()
    g = stencil_grid(44, 36, 30, 10.0, 10.0, 10.0, 1e-3)
    cfg = stencil_fwi_config(24, 20, 18, 15, 6, 0, 1)
    work = C.c_void_p(0x10000)
    assert L.stencil_tune_fwi_f32(C.byref(g), C.byref(cfg), None, 3, None, None, None, None, None) == 1
    assert L.stencil_tune_fwi_f32(C.byref(g), C.byref(cfg), work, 0, None, None, None, None, None) == 1
    assert L.stencil_tune_fwi_f32(C.byref(g), C.byref(cfg), C.c_void_p(0x10010), 3, None, None, None, None,
                                  None) == 1  # work not 256-byte aligned
    bad = stencil_fwi_config(24, 2, 18, 15, 6, 0, 1)  # source in the halo
    assert L.stencil_tune_fwi_f32(C.byref(g), C.byref(bad), work, 3, None, None, None, None, None) == 1


def test_set_tuned_shape_round_trip_and_validation():
    import stencil_solver_amd as ss

    g = ss.GridDims(72, 44, 30)
    ss.clear_tuning()
    shp = {"tile_x": 128, "tile_y": 16, "n_chunks": 3, "xcd_remap": 0, "nontemporal": 1}
    for kind in ("wave", "heat", "bfp16", "fwi_forward", "fwi_adjoint", "fwi_recompute"):
        ss.set_tuned_shape(g, kind, shp)
        got, hit = ss.tuned_shape(g, kind)
        assert hit and got == shp, kind
    from stencil_solver_amd._lib import StencilError

    with pytest.raises(StencilError):
        ss.set_tuned_shape(g, "wave", dict(shp, tile_y=7))  # not a built tile
    with pytest.raises(StencilError):
        ss.set_tuned_shape(g, "wave", dict(shp, xcd_remap=2))
    # the step pair is built for four tiles only (128 x 8 / 16, 256 x 4 / 8 cells)
    pair = dict(shp, tile_x=256, tile_y=4)
    ss.set_tuned_shape(g, "fwi_pair", pair)
    assert ss.tuned_shape(g, "fwi_pair") == (pair, True)
    with pytest.raises(StencilError):
        ss.set_tuned_shape(g, "fwi_pair", dict(shp, tile_x=256, tile_y=16))
    ss.clear_tuning()
    assert not ss.tuned_shape(g, "fwi_adjoint")[1]
    # untuned, the pair maps the default shape onto a built one
    got, hit = ss.tuned_shape(g, "fwi_pair")
    assert not hit and (got["tile_x"], got["tile_y"]) in {(128, 8), (128, 16), (256, 4), (256, 8)}


def test_hip_errors_are_reported_not_raised():
    """A HIP failure comes back as STENCIL_ERR_HIP with the runtime's code
    kept for stencil_last_hip_error (no exit, no exception); with no device
    that is the coefficient upload's device query (hipErrorNoDevice)."""
    import torch

    import stencil_solver_amd as ss

    L = ss.lib()
    r = L.stencil_upload_fd_coefficients()
    if torch.cuda.device_count() == 0:
        assert r == 2 and L.stencil_strerror(r) == b"HIP runtime error"
        assert L.stencil_last_hip_error() == 100  # hipErrorNoDevice
    else:
        assert r == 0
