"""Python mirror of the reference's operator and solver API over the C-ABI.

* ``launch_naive / launch_tiled / launch / upload_fd_coefficients`` -- the
  launcher API (reference include/stencil/kernels.cuh:33-83) on torch device
  tensors; asynchronous on the current torch stream unless ``stream`` is given.
* ``heat_step`` -- the dedicated heat kernel (replaces the heat step of
  src/solver/heat_diffusion.cpp:103-122).
* ``Config`` / ``make_solver`` / ``Solver`` -- the solver API (reference
  include/stencil/config.hpp:34-88, include/stencil/solver.hpp:13-24).

Fields are torch tensors of shape (nz, ny, nx), C-contiguous, so that the
flat index is z*nx*ny + y*nx + x as in the reference layout.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields
from typing import Optional

import numpy as np

from . import _lib
from ._lib import (StencilError, check, lib, stencil_config, stencil_field16, stencil_grid, stencil_perf,
                   stencil_tile_shape)

VARIANTS = {"naive": 0, "tiled": 1, "multigpu": 2}
RADIUS = 4


def _variant(v) -> int:
    if isinstance(v, str):
        if v not in VARIANTS:
            raise ValueError(f"unknown kernel variant {v!r}")
        return VARIANTS[v]
    return int(v)


@dataclass
class GridDims:
    """Reference GridDims (include/stencil/types.hpp:115-140)."""

    nx: int
    ny: int
    nz: int
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0
    dt: float = 1.0

    @classmethod
    def cube(cls, n: int, h: float = 1.0, dt: float = 1.0) -> "GridDims":
        return cls(n, n, n, h, h, h, dt)

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    def size(self) -> int:
        return self.nx * self.ny * self.nz

    def interior_size(self) -> int:
        return max(self.nx - 2 * RADIUS, 0) * max(self.ny - 2 * RADIUS, 0) * max(self.nz - 2 * RADIUS, 0)

    def to_c(self) -> stencil_grid:
        return stencil_grid(self.nx, self.ny, self.nz, float(self.dx), float(self.dy), float(self.dz),
                            float(self.dt))


def _dev_ptr(t, grid: GridDims, dtype, name: str) -> int:
    import torch

    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a torch tensor on a HIP device")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous() or t.numel() != grid.size():
        raise ValueError(f"{name} must be contiguous with nx*ny*nz = {grid.size()} elements")
    return t.data_ptr()


def _stream(stream) -> int:
    import torch

    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def upload_fd_coefficients() -> None:
    check(lib().stencil_upload_fd_coefficients(), "upload_fd_coefficients")


def _wave(fn32, fn64, p_cur, p_prev, vel2, p_next, grid: GridDims, stream, what, extra=()):
    import torch

    dtype = p_cur.dtype
    if dtype not in (torch.float32, torch.float64):
        raise TypeError("fields must be float32 or float64")
    fn = fn32 if dtype == torch.float32 else fn64
    if fn is None:
        raise TypeError(f"{what} has no {dtype} form")
    ptrs = [_dev_ptr(t, grid, dtype, n) for t, n in
            ((p_cur, "p_cur"), (p_prev, "p_prev"), (vel2, "vel2"), (p_next, "p_next"))]
    g = grid.to_c()
    check(fn(*ptrs, C.byref(g), *extra, _stream(stream)), what)


def launch_naive(p_cur, p_prev, vel2, p_next, grid: GridDims, stream=None) -> None:
    """One acoustic step, coalesced kernel, reference association (kernels.cuh:42)."""
    L = lib()
    _wave(L.stencil_launch_naive_f32, L.stencil_launch_naive_f64, p_cur, p_prev, vel2, p_next, grid,
          stream, "launch_naive")


def launch_tiled(p_cur, p_prev, vel2, p_next, grid: GridDims, stream=None) -> None:
    """One acoustic step, LDS + z-register-queue kernel (kernels.cuh:54)."""
    L = lib()
    _wave(L.stencil_launch_tiled_f32, L.stencil_launch_tiled_f64, p_cur, p_prev, vel2, p_next, grid,
          stream, "launch_tiled")


def launch(variant, p_cur, p_prev, vel2, p_next, grid: GridDims, stream=None) -> None:
    """Dispatch by KernelVariant (kernels.cuh:68-83)."""
    v = _variant(variant)
    if v == 0:
        launch_naive(p_cur, p_prev, vel2, p_next, grid, stream)
    elif v in (1, 2):
        launch_tiled(p_cur, p_prev, vel2, p_next, grid, stream)
    else:
        raise ValueError(f"unknown kernel variant {variant!r}")


def tile_shape(tx4=0, ty=0, n_chunks=0, xcd_remap=-1, nontemporal=-1) -> stencil_tile_shape:
    """Launch shape of the tiled kernel (see include/stencil_capi.h)."""
    return stencil_tile_shape(int(tx4), int(ty), int(n_chunks), int(xcd_remap), int(nontemporal))


def launch_tiled_ex(p_cur, p_prev, vel2, p_next, grid: GridDims, exact=True, tx4=0, ty=0, n_chunks=0,
                    xcd_remap=-1, nontemporal=-1, stream=None) -> None:
    """The tiled step with an explicit launch shape (no reference counterpart)."""
    L = lib()
    shp = tile_shape(tx4, ty, n_chunks, xcd_remap, nontemporal)
    _wave(L.stencil_launch_tiled_ex_f32, None, p_cur, p_prev, vel2, p_next, grid, stream,
          "launch_tiled_ex", extra=(int(bool(exact)), C.byref(shp)))


def heat_step(t, t_next, alpha: float, grid: GridDims, variant="tiled", stream=None, exact=None,
              tx4=0, ty=0, n_chunks=0, xcd_remap=-1, nontemporal=-1) -> None:
    """t_next = t + (dt^2 alpha) Lap(t); grid.dt is the heat dt_eff."""
    import torch

    L = lib()
    g = grid.to_c()
    if t.dtype == torch.float64:
        ptrs = [_dev_ptr(x, grid, torch.float64, n) for x, n in ((t, "t"), (t_next, "t_next"))]
        check(L.stencil_heat_step_f64(*ptrs, float(alpha), C.byref(g), _variant(variant), _stream(stream)),
              "heat_step")
        return
    ptrs = [_dev_ptr(x, grid, torch.float32, n) for x, n in ((t, "t"), (t_next, "t_next"))]
    if exact is None:
        check(L.stencil_heat_step_f32(*ptrs, float(alpha), C.byref(g), _variant(variant), _stream(stream)),
              "heat_step")
    else:
        shp = tile_shape(tx4, ty, n_chunks, xcd_remap, nontemporal)
        check(L.stencil_heat_step_ex_f32(*ptrs, float(alpha), C.byref(g), int(bool(exact)), C.byref(shp),
                                         _stream(stream)), "heat_step_ex")


def _shape_dict(shp) -> dict:
    return {"tile_x": 4 * shp.tile_x4, "tile_y": shp.tile_y, "n_chunks": shp.n_chunks,
            "xcd_remap": shp.xcd_remap, "nontemporal": shp.nontemporal}


def tune_tiled(p_cur, p_prev, vel2, p_next, grid: GridDims, reps: int = 5, csv_path=None, stream=None) -> dict:
    """Autotune the tiled wave kernel for this grid size (p_next is
    overwritten); later launch_tiled calls on these dims use the result."""
    import torch

    shp = stencil_tile_shape()
    ptrs = [_dev_ptr(t, grid, torch.float32, n) for t, n in
            ((p_cur, "p_cur"), (p_prev, "p_prev"), (vel2, "vel2"), (p_next, "p_next"))]
    g = grid.to_c()
    path = str(csv_path).encode() if csv_path else None
    check(lib().stencil_tune_tiled_f32(*ptrs, C.byref(g), int(reps), C.byref(shp), path, _stream(stream)),
          "tune_tiled")
    return _shape_dict(shp)


def tune_heat(t, t_next, grid: GridDims, reps: int = 5, csv_path=None, stream=None) -> dict:
    import torch

    shp = stencil_tile_shape()
    ptrs = [_dev_ptr(x, grid, torch.float32, n) for x, n in ((t, "t"), (t_next, "t_next"))]
    g = grid.to_c()
    path = str(csv_path).encode() if csv_path else None
    check(lib().stencil_tune_heat_f32(*ptrs, C.byref(g), int(reps), C.byref(shp), path, _stream(stream)),
          "tune_heat")
    return _shape_dict(shp)


# Tuner-cache kinds (the C-ABI's kind numbers).
_SHAPE_KINDS = {"wave": 0, "heat": 1, "bf16": 3, "fp16": 4, "bfp16": 5, "fwi_forward": 6, "fwi_adjoint": 7,
                "fwi_recompute": 8, "fwi_pair": 9}


def tuned_shape(grid: GridDims, kind: str = "wave"):
    """(shape dict, from_tuner) a NULL-shape launch would use on these dims;
    kind "wave", "heat", a 16-bit storage name for launch_tiled_16, or
    "fwi_forward" / "fwi_adjoint" / "fwi_recompute" / "fwi_pair" (the
    backward pass's recompute + adjoint step pair) for the FWI steps."""
    shp = stencil_tile_shape()
    g = grid.to_c()
    r = lib().stencil_tuned_shape(C.byref(g), _SHAPE_KINDS[kind], C.byref(shp))
    if r < 0 or r > 1:
        check(r, "tuned_shape")
    return _shape_dict(shp), bool(r)


def set_tuned_shape(grid: GridDims, kind: str, shape: dict) -> None:
    """Put a launch shape (a dict as tuned_shape returns; tile_x in cells)
    into the tuner cache for these dims and kind, e.g. one saved from an
    earlier run's tuning."""
    if shape["tile_x"] % 4:
        raise ValueError("tile_x must be a multiple of 4 cells")
    shp = stencil_tile_shape(shape["tile_x"] // 4, shape["tile_y"], shape.get("n_chunks", 0),
                             shape.get("xcd_remap", -1), shape.get("nontemporal", -1))
    g = grid.to_c()
    check(lib().stencil_set_tuned_shape(C.byref(g), _SHAPE_KINDS[kind], C.byref(shp)), "set_tuned_shape")


def clear_tuning() -> None:
    lib().stencil_clear_tuning()


def wave_bytes_per_step(grid: GridDims, bytes_per_value: int = 4) -> float:
    g = grid.to_c()
    return float(lib().stencil_wave_bytes_per_step(C.byref(g), bytes_per_value))


def heat_bytes_per_step(grid: GridDims, bytes_per_value: int = 4) -> float:
    g = grid.to_c()
    return float(lib().stencil_heat_bytes_per_step(C.byref(g), bytes_per_value))


# ── 16-bit field storage ────────────────────────────────────────────────────
STORAGE = {"bf16": 1, "fp16": 2, "bfp16": 3}


class Field16:
    """A field in 16-bit storage on the device (include/stencil_capi.h,
    ``stencil_field16``): ``data`` int16 tensor of the grid's shape, ``exp``
    int8 block exponents (block floating point only), ``scale`` the
    power-of-two field scale (fp16 only)."""

    def __init__(self, fmt: str, shape, device, scale: float = 1.0):
        import torch

        if fmt not in STORAGE:
            raise ValueError(f"unknown 16-bit storage {fmt!r}")
        self.fmt, self.shape, self.scale = fmt, tuple(shape), float(scale)
        n = int(np.prod(shape))
        self.data = torch.zeros(shape, dtype=torch.int16, device=device)
        self.exp = torch.zeros(n // 16, dtype=torch.int8, device=device) if fmt == "bfp16" else None

    def c(self) -> stencil_field16:
        return stencil_field16(self.data.data_ptr(), self.exp.data_ptr() if self.exp is not None else None)

    def nbytes(self) -> int:
        return self.data.numel() * 2 + (self.exp.numel() if self.exp is not None else 0)


def pow2_scale_for(max_abs: float, target: float = 2.0 ** 14) -> float:
    """Power-of-two scale that maps max_abs to at most `target` (fp16 storage)."""
    import math

    if not max_abs > 0:
        return 1.0
    return float(2.0 ** math.floor(math.log2(target / max_abs)))


def encode16(fmt: str, src, dst: Optional[Field16] = None, scale: float = 1.0, stream=None) -> Field16:
    import torch

    if dst is None:
        dst = Field16(fmt, tuple(src.shape), src.device, scale)
    if src.dtype != torch.float32 or not src.is_contiguous():
        raise TypeError("encode16 takes a contiguous float32 tensor")
    f = dst.c()
    check(lib().stencil_encode16(STORAGE[fmt], src.data_ptr(), C.byref(f), src.numel(), float(dst.scale),
                                 _stream(stream)), "encode16")
    return dst


def decode16(field: Field16, out=None, stream=None):
    import torch

    if out is None:
        out = torch.empty(field.shape, dtype=torch.float32, device=field.data.device)
    f = field.c()
    check(lib().stencil_decode16(STORAGE[field.fmt], C.byref(f), out.data_ptr(), out.numel(), float(field.scale),
                                 _stream(stream)), "decode16")
    return out


def launch_tiled_16(p_cur: Field16, p_prev: Field16, vel2: Field16, p_next: Field16, grid: GridDims,
                    tile=None, stream=None, exact: bool = False) -> None:
    """One acoustic step on 16-bit fields, FP32 arithmetic.  exact=True uses the
    reference association (bit-reproducible against the CPU restatement);
    the default fused-weight form differs by FP32 rounding only."""
    fmt = p_cur.fmt
    if any(f.fmt != fmt for f in (p_prev, vel2, p_next)):
        raise ValueError("all four fields must share one storage format")
    if p_prev.scale != p_cur.scale or p_next.scale != p_cur.scale:
        raise ValueError("p_cur, p_prev and p_next must share one scale")
    fs = [f.c() for f in (p_cur, p_prev, vel2, p_next)]
    g = grid.to_c()
    shp = C.byref(tile) if tile is not None else None
    check(lib().stencil_launch_tiled_16(STORAGE[fmt], *[C.byref(f) for f in fs], C.byref(g), float(p_cur.scale),
                                        float(vel2.scale), int(bool(exact)), shp, _stream(stream)),
          "launch_tiled_16")


def tune_tiled_16(p_cur: Field16, p_prev: Field16, vel2: Field16, p_next: Field16, grid: GridDims,
                  reps: int = 5, csv_path=None, stream=None) -> dict:
    """Autotune launch_tiled_16 (fused form) for this grid and storage format
    (p_next is overwritten); later NULL-shape calls use the result."""
    fmt = p_cur.fmt
    if any(f.fmt != fmt for f in (p_prev, vel2, p_next)):
        raise ValueError("all four fields must share one storage format")
    shp = stencil_tile_shape()
    fs = [f.c() for f in (p_cur, p_prev, vel2, p_next)]
    g = grid.to_c()
    path = str(csv_path).encode() if csv_path else None
    check(lib().stencil_tune_tiled_16(STORAGE[fmt], *[C.byref(f) for f in fs], C.byref(g), float(p_cur.scale),
                                      float(vel2.scale), int(reps), C.byref(shp), path, _stream(stream)),
          "tune_tiled_16")
    return _shape_dict(shp)


class Reducer:
    """Deterministic FP64 reductions over the interior (include/stencil_capi.h)."""

    def __init__(self, device):
        import torch

        nb = int(lib().stencil_reduce_workspace_bytes())
        self.work = torch.empty(nb // 8, dtype=torch.float64, device=device)
        self.out = torch.zeros(3, dtype=torch.float64, device=device)

    def _run(self, fn, ptrs, grid, stream):
        g = grid.to_c()
        check(fn(*ptrs, C.byref(g), self.out.data_ptr(), self.work.data_ptr(), _stream(stream)), fn.__name__)
        return self.out.cpu().tolist()

    def energy(self, p_n, p_np1, vel2, grid: GridDims, stream=None):
        """(kinetic, potential) of the leapfrog pair; their sum is the conserved energy."""
        import torch

        fn = lib().stencil_energy_f64 if p_n.dtype == torch.float64 else lib().stencil_energy_f32
        k, p, _ = self._run(fn, [p_n.data_ptr(), p_np1.data_ptr(), vel2.data_ptr()], grid, stream)
        return k, p

    def rel_l2(self, a32, b64, grid: GridDims, stream=None) -> float:
        num, den, _ = self._run(lib().stencil_diff_norms_f32_f64, [a32.data_ptr(), b64.data_ptr()], grid, stream)
        return (num / den) ** 0.5 if den > 0 else num ** 0.5

    def dot(self, a, b, grid: GridDims, stream=None) -> float:
        import torch

        fn = lib().stencil_dot_f64 if a.dtype == torch.float64 else lib().stencil_dot_f32
        return self._run(fn, [a.data_ptr(), b.data_ptr()], grid, stream)[0]


# ── Seeded host inputs ──────────────────────────────────────────────────────
def host_gaussian(shape, cx=0.5, cy=0.5, cz=0.5, sigma=5.0) -> np.ndarray:
    """init_gaussian (reference include/stencil/grid.hpp:57-78) into a new array."""
    nz, ny, nx = shape
    a = np.zeros(shape, dtype=np.float32)
    check(lib().stencil_host_init_gaussian_f32(a.ctypes.data, nx, ny, nz, cx, cy, cz, sigma), "host_gaussian")
    return a


def host_layered_noise_vel2(shape, n_layers=16, vmin=1500.0, vmax=4000.0, noise=0.05, seed=2) -> np.ndarray:
    nz, ny, nx = shape
    a = np.empty(shape, dtype=np.float32)
    check(lib().stencil_host_layered_noise_vel2_f32(a.ctypes.data, nx, ny, nz, n_layers, vmin, vmax, noise,
                                                    seed), "host_layered_noise_vel2")
    return a


def host_noise_cube(shape, side=64, hot=10.0, seed=2) -> np.ndarray:
    nz, ny, nx = shape
    a = np.empty(shape, dtype=np.float32)
    check(lib().stencil_host_noise_cube_f32(a.ctypes.data, nx, ny, nz, side, hot, seed), "host_noise_cube")
    return a


# ── Solver API ──────────────────────────────────────────────────────────────
@dataclass
class Config:
    """Reference Config (include/stencil/config.hpp:34-88) plus the additions
    of include/stencil/config.hpp (dt, velocity model, initial condition)."""

    nx: int = 256
    ny: int = 256
    nz: int = 256
    dx: float = 10.0
    dy: float = 10.0
    dz: float = 10.0
    dt: float = 0.0
    steps: int = 100
    warmup_steps: int = 10
    velocity: float = 1500.0
    diffusivity: float = float(np.float32(1.0e-4))
    pulse_cx: float = 0.5
    pulse_cy: float = 0.5
    pulse_cz: float = 0.5
    pulse_sigma: float = 5.0
    physics: str = "wave"
    variant: str = "naive"
    device: int = 0
    vmodel: str = "constant"
    n_layers: int = 16
    vmin: float = 1500.0
    vmax: float = 4000.0
    vnoise: float = 0.05
    ic: str = "gaussian"
    cube_side: int = 64
    cube_value: float = 10.0
    seed: int = 2
    fp64: bool = False

    def to_c(self) -> stencil_config:
        c = stencil_config()
        lib().stencil_config_default(C.byref(c))
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name == "physics":
                v = {"wave": 0, "heat": 1}[v]
            elif f.name == "variant":
                v = _variant(v)
            elif f.name == "vmodel":
                v = {"constant": 0, "layered-noise": 1}[v]
            elif f.name == "ic":
                v = {"gaussian": 0, "noise-cube": 1}[v]
            elif f.name == "fp64":
                v = int(bool(v))
            setattr(c, f.name, v)
        return c


@dataclass
class PerfResult:
    """Reference PerfResult (include/stencil/timer.hpp:88-96)."""

    label: str = ""
    time_ms: float = 0.0
    bandwidth_gbs: float = 0.0
    gflops: float = 0.0
    model_bandwidth_gbs: float = 0.0
    grid_size: int = 0


class Solver:
    """ISolver over a C-ABI handle (reference include/stencil/solver.hpp:13-22)."""

    def __init__(self, cfg: Config):
        self._h = C.c_void_p()
        c = cfg.to_c()
        check(lib().stencil_solver_create(C.byref(c), C.byref(self._h)), "make_solver")
        self.cfg = cfg

    def close(self):
        if self._h:
            lib().stencil_solver_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self) -> None:
        check(lib().stencil_solver_step(self._h), "step")

    def run(self, n_steps: int, warmup: int = 0) -> PerfResult:
        p = stencil_perf()
        check(lib().stencil_solver_run(self._h, n_steps, warmup, C.byref(p)), "run")
        return PerfResult(p.label.decode(), p.time_ms, p.bandwidth_gbs, p.gflops, p.model_bandwidth_gbs,
                          p.grid_size)

    def dims(self) -> GridDims:
        g = stencil_grid()
        check(lib().stencil_solver_dims(self._h, C.byref(g)), "dims")
        return GridDims(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.dt)

    def name(self) -> str:
        buf = C.create_string_buffer(64)
        check(lib().stencil_solver_name(self._h, buf, 64), "name")
        return buf.value.decode()

    def download(self) -> np.ndarray:
        d = self.dims()
        out = np.empty(d.shape, dtype=np.float64 if self.cfg.fp64 else np.float32)
        check(lib().stencil_solver_download(self._h, out.ctypes.data, out.nbytes), "download")
        return out


def make_solver(cfg: Config) -> Solver:
    """Reference make_solver (src/solver/solver_base.cpp:119-133)."""
    return Solver(cfg)


__all__ = ["GridDims", "Config", "PerfResult", "Solver", "StencilError", "make_solver", "upload_fd_coefficients",
           "launch_naive", "launch_tiled", "launch", "launch_tiled_ex", "tile_shape", "tune_tiled", "tune_heat",
           "tuned_shape", "clear_tuning", "Field16", "STORAGE", "encode16", "decode16", "launch_tiled_16",
           "tune_tiled_16",
           "pow2_scale_for", "Reducer", "heat_step", "wave_bytes_per_step",
           "heat_bytes_per_step", "host_gaussian", "host_layered_noise_vel2", "host_noise_cube", "VARIANTS"]
