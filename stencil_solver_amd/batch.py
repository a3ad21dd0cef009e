"""Batched multi-shot forward modelling (include/stencil_capi.h,
stencil_batch_*): many shots on their own grids stepped together -- one
launch of the tiled kernel per time step over a work-item table spanning
every shot, the whole run one hipGraph launch.  No reference counterpart (the
reference runs one grid per process).

Each shot is the FWI forward scheme (stencil_solver_amd.fwi) from zero fields:
p^{n+1} = 2 p^n - p^{n-1} + dt^2 vel2 L p^n + wavelet[n] at the source cell.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, stencil_shot, stencil_tile_shape
from .api import GridDims, _shape_dict, _stream


class ShotBatch:
    """Shots = list of dicts with keys ``grid`` (GridDims), ``vel2`` (device
    FP32 tensor (nz, ny, nx)), ``wavelet`` (n_steps samples or None) and
    ``src`` ((z, y, x) interior cell).  The three time levels of every shot are
    allocated here; the run graph is captured at construction."""

    def __init__(self, shots, n_steps: int, device, tune: bool = True, stream=None):
        import torch

        self.torch = torch
        self.device = torch.device(device)
        self.n_steps = int(n_steps)
        self.rings, self.vel2, self.wavelets = [], [], []
        arr = (stencil_shot * len(shots))()
        for i, sh in enumerate(shots):
            g: GridDims = sh["grid"]
            shape = (g.nz, g.ny, g.nx)
            v = sh["vel2"]
            if tuple(v.shape) != shape or v.dtype != torch.float32 or not v.is_contiguous():
                raise ValueError(f"shot {i}: vel2 must be a contiguous float32 tensor of shape {shape}")
            ring = [torch.empty(shape, dtype=torch.float32, device=self.device) for _ in range(3)]
            w = None
            if sh.get("wavelet") is not None:
                w = torch.as_tensor(np.asarray(sh["wavelet"], dtype=np.float32)).to(self.device).contiguous()
                if w.numel() != self.n_steps:
                    raise ValueError(f"shot {i}: wavelet needs n_steps = {self.n_steps} samples")
            self.rings.append(ring)
            self.vel2.append(v)
            self.wavelets.append(w)
            sz, sy, sx = sh.get("src", (0, 0, 0))
            arr[i].grid = g.to_c()
            for r in range(3):
                arr[i].ring[r] = ring[r].data_ptr()
            arr[i].vel2 = v.data_ptr()
            arr[i].wavelet = w.data_ptr() if w is not None else None
            arr[i].src_x, arr[i].src_y, arr[i].src_z = int(sx), int(sy), int(sz)
        self._shots = arr
        h = C.c_void_p()
        check(lib().stencil_batch_create(arr, len(shots), self.n_steps, int(bool(tune)), C.byref(h),
                                         _stream(stream)), "batch_create")
        self._h = h

    def run(self, stream=None) -> None:
        """All shots, n_steps each, from zero fields: one graph launch."""
        check(lib().stencil_batch_run(self._h, _stream(stream)), "batch_run")

    def check(self, stream=None) -> None:
        """Waits for the stream; raises StencilError when a grid barrier of
        the last persistent run timed out (its fields are then invalid)."""
        check(lib().stencil_batch_check(self._h, _stream(stream)), "batch_check")

    @property
    def ragged(self) -> bool:
        """True when every step of every shot is one launch (the default);
        False for the grouped mode (STENCIL_BATCH_MODE=grouped at creation)."""
        return lib().stencil_batch_is_ragged(self._h) == 1

    @property
    def persistent_grid(self) -> int:
        """Workgroups of the one persistent launch that steps every shot
        (STENCIL_BATCH_MODE=persistent at creation); 0 when the run is one
        launch per step."""
        r = lib().stencil_batch_persistent_grid(self._h)
        if r < 0:
            check(r, "batch_persistent_grid")
        return r

    def set_plan(self, tile_x: int, tile_y: int, chunk_planes: int, xcd_runs: bool = True, stream=None) -> None:
        """Re-plan the ragged launch: tiles of tile_x x tile_y cells (a built
        shape), z-chunks of about chunk_planes planes, the per-XCD runs on or
        off; recaptures the graph (synchronises)."""
        if tile_x % 4:
            raise ValueError("tile_x must be a multiple of 4 cells")
        shp = stencil_tile_shape(tile_x // 4, int(tile_y), 0, int(bool(xcd_runs)), 1)
        check(lib().stencil_batch_set_plan(self._h, C.byref(shp), int(chunk_planes), _stream(stream)),
              "batch_set_plan")

    def newest(self, i: int):
        """p^{n_steps} of shot i (valid after run())."""
        r = lib().stencil_batch_newest(self._h, int(i))
        if r < 0:
            check(r, "batch_newest")
        return self.rings[i][r]

    def groups(self) -> list[dict]:
        out = []
        for gi in range(lib().stencil_batch_groups(self._h)):
            g = type(self._shots[0].grid)()
            shp = stencil_tile_shape()
            n = C.c_int()
            check(lib().stencil_batch_group(self._h, gi, C.byref(g), C.byref(shp), C.byref(n)), "batch_group")
            out.append({"nx": g.nx, "ny": g.ny, "nz": g.nz, "shots": n.value, "shape": _shape_dict(shp)})
        return out

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().stencil_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["ShotBatch"]
