"""Adjoint-state FWI gradient through the C-ABI (include/stencil_capi.h,
stencil_fwi_*), and its torch.autograd binding.

The scheme (no reference counterpart; the reference models forward only): with
m = vel2 = v^2, a point source of amplitude wavelet[n] added to p^{n+1}, and
the receiver plane rec_z recorded at every step,

    p^{n+1} = 2 p^n - p^{n-1} + dt^2 m L p^n + wavelet[n] e_src,   p^0 = p^-1 = 0
    J(m)    = 1/2 sum_k |p^k(rec_z) - d_obs^k|^2

``FwiProblem.misfit`` runs the checkpointed forward pass, ``gradient`` the
adjoint sweep (dJ/dm on the full grid, zero on the halo).  ``FwiMisfit`` wraps
the pair as a torch.autograd.Function of vel2, so ``J.backward()`` fills
``vel2.grad`` (and a velocity v with vel2 = v * v gets dJ/dv = 2 v dJ/dm by
autograd).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import check, lib, stencil_fwi_config, stencil_tile_shape
from .api import GridDims, _dev_ptr, _shape_dict, _stream, _variant, tuned_shape


def ricker(n_steps: int, dt: float, f_peak: float, t0: float | None = None, amplitude: float = 1.0) -> np.ndarray:
    """Ricker wavelet (1 - 2 a) exp(-a), a = (pi f (t - t0))^2, sampled at
    t = n dt; t0 defaults to 1.2 / f (the wavelet starts near zero)."""
    t0 = 1.2 / f_peak if t0 is None else t0
    t = np.arange(n_steps, dtype=np.float64) * dt - t0
    a = (np.pi * f_peak * t) ** 2
    return amplitude * (1.0 - 2.0 * a) * np.exp(-a)


class FwiProblem:
    """One acquisition (grid, steps, source cell and wavelet, receiver plane)
    with its device workspace.  ``dtype`` is torch.float32 or torch.float64;
    ``variant`` picks the FP32 step kernels ("tiled": the LDS z-queue kernel,
    "naive": the global-memory one; both give bit-identical results, FP64
    always uses the latter).  ``tune`` runs the FWI kernels' launch-shape
    autotuner on this grid first (FP32 tiled only); the steps then use the
    tuned shapes."""

    def __init__(self, grid: GridDims, n_steps: int, src: tuple[int, int, int], rec_z: int, wavelet,
                 device, dtype=None, ckpt_every: int = 0, variant: str = "tiled", tune: bool = False):
        import torch

        self.torch = torch
        self.grid = grid
        self.dtype = dtype or torch.float32
        if self.dtype not in (torch.float32, torch.float64):
            raise ValueError("FwiProblem dtype must be float32 or float64")
        self.device = torch.device(device)
        self.n_steps = int(n_steps)
        sz, sy, sx = (int(v) for v in src)
        self.src = (sz, sy, sx)
        self.rec_z = int(rec_z)
        if variant not in ("naive", "tiled"):
            raise ValueError("variant must be 'naive' or 'tiled'")
        self.variant = variant
        self.cfg = stencil_fwi_config(self.n_steps, sx, sy, sz, self.rec_z, int(ckpt_every),
                                      0 if variant == "naive" else 1)
        self._g = grid.to_c()
        esz = 4 if self.dtype == torch.float32 else 8
        nb = int(lib().stencil_fwi_workspace_bytes(C.byref(self._g), C.byref(self.cfg), esz))
        if nb == 0:
            raise ValueError("invalid FWI configuration (source / receiver plane must be interior cells)")
        self.workspace_bytes = nb
        # 256-byte aligned: torch's caching allocator returns 512-byte aligned blocks
        self.work = torch.empty(nb, dtype=torch.uint8, device=self.device)
        assert self.work.data_ptr() % 256 == 0
        w = torch.as_tensor(np.asarray(wavelet, dtype=np.float64), dtype=self.dtype)
        if w.numel() != self.n_steps:
            raise ValueError(f"wavelet has {w.numel()} samples, need n_steps = {self.n_steps}")
        self.wavelet = w.to(self.device).contiguous()
        self.misfit_dev = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._have_residual = False
        # Bumped by every call that overwrites the workspace's checkpoints and
        # residual (misfit, traces): a gradient belongs to the misfit call of
        # the same generation, and to the vel2 tensor (storage and version)
        # that call saw.
        self.generation = 0
        self._misfit_vel2 = None
        self._sfx = "f32" if self.dtype == torch.float32 else "f64"
        self.shape = None
        if tune and self.dtype == torch.float32 and variant == "tiled":
            self.shape = self.tune()

    def tune(self, reps: int = 3, csv_path=None) -> dict:
        """Autotune the FP32 FWI kernels for this grid on the problem's own
        workspace (a forward pass's checkpoints, traces and stored L p stay
        intact); returns {"forward", "adjoint", "recompute", "pair"} shapes:
        the forward pass's steps, the adjoint steps, the forward steps the
        backward pass recomputes beside them on a second stream, and the step
        pair (a recompute and an adjoint step in one launch; None with a
        single segment, where nothing is recomputed)."""
        if self.dtype != self.torch.float32 or self.variant != "tiled":
            raise ValueError("tune() applies to the FP32 tiled FWI kernels")
        shp = [stencil_tile_shape() for _ in range(3)]
        path = str(csv_path).encode() if csv_path else None
        check(lib().stencil_tune_fwi_f32(C.byref(self._g), C.byref(self.cfg), self.work.data_ptr(), int(reps),
                                         *[C.byref(x) for x in shp], path, _stream(None)), "tune_fwi")
        out = dict(zip(("forward", "adjoint", "recompute"), (_shape_dict(x) for x in shp)))
        pair, hit = tuned_shape(self.grid, "fwi_pair")
        out["pair"] = pair if hit else None
        return out

    @property
    def ckpt_every(self) -> int:
        return self.cfg.ckpt_every or int(math.ceil(math.sqrt(self.n_steps)))

    def _check_field(self, t, name):
        shape = (self.grid.nz, self.grid.ny, self.grid.nx)
        if tuple(t.shape) != shape or t.dtype != self.dtype or t.device != self.device or not t.is_contiguous():
            raise ValueError(f"{name}: need a contiguous {self.dtype} tensor of shape {shape} on {self.device}")

    def _check_traces(self, t, name):
        shape = (self.n_steps, self.grid.ny, self.grid.nx)
        if tuple(t.shape) != shape or t.dtype != self.dtype or t.device != self.device or not t.is_contiguous():
            raise ValueError(f"{name}: need a contiguous {self.dtype} tensor of shape {shape} on {self.device}")

    def traces(self, vel2, stream=None):
        """Forward modelling: the receiver traces (n_steps, ny, nx)."""
        self._check_field(vel2, "vel2")
        out = self.torch.empty((self.n_steps, self.grid.ny, self.grid.nx), dtype=self.dtype, device=self.device)
        fn = getattr(lib(), f"stencil_fwi_forward_{self._sfx}")
        check(fn(vel2.data_ptr(), C.byref(self._g), C.byref(self.cfg), self.wavelet.data_ptr(), None,
                 out.data_ptr(), None, self.work.data_ptr(), _stream(stream)), "fwi_forward")
        self.generation += 1
        self._have_residual = False
        self._misfit_vel2 = None
        return out

    def misfit(self, vel2, d_obs, stream=None):
        """J = 1/2 |traces(vel2) - d_obs|^2 as a 1-element float64 device tensor;
        keeps the checkpoints and residual for ``gradient``."""
        self._check_field(vel2, "vel2")
        self._check_traces(d_obs, "d_obs")
        fn = getattr(lib(), f"stencil_fwi_forward_{self._sfx}")
        check(fn(vel2.data_ptr(), C.byref(self._g), C.byref(self.cfg), self.wavelet.data_ptr(), d_obs.data_ptr(),
                 None, self.misfit_dev.data_ptr(), self.work.data_ptr(), _stream(stream)), "fwi_forward")
        self.generation += 1
        self._have_residual = True
        self._misfit_vel2 = (vel2.data_ptr(), vel2._version)
        return self.misfit_dev.clone()

    def gradient(self, vel2, stream=None, src_trace: bool = False):
        """dJ/dvel2 of the last ``misfit`` call, which must have run on this
        same vel2 tensor, unmodified since; with src_trace the adjoint field
        at the source cell, lambda^{n+1}, is returned as well."""
        if not self._have_residual:
            raise RuntimeError("gradient() needs a preceding misfit() on this problem")
        self._check_field(vel2, "vel2")
        if self._misfit_vel2 != (vel2.data_ptr(), vel2._version):
            raise RuntimeError("gradient(): vel2 is not the tensor of the last misfit() call on this problem "
                               "(or it changed since); run misfit() on it first")
        grad = self.torch.empty_like(vel2)
        tr = self.torch.empty(self.n_steps, dtype=self.dtype, device=self.device) if src_trace else None
        fn = getattr(lib(), f"stencil_fwi_backward_{self._sfx}")
        check(fn(vel2.data_ptr(), C.byref(self._g), C.byref(self.cfg), self.wavelet.data_ptr(), grad.data_ptr(),
                 tr.data_ptr() if tr is not None else None, self.work.data_ptr(), _stream(stream)),
              "fwi_backward")
        return (grad, tr) if src_trace else grad


def _sfx_of(t) -> str:
    import torch

    if t.dtype == torch.float32:
        return "f32"
    if t.dtype == torch.float64:
        return "f64"
    raise TypeError("fields must be float32 or float64")


def _aux_ptr(t, like, name: str, min_numel: int) -> int:
    """Device pointer of an auxiliary array that the kernels index by a
    caller-chosen offset: it must sit on the fields' device, be contiguous,
    of the field dtype, and hold at least min_numel values."""
    import torch

    if not isinstance(t, torch.Tensor) or t.device != like.device:
        raise ValueError(f"{name} must be a tensor on {like.device}")
    if t.dtype != like.dtype or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous {like.dtype} tensor")
    if t.numel() < min_numel:
        raise ValueError(f"{name} has {t.numel()} values, needs at least {min_numel}")
    return t.data_ptr()


def forward_step(p_cur, p_prev, vel2, p_next, grid: GridDims, lap_out=None, src_index: int = -1, wavelet=None,
                 n: int = 0, variant: str = "tiled", stream=None) -> None:
    """One forward FWI step (stencil_forward_step_*): p_next = 2 p_cur - p_prev
    + dt^2 vel2 L p_cur in the reference association, plus wavelet[n] at the
    linear cell src_index (< 0: no source); L p_cur into lap_out when given."""
    sfx = _sfx_of(p_cur)
    dt = p_cur.dtype
    ptr = [_dev_ptr(t, grid, dt, nm) for t, nm in ((p_cur, "p_cur"), (p_prev, "p_prev"), (vel2, "vel2"),
                                                       (p_next, "p_next"))]
    lp = _dev_ptr(lap_out, grid, dt, "lap_out") if lap_out is not None else None
    w = None
    if int(src_index) >= 0:
        if int(src_index) >= grid.size():
            raise ValueError(f"src_index {src_index} is outside the grid ({grid.size()} cells)")
        if wavelet is None or int(n) < 0:
            raise ValueError("a source (src_index >= 0) needs a wavelet and a sample index n >= 0")
        w = _aux_ptr(wavelet, p_cur, "wavelet", int(n) + 1)
    elif wavelet is not None:
        w = _aux_ptr(wavelet, p_cur, "wavelet", 1)
    g = grid.to_c()
    check(getattr(lib(), f"stencil_forward_step_{sfx}")(*ptr, lp, C.byref(g), int(src_index), w, int(n),
                                                        _variant(variant), _stream(stream)), "forward_step")


def adjoint_step(psi1, psi2, vel2, psi_out, grid: GridDims, lap=None, grad=None, res_plane=None, rec_z: int = 0,
                 src_index: int = -1, src_value=None, variant: str = "tiled", stream=None) -> None:
    """One adjoint step in psi = vel2 * lambda (stencil_adjoint_step_*):
    psi_out = 2 psi1 - psi2 + dt^2 vel2 L psi1 (+ vel2 * res_plane on plane
    rec_z), grad += dt^2 lap psi1 when lap is given, src_value[0] =
    psi_out / vel2 at src_index when given."""
    sfx = _sfx_of(psi1)
    dt = psi1.dtype
    ptr = [_dev_ptr(t, grid, dt, nm) for t, nm in ((psi1, "psi1"), (psi2, "psi2"), (vel2, "vel2"))]
    lp = _dev_ptr(lap, grid, dt, "lap") if lap is not None else None
    gp = _dev_ptr(grad, grid, dt, "grad") if grad is not None else None
    op = _dev_ptr(psi_out, grid, dt, "psi_out")
    rp = None
    if res_plane is not None:
        if res_plane.numel() != grid.nx * grid.ny:
            raise ValueError(f"res_plane must hold nx*ny = {grid.nx * grid.ny} values")
        rp = _aux_ptr(res_plane, psi1, "res_plane", grid.nx * grid.ny)
    sv = None
    if src_value is not None:
        if not 0 <= int(src_index) < grid.size():
            raise ValueError(f"src_value needs a src_index inside the grid, got {src_index}")
        sv = _aux_ptr(src_value, psi1, "src_value", 1)
    g = grid.to_c()
    check(getattr(lib(), f"stencil_adjoint_step_{sfx}")(*ptr, lp, gp, op, rp, int(rec_z), C.byref(g),
                                                        int(src_index), sv, _variant(variant), _stream(stream)),
          "adjoint_step")


def _autograd_function():
    import torch

    class FwiMisfit(torch.autograd.Function):
        """J(vel2) = 1/2 |traces(vel2) - d_obs|^2 with the adjoint-state gradient."""

        @staticmethod
        def forward(ctx, vel2, problem: FwiProblem, d_obs):
            v = vel2.detach().contiguous()
            J = problem.misfit(v, d_obs)
            ctx.problem = problem
            ctx.generation = problem.generation
            ctx.save_for_backward(v, d_obs)
            return J.reshape(())

        @staticmethod
        def backward(ctx, grad_out):
            v, d_obs = ctx.saved_tensors
            pr = ctx.problem
            if pr.generation != ctx.generation:
                # Another misfit/traces call on this problem overwrote the
                # checkpoints and residual since this forward: rebuild them.
                pr.misfit(v, d_obs)
            g = pr.gradient(v)
            return grad_out.to(g.dtype) * g, None, None

    return FwiMisfit


_FWI_FN = None


def fwi_misfit(vel2, problem: FwiProblem, d_obs):
    """Differentiable misfit J(vel2) (torch.autograd.Function)."""
    global _FWI_FN
    if _FWI_FN is None:
        _FWI_FN = _autograd_function()
    return _FWI_FN.apply(vel2, problem, d_obs)


__all__ = ["ricker", "FwiProblem", "fwi_misfit", "forward_step", "adjoint_step"]
