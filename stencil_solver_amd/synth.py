"""Seeded synthetic inputs for the batched-shot and FWI workloads (no
reference counterpart: the reference only builds constant / two-layer models
and a Gaussian pulse, include/stencil/grid.hpp:57-99).

Noise is counter-based (SplitMix64 of seed, stream and cell index, the same
generator as include/stencil/grid.hpp's host helpers in this package), so a
model depends only on its seed and shape.  The smooth models are that white
noise passed through a separable Gaussian filter (sigma in cells, truncated at
4 sigma, edges clamped), then mapped linearly onto [vmin, vmax].
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def u01(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """U[0, 1) float32 for cell indices start .. start+n-1: the top 24 bits of
    splitmix64(splitmix64(seed ^ splitmix64(stream)) + index)."""
    key = splitmix64(np.uint64(seed) ^ splitmix64(np.array([stream], dtype=np.uint64)))[0]
    with np.errstate(over="ignore"):
        idx = np.arange(start, start + n, dtype=np.uint64) + key
    bits = splitmix64(idx) >> np.uint64(40)
    return (bits.astype(np.float64) * 2.0 ** -24).astype(np.float32)


def gaussian_taps(sigma: float, truncate: float = 4.0) -> np.ndarray:
    """Normalised taps exp(-r^2 / 2 sigma^2), |r| <= int(truncate * sigma + 0.5)."""
    rad = int(truncate * sigma + 0.5)
    r = np.arange(-rad, rad + 1, dtype=np.float64)
    w = np.exp(-0.5 * (r / sigma) ** 2)
    return w / w.sum()


def gaussian_filter(x, sigma: float, truncate: float = 4.0):
    """Separable Gaussian filter of a (nz, ny, nx) torch tensor, edges clamped
    (scipy.ndimage.gaussian_filter with mode='nearest').  Each axis pass is a
    sum of shifted slices of the edge-padded field (one multiply-add per tap),
    so it runs as plain elementwise kernels on any device."""
    import torch

    taps = gaussian_taps(sigma, truncate)
    rad = (len(taps) - 1) // 2
    out = x
    for axis in (2, 1, 0):
        n = out.shape[axis]
        lo = out.narrow(axis, 0, 1)
        hi = out.narrow(axis, n - 1, 1)
        rep = [1, 1, 1]
        rep[axis] = rad
        padded = torch.cat([lo.repeat(rep), out, hi.repeat(rep)], dim=axis)
        acc = torch.zeros_like(out)
        for k, w in enumerate(taps):
            acc.add_(padded.narrow(axis, k, n), alpha=float(w))
        out = acc
    return out


def _s64(c: int) -> int:
    """A uint64 constant as the int64 with the same bits."""
    return c - (1 << 64) if c >= (1 << 63) else c


def _splitmix64_torch(z):
    """splitmix64 on int64 tensors holding uint64 bit patterns (wrapping
    multiply; right shifts made logical by masking)."""
    def srl(v, k):
        return (v >> k) & ((1 << (64 - k)) - 1)

    z = z + _s64(0x9E3779B97F4A7C15)
    z = (z ^ srl(z, 30)) * _s64(0xBF58476D1CE4E5B9)
    z = (z ^ srl(z, 27)) * _s64(0x94D049BB133111EB)
    return z ^ srl(z, 31)


def u01_torch(seed: int, stream: int, n: int, device="cpu"):
    """u01() computed on `device` (same values)."""
    import torch

    key = int(splitmix64(np.uint64(seed) ^ splitmix64(np.array([stream], dtype=np.uint64)))[0])
    idx = torch.arange(n, dtype=torch.int64, device=device) + _s64(key)
    bits = (_splitmix64_torch(idx) >> 40) & ((1 << 24) - 1)
    return bits.to(torch.float32) * (2.0 ** -24)


def smooth_velocity(shape, sigma: float, vmin: float, vmax: float, seed: int, device="cpu", stream: int = 0):
    """(nz, ny, nx) float32 velocity: seeded white noise, Gaussian-filtered,
    mapped linearly so its minimum is vmin and its maximum vmax."""
    import torch

    nz, ny, nx = shape
    noise = u01_torch(seed, stream, nz * ny * nx, device).reshape(shape)
    f = gaussian_filter(noise, sigma)
    lo, hi = f.min(), f.max()
    return (vmin + (vmax - vmin) * (f - lo) / (hi - lo)).to(torch.float32).contiguous()


def sphere_mask(shape, center, radius: float, device="cpu"):
    """Cells whose centre lies within `radius` cells of `center` (z, y, x)."""
    import torch

    nz, ny, nx = shape
    z = torch.arange(nz, device=device).view(-1, 1, 1)
    y = torch.arange(ny, device=device).view(1, -1, 1)
    x = torch.arange(nx, device=device).view(1, 1, -1)
    cz, cy, cx = center
    return (z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2 <= radius * radius


__all__ = ["splitmix64", "u01", "gaussian_taps", "gaussian_filter", "smooth_velocity", "sphere_mask"]
