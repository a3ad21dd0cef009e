"""Synthetic inputs of the batched-shot and FWI workloads: the noise generator
equals the oracle's counter-based generator, and the smoothing equals
scipy.ndimage.gaussian_filter (mode 'nearest', truncate 4)."""
from __future__ import annotations

import numpy as np
import pytest

from stencil_solver_amd import synth


def test_u01_matches_oracle_generator(oracle):
    rng = np.random.default_rng(0)
    for seed, stream in ((0, 0), (4, 1), (2 ** 40 + 7, 12345)):
        start = int(rng.integers(0, 2 ** 30))
        got = synth.u01(seed, stream, 64, start)
        want = np.array([oracle.lib.oracle_u01(seed, stream, start + i) for i in range(64)], np.float32)
        assert np.array_equal(got, want)


def test_u01_torch_matches_numpy():
    for seed, stream in ((0, 0), (4, 1), (2 ** 40 + 7, 12345), (2 ** 63 + 5, 2 ** 64 - 1)):
        assert np.array_equal(synth.u01_torch(seed, stream, 5000).numpy(), synth.u01(seed, stream, 5000))


def test_gaussian_filter_matches_scipy():
    import torch
    from scipy import ndimage

    x = np.random.default_rng(1).random((23, 19, 31))
    for sigma in (1.5, 3.0):
        got = synth.gaussian_filter(torch.from_numpy(x), sigma).numpy()
        want = ndimage.gaussian_filter(x, sigma, mode="nearest", truncate=4.0)
        assert np.max(np.abs(got - want)) < 1e-12


def test_smooth_velocity_range_and_determinism():
    v = synth.smooth_velocity((24, 20, 28), 3.0, 1500.0, 4500.0, seed=4)
    assert v.dtype.is_floating_point and tuple(v.shape) == (24, 20, 28)
    assert float(v.min()) == pytest.approx(1500.0, abs=1e-3)
    assert float(v.max()) == pytest.approx(4500.0, abs=1e-3)
    assert np.array_equal(v.numpy(), synth.smooth_velocity((24, 20, 28), 3.0, 1500.0, 4500.0, seed=4).numpy())
    assert not np.array_equal(v.numpy(), synth.smooth_velocity((24, 20, 28), 3.0, 1500.0, 4500.0, seed=5).numpy())
    # smooth: neighbour differences far below the range
    assert float((v[:, :, 1:] - v[:, :, :-1]).abs().max()) < 0.25 * 3000.0


def test_sphere_mask():
    m = synth.sphere_mask((11, 11, 11), (5, 5, 5), 2.0)
    assert int(m.sum()) == 33  # lattice points with r^2 <= 4
    assert bool(m[5, 5, 7]) and not bool(m[5, 5, 8])
