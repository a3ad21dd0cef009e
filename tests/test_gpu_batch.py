"""Batched multi-shot forward modelling on the GPU: every shot of a ragged
batch equals the oracle's single-shot run bit for bit (default, tuned and every
launch plan: each built tile shape with short, medium and whole-column
z-chunks, with and without the per-XCD runs, as one launch per step and as one
persistent launch; and the grouped mode at every built shape), reruns of the
captured graph are identical, and bad shots are rejected."""
from __future__ import annotations

import numpy as np
import pytest

from test_oracle_fwi import smooth_model

pytestmark = pytest.mark.gpu
R = 4
SHAPES = [(24, 24, 24), (28, 32, 36), (28, 32, 36), (40, 20, 28), (24, 24, 24), (22, 30, 44)]


@pytest.fixture(scope="module")
def ss(dev):
    import stencil_solver_amd

    return stencil_solver_amd


def _shots(ss, dev, n_steps, seed=0, with_source=True):
    import torch

    rng = np.random.default_rng(seed)
    shots, host = [], []
    for i, shape in enumerate(SHAPES):
        nz, ny, nx = shape
        m = smooth_model(shape, 100 + i, 1500.0, 3000.0, np.float32)
        src = tuple(int(rng.integers(R, s - R)) for s in shape)
        w = (1e3 * ss.ricker(n_steps, 1e-3, 25.0 + i)).astype(np.float32) if with_source or i % 2 else None
        grid = ss.GridDims(nx, ny, nz, np.float32(10.0), np.float32(10.0), np.float32(10.0), np.float32(1e-3))
        shots.append({"grid": grid, "vel2": torch.from_numpy(m).to(dev), "wavelet": w, "src": src})
        host.append((m, w, src))
    return shots, host


@pytest.fixture(params=["ragged", "persistent", "grouped"])
def mode(request, monkeypatch):
    if request.param == "ragged":
        monkeypatch.delenv("STENCIL_BATCH_MODE", raising=False)
    else:
        monkeypatch.setenv("STENCIL_BATCH_MODE", request.param)
    return request.param


def test_batch_matches_oracle_per_shot(ss, dev, oracle, mode):
    n = 30
    shots, host = _shots(ss, dev, n)
    b = ss.ShotBatch(shots, n, dev, tune=False)
    assert b.ragged == (mode != "grouped")
    assert (b.persistent_grid > 0) == (mode == "persistent")
    assert b.persistent_grid % 8 == 0
    b.run()
    for i, (m, w, src) in enumerate(host):
        want = oracle.shot_run(m, np.float32(10.0), np.float32(1e-3), n, src, w)
        got = b.newest(i).cpu().numpy()
        assert np.abs(want).max() > 0
        assert np.array_equal(got, want), i
    groups = b.groups()
    assert sorted(g["shots"] for g in groups) == [1, 1, 2, 2]
    b.close()


def test_persistent_barrier_timeout_is_reported(ss, dev, oracle, monkeypatch):
    """A persistent run whose barriers give up at their first poll
    (STENCIL_BATCH_BARRIER_POLLS=0, a limit set from the host) still ends --
    every workgroup leaves at its next barrier -- and check() reports the
    timeout; the same shots in a batch made without the limit, run after it,
    match the oracle."""
    from stencil_solver_amd._lib import StencilError

    n = 12
    shots, host = _shots(ss, dev, n)
    monkeypatch.setenv("STENCIL_BATCH_MODE", "persistent")
    monkeypatch.setenv("STENCIL_BATCH_BARRIER_POLLS", "0")
    b = ss.ShotBatch(shots, n, dev, tune=False)
    assert b.persistent_grid >= 8
    b.run()
    with pytest.raises(StencilError, match="timed out|TIMEOUT|time"):
        b.check()
    b.close()
    monkeypatch.delenv("STENCIL_BATCH_BARRIER_POLLS")
    b = ss.ShotBatch(shots, n, dev, tune=False)
    b.run()
    b.check()
    for i, (m, w, src) in enumerate(host):
        want = oracle.shot_run(m, np.float32(10.0), np.float32(1e-3), n, src, w)
        assert np.array_equal(b.newest(i).cpu().numpy(), want), i
    b.close()


def test_batch_rerun_and_tuned_shapes_are_identical(ss, dev, mode):
    n = 25
    shots, _ = _shots(ss, dev, n, seed=3)
    b1 = ss.ShotBatch(shots, n, dev, tune=False)
    b1.run()
    first = [b1.newest(i).clone() for i in range(len(shots))]
    b1.run()  # the graph starts again from zero fields
    for i in range(len(shots)):
        assert bool((b1.newest(i) == first[i]).all()), i
    b1.close()
    ss.clear_tuning()
    b2 = ss.ShotBatch(shots, n, dev, tune=True)
    if mode == "grouped":
        assert all(ss.tuned_shape(s["grid"])[1] for s in shots)
    else:  # one tuned tile shape for the whole batch
        tiles = {(g["shape"]["tile_x"], g["shape"]["tile_y"]) for g in b2.groups()}
        assert len(tiles) == 1
    b2.run()
    for i in range(len(shots)):
        assert bool((b2.newest(i) == first[i]).all()), i
    b2.close()
    ss.clear_tuning()


TILES = [(16, 16), (16, 32), (32, 8), (32, 16), (64, 4), (64, 8), (64, 12), (64, 16)]


@pytest.mark.parametrize("persistent", [False, True])
def test_ragged_batch_every_plan_matches_oracle(ss, dev, oracle, monkeypatch, persistent):
    """The ragged batch -- one launch per step, or one persistent launch for
    the run -- at every built tile shape, with z-chunks of 9 planes (several
    chunks per column, ragged last chunk), 18 and whole columns, with and
    without the per-XCD runs: every shot equals the oracle's single-shot run
    bit for bit."""
    if persistent:
        monkeypatch.setenv("STENCIL_BATCH_MODE", "persistent")
    else:
        monkeypatch.delenv("STENCIL_BATCH_MODE", raising=False)
    n = 12
    shots, host = _shots(ss, dev, n, seed=5)
    want = [oracle.shot_run(m, np.float32(10.0), np.float32(1e-3), n, src, w) for m, w, src in host]
    b = ss.ShotBatch(shots, n, dev, tune=False)
    assert b.ragged and (b.persistent_grid > 0) == persistent
    for tx4, ty in TILES:
        for chunk in (9, 18, 1000):
            for runs in (True, False):
                b.set_plan(4 * tx4, ty, chunk, runs)
                gs = b.groups()
                assert all(g["shape"]["tile_x"] == 4 * tx4 and g["shape"]["tile_y"] == ty for g in gs)
                if chunk == 1000:
                    assert all(g["shape"]["n_chunks"] == 1 for g in gs)
                b.run()
                for i in range(len(shots)):
                    assert np.array_equal(b.newest(i).cpu().numpy(), want[i]), (tx4, ty, chunk, runs, i)
    b.close()
    with pytest.raises(ss.StencilError):
        b2 = ss.ShotBatch(shots, n, dev, tune=False)
        try:
            b2.set_plan(4 * 24, 8, 18)  # not a built shape
        finally:
            b2.close()


def test_batch_every_tile_shape_matches_oracle(ss, dev, oracle, monkeypatch):
    """The grouped mode's batched forward kernel at every built tile shape (1
    to 3 z-chunks, both workgroup orders), given to the groups through the
    tuner cache: every shot equals the oracle's single-shot run bit for bit."""
    monkeypatch.setenv("STENCIL_BATCH_MODE", "grouped")
    n = 12
    shots, host = _shots(ss, dev, n, seed=5)
    want = [oracle.shot_run(m, np.float32(10.0), np.float32(1e-3), n, src, w) for m, w, src in host]
    for j, (tx4, ty) in enumerate(TILES):
        for chunks in (1, 3):
            ss.clear_tuning()
            shp = {"tile_x": 4 * tx4, "tile_y": ty, "n_chunks": chunks, "xcd_remap": (j + chunks) % 2,
                   "nontemporal": 1}
            for s in shots:
                ss.set_tuned_shape(s["grid"], "wave", shp)
            b = ss.ShotBatch(shots, n, dev, tune=False)
            assert not b.ragged
            assert all(g["shape"] == shp for g in b.groups())
            b.run()
            for i in range(len(shots)):
                assert np.array_equal(b.newest(i).cpu().numpy(), want[i]), (tx4, ty, chunks, i)
            b.close()
    ss.clear_tuning()


def test_shot_without_source_stays_zero(ss, dev, mode):
    n = 12
    shots, _ = _shots(ss, dev, n, seed=5, with_source=False)
    b = ss.ShotBatch(shots, n, dev, tune=False)
    b.run()
    for i, s in enumerate(shots):
        if s["wavelet"] is None:
            assert not bool(b.newest(i).any())
        else:
            assert bool(b.newest(i).any())
    b.close()


def test_batch_rejects_bad_shots(ss, dev):
    import torch

    from stencil_solver_amd import StencilError

    grid = ss.GridDims(26, 24, 24, 10.0, 10.0, 10.0, 1e-3)  # nx % 4 != 0
    v = torch.ones((24, 24, 26), device=dev)
    with pytest.raises(StencilError):
        ss.ShotBatch([{"grid": grid, "vel2": v, "wavelet": None}], 4, dev, tune=False)
    grid = ss.GridDims(24, 24, 24, 10.0, 10.0, 10.0, 1e-3)
    v = torch.ones((24, 24, 24), device=dev)
    with pytest.raises(StencilError):  # source in the halo
        ss.ShotBatch([{"grid": grid, "vel2": v, "wavelet": np.zeros(4), "src": (2, 10, 10)}], 4, dev, tune=False)
    with pytest.raises(ValueError):
        ss.ShotBatch([{"grid": grid, "vel2": v, "wavelet": np.zeros(3), "src": (10, 10, 10)}], 4, dev, tune=False)


def test_device_noise_matches_host(ss, dev):
    """The on-device counter noise (int64 wrapping arithmetic) equals the
    numpy generator, so the synthetic models are the same on any device."""
    from stencil_solver_amd import synth

    for seed, stream in ((4, 1), (2 ** 63 + 5, 7)):
        got = synth.u01_torch(seed, stream, 200000, dev).cpu().numpy()
        assert np.array_equal(got, synth.u01(seed, stream, 200000))
