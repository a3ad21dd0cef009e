"""The solver API (reference include/stencil/solver.hpp:13-24) through the C-ABI
handles: names, dims, metrics, FP64 build, seeded workloads."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ss(dev):
    import stencil_solver_amd

    return stencil_solver_amd


@pytest.mark.parametrize("physics,variant,name", [("wave", "naive", "wave/naive"), ("wave", "tiled", "wave/tiled"),
                                                  ("wave", "multigpu", "wave/multigpu"),
                                                  ("heat", "naive", "heat/naive"), ("heat", "tiled", "heat/tiled")])
def test_names_and_dims(ss, physics, variant, name):
    s = ss.make_solver(ss.Config(nx=32, ny=32, nz=32, physics=physics, variant=variant, diffusivity=1.0))
    assert s.name() == name
    d = s.dims()
    assert (d.nx, d.ny, d.nz) == (32, 32, 32)
    assert d.dt == pytest.approx(float(np.float32(0.45) * np.float32(10.0) / np.float32(1500.0)))


def test_run_metrics(ss):
    s = ss.make_solver(ss.Config(nx=64, ny=64, nz=64, variant="tiled"))
    r = s.run(10, 2)
    assert r.label == "wave/tiled" and r.grid_size == 64
    assert r.time_ms > 0 and r.bandwidth_gbs > 0 and r.gflops > 0
    assert r.model_bandwidth_gbs > r.bandwidth_gbs  # (6R+4) model > compulsory 4 values


def test_fp64_solver_matches_oracle(ss, oracle):
    s = ss.make_solver(ss.Config(nx=40, ny=40, nz=40, velocity=1500.0, pulse_sigma=4.0, fp64=True))
    for _ in range(10):
        s.step()
    out = s.download()
    assert out.dtype == np.float64
    p0 = oracle.gaussian(40, sigma=4.0, dtype=np.float64)
    dt = float(0.45 * 10.0 / 1500.0)
    ref = oracle.wave_run(p0, np.full_like(p0, 1500.0 * 1500.0), 10.0, dt, 10)
    assert np.array_equal(out, ref)


def test_explicit_dt_and_layered_model(ss, oracle):
    cfg = ss.Config(nx=48, ny=48, nz=48, dt=float(np.float32(1e-3)), vmodel="layered-noise", n_layers=4,
                    seed=9, pulse_sigma=3.0, variant="naive")
    s = ss.make_solver(cfg)
    assert s.dims().dt == float(np.float32(1e-3))
    for _ in range(12):
        s.step()
    p0 = oracle.gaussian(48, sigma=3.0)
    vel2 = oracle.layered_noise_vel2(p0.shape, n_layers=4, vmin=1500.0, vmax=4000.0, noise=0.05, seed=9)
    ref = oracle.wave_run(p0, vel2, np.float32(10.0), np.float32(1e-3), 12)
    assert np.array_equal(s.download(), ref)


def test_heat_noise_cube_workload(ss, oracle):
    """Heat with the noise + hot-cube IC at r = alpha dt / h^2 = 0.05 (h = 10, dt = 5)."""
    dt = float(np.float32(5.0))
    cfg = ss.Config(nx=40, ny=40, nz=40, physics="heat", ic="noise-cube", cube_side=8, cube_value=10.0, seed=2,
                    diffusivity=1.0, dt=dt, variant="naive")
    s = ss.make_solver(cfg)
    for _ in range(15):
        s.step()
    t0 = oracle.heat_noise_cube((40, 40, 40), side=8, hot=10.0, seed=2)
    ref = oracle.heat_run(t0, np.float32(10.0), oracle.heat_dt_eff(np.float32(dt)), np.float32(1.0), 15)
    assert np.array_equal(s.download(), ref)
