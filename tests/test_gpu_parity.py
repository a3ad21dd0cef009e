"""Parity of the gfx950 kernels with the oracle and the reference's golden vectors.

Bars (tolerances written in each test):
* naive and tiled kernels (FP32 and FP64): bit-identical to the oracle / to the
  reference CPU path built without FMA contraction, and within 1e-5 relative
  L2 of the reference's own Release (FMA-contracted) build.
* the opt-in FMA-fused tiled form: relative L2 <= 1e-5 after 5 steps, <= 1e-4
  after 200 (its rounding drifts from the reference association).
* at the full 512^3 / 1024^3 sizes (too large for the oracle): size-independent
  properties -- exact-mode tiled == naive bit for bit, fused within 1e-5 of
  naive, halos untouched, mirror symmetry of a centred pulse.
"""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu
R = 4


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def to_dev(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def ring_run(ss, variant, p0, vel2, grid, n_steps, dev, **kw):
    """n_steps over a 3-buffer ring as WavePropSolver::step_impl does; returns cur."""
    import torch

    bufs = [to_dev(p0, dev), to_dev(p0, dev), torch.zeros(p0.shape, dtype=to_dev(p0, dev).dtype, device=dev)]
    v = to_dev(vel2, dev)
    i = 0
    for _ in range(n_steps):
        if variant == "tiled_ex":
            ss.launch_tiled_ex(bufs[(i + 1) % 3], bufs[i], v, bufs[(i + 2) % 3], grid, **kw)
        else:
            ss.launch(variant, bufs[(i + 1) % 3], bufs[i], v, bufs[(i + 2) % 3], grid)
        i = (i + 1) % 3
    torch.cuda.synchronize()
    return bufs[(i + 1) % 3].cpu().numpy()


@pytest.fixture(scope="module")
def ss(dev):
    import stencil_solver_amd

    return stencil_solver_amd


# ── golden vectors of the reference CPU path ──────────────────────────────────
def test_naive_step_matches_reference_golden(ss, dev):
    gd = load_golden("kernel_24.npz")
    g = gd["init"]
    vel2 = np.full_like(g, np.float32(1500.0) * np.float32(1500.0))
    grid = ss.GridDims.cube(24, np.float32(10.0), np.float32(1e-3))
    import torch

    nxt = torch.zeros(g.shape, device=dev)
    ss.launch_naive(to_dev(g, dev), to_dev(g, dev), to_dev(vel2, dev), nxt, grid)
    out = nxt.cpu().numpy()
    assert np.array_equal(out, gd["next_nofma"])  # bit-exact
    t = torch.zeros(g.shape, device=dev)
    ss.launch_tiled(to_dev(g, dev), to_dev(g, dev), to_dev(vel2, dev), t, grid)
    tt = t.cpu().numpy()
    rel = np.linalg.norm(tt - gd["next_nofma"]) / np.linalg.norm(gd["next_nofma"])
    assert rel <= 1e-5 and np.max(np.abs(tt - out)) < 1e-5


@pytest.mark.parametrize("variant", ["naive", "tiled", "multigpu"])
@pytest.mark.parametrize("steps", [1, 20])
def test_wave_solver_matches_reference_golden(ss, dev, variant, steps):
    gd = load_golden("solver_wave_32.npz")
    s = ss.make_solver(ss.Config(nx=32, ny=32, nz=32, velocity=1500.0, pulse_sigma=4.0, variant=variant))
    for _ in range(steps):
        s.step()
    out = s.download()
    ref = gd[f"naive{steps}_nofma"]
    assert np.array_equal(out, ref)
    nat = gd[f"naive{steps}_native"].astype(np.float64)
    assert np.linalg.norm(out - nat) / np.linalg.norm(nat) <= 1e-5


def test_run_with_warmup_matches_golden(ss, dev):
    gd = load_golden("solver_wave_32.npz")
    s = ss.make_solver(ss.Config(nx=32, ny=32, nz=32, velocity=1500.0, pulse_sigma=4.0, variant="naive"))
    r = s.run(5, 2)
    assert r.label == "wave/naive" and r.time_ms > 0 and r.grid_size == 32
    assert np.array_equal(s.download(), gd["run5_warm2_nofma"])


@pytest.mark.parametrize("variant", ["naive", "tiled"])
def test_heat_solver_matches_reference_golden(ss, dev, variant):
    gd = load_golden("solver_heat_48.npz")
    s = ss.make_solver(ss.Config(nx=48, ny=48, nz=48, dx=1.0, dy=1.0, dz=1.0, velocity=500.0,
                                 diffusivity=float(np.float32(1e-3)), pulse_sigma=5.0, physics="heat",
                                 variant=variant))
    assert s.name() == f"heat/{variant}"
    for _ in range(20):
        s.step()
    out = s.download()
    assert np.array_equal(out, gd["heat20_nofma"])
    nat = gd["heat20_native"].astype(np.float64)
    assert np.linalg.norm(out - nat) / np.linalg.norm(nat) <= 1e-5


@pytest.mark.parametrize("variant", ["naive", "tiled"])
def test_acoustic128_matches_reference_golden(ss, dev, variant, golden_meta):
    """128^3, h=10, v=2000, dt=1e-3, sigma=3, 200 steps."""
    gd = load_golden("acoustic_128.npz")
    p0 = ss.host_gaussian((128, 128, 128), 0.5, 0.5, 0.5, 3.0)
    assert sha(p0) == golden_meta["acoustic128_nofma"]["sha_init"]
    vel2 = np.full_like(p0, np.float32(2000.0) * np.float32(2000.0))
    grid = ss.GridDims.cube(128, np.float32(10.0), np.float32(1e-3))
    out = ring_run(ss, variant, p0, vel2, grid, 200, dev)
    assert sha(out) == golden_meta["acoustic128_nofma"]["sha_final"]
    mid = gd["mid_z_nofma"].astype(np.float64)
    assert np.linalg.norm(out[64] - mid) / np.linalg.norm(mid) <= 1e-5
    nat = gd["mid_z_native"].astype(np.float64)
    assert np.linalg.norm(out[64] - nat) / np.linalg.norm(nat) <= 1e-5


def test_acoustic128_fp64_matches_reference_golden(ss, dev, golden_meta):
    from oracle import Oracle

    p0 = Oracle().gaussian(128, sigma=3.0, dtype=np.float64)
    vel2 = np.full_like(p0, 2000.0 * 2000.0)
    grid = ss.GridDims.cube(128, 10.0, 1e-3)
    for variant in ("naive", "tiled"):
        out = ring_run(ss, variant, p0, vel2, grid, 200, dev)
        assert sha(out) == golden_meta["acoustic128_f64"]["sha_final"]


# ── against the oracle on ragged shapes, seeded inputs ───────────────────────
SHAPES = [(36, 52, 40), (40, 44, 72), (24, 24, 24), (33, 29, 136), (20, 61, 12), (28, 30, 30)]


def _inputs(oracle, shape, seed):
    p0 = oracle.gaussian(0, 0.45, 0.55, 0.5, 3.0, shape=shape)
    vel2 = oracle.layered_noise_vel2(shape, n_layers=5, vmin=1500.0, vmax=4000.0, noise=0.05, seed=seed)
    return p0, vel2


@pytest.mark.parametrize("shape", SHAPES)
def test_naive_bit_exact_vs_oracle(ss, dev, oracle, shape):
    p0, vel2 = _inputs(oracle, shape, 11)
    grid = ss.GridDims(shape[2], shape[1], shape[0], np.float32(10.0), np.float32(12.0), np.float32(8.0),
                       np.float32(8e-4))
    ref = oracle.wave_run(p0, vel2, grid.dx, grid.dt, 6, grid.dy, grid.dz)
    assert np.array_equal(ring_run(ss, "naive", p0, vel2, grid, 6, dev), ref)


@pytest.mark.parametrize("shape", [s for s in SHAPES if s[2] % 4 == 0])
@pytest.mark.parametrize("tx4,ty,n_chunks,xcd,nt", [(16, 16, 0, -1, -1), (16, 32, 1, 0, 1), (32, 8, 3, 1, 0),
                                                   (32, 16, 5, 0, 1), (64, 4, 2, 1, 0), (64, 8, 0, -1, 1),
                                                   (64, 12, 3, 0, 0), (64, 16, 2, 1, 1)])
def test_tiled_exact_bit_exact_vs_oracle(ss, dev, oracle, shape, tx4, ty, n_chunks, xcd, nt):
    p0, vel2 = _inputs(oracle, shape, 12)
    grid = ss.GridDims(shape[2], shape[1], shape[0], np.float32(10.0), np.float32(12.0), np.float32(8.0),
                       np.float32(8e-4))
    ref = oracle.wave_run(p0, vel2, grid.dx, grid.dt, 5, grid.dy, grid.dz)
    kw = dict(tx4=tx4, ty=ty, n_chunks=n_chunks, xcd_remap=xcd, nontemporal=nt)
    out = ring_run(ss, "tiled_ex", p0, vel2, grid, 5, dev, exact=True, **kw)
    assert np.array_equal(out, ref)
    fused = ring_run(ss, "tiled_ex", p0, vel2, grid, 5, dev, exact=False, **kw)
    assert oracle.rel_l2_interior(fused, ref) <= 1e-5


@pytest.mark.parametrize("shape", SHAPES)
def test_tiled_default_vs_oracle(ss, dev, oracle, shape):
    """launch_tiled takes any shape (nx % 4 != 0 runs the coalesced kernel)."""
    p0, vel2 = _inputs(oracle, shape, 13)
    grid = ss.GridDims(shape[2], shape[1], shape[0], np.float32(10.0), np.float32(10.0), np.float32(10.0),
                       np.float32(1e-3))
    ref = oracle.wave_run(p0, vel2, grid.dx, grid.dt, 8)
    assert np.array_equal(ring_run(ss, "tiled", p0, vel2, grid, 8, dev), ref)


@pytest.mark.parametrize("exact", [True, False])
def test_alternating_sweep_bit_identical(ss, dev, oracle, exact):
    """Chained steps sweep z alternately down and up (each step starts on the
    planes the previous one wrote last); steps whose input is a fresh copy
    always sweep upwards.  Both give the same bits."""
    import torch

    shape = (44, 36, 72)
    p0, vel2 = _inputs(oracle, shape, 21)
    grid = ss.GridDims(shape[2], shape[1], shape[0], np.float32(10.0), np.float32(12.0), np.float32(8.0),
                       np.float32(8e-4))
    kw = dict(tx4=32, ty=8, n_chunks=3, exact=exact)
    chained = ring_run(ss, "tiled_ex", p0, vel2, grid, 9, dev, **kw)
    prev, cur, v = to_dev(p0, dev), to_dev(p0, dev), to_dev(vel2, dev)
    for _ in range(9):
        nxt = torch.zeros_like(cur)
        ss.launch_tiled_ex(cur.clone(), prev.clone(), v, nxt, grid, **kw)
        prev, cur = cur, nxt
    torch.cuda.synchronize()
    assert np.array_equal(chained, cur.cpu().numpy())
    if exact:
        assert np.array_equal(chained, oracle.wave_run(p0, vel2, grid.dx, grid.dt, 9, grid.dy, grid.dz))


def test_layered_model_128_50_steps(ss, dev, oracle):
    """The 512^3 workload's velocity model at 128^3: 16 layers, +-5 % noise."""
    shape = (128, 128, 128)
    p0 = ss.host_gaussian(shape, 0.5, 0.5, 0.5, 3.0)
    vel2 = ss.host_layered_noise_vel2(shape, 16, 1500.0, 4000.0, 0.05, 2)
    assert np.array_equal(vel2, oracle.layered_noise_vel2(shape, 16, 1500.0, 4000.0, 0.05, 2))
    grid = ss.GridDims.cube(128, np.float32(10.0), np.float32(1e-3))
    ref = oracle.wave_run(p0, vel2, grid.dx, grid.dt, 50)
    assert np.array_equal(ring_run(ss, "naive", p0, vel2, grid, 50, dev), ref)
    assert np.array_equal(ring_run(ss, "tiled", p0, vel2, grid, 50, dev), ref)


def test_fused_form_drift_bound(ss, dev, golden_meta):
    """The opt-in FMA form after 200 steps of the 128^3 workload: <= 1e-4."""
    gd = load_golden("acoustic_128.npz")
    p0 = ss.host_gaussian((128, 128, 128), 0.5, 0.5, 0.5, 3.0)
    vel2 = np.full_like(p0, np.float32(2000.0) * np.float32(2000.0))
    grid = ss.GridDims.cube(128, np.float32(10.0), np.float32(1e-3))
    out = ring_run(ss, "tiled_ex", p0, vel2, grid, 200, dev, exact=False)
    mid = gd["mid_z_nofma"].astype(np.float64)
    rel = np.linalg.norm(out[64] - mid) / np.linalg.norm(mid)
    print(f"fused-form rel-L2 after 200 steps: {rel:.3e}")
    assert rel <= 1e-4


@pytest.mark.parametrize("exact", [None, True, False])
def test_heat_tiled_vs_oracle(ss, dev, oracle, exact):
    import torch

    shape = (40, 36, 48)
    t0 = oracle.heat_noise_cube(shape, side=8, hot=10.0, seed=2)
    dte = oracle.heat_dt_eff(np.float32(5.0))
    grid = ss.GridDims(shape[2], shape[1], shape[0], np.float32(10.0), np.float32(10.0), np.float32(10.0), dte)
    ref = oracle.heat_run(t0, np.float32(10.0), dte, np.float32(1.0), 10)
    a, b = to_dev(t0, dev), torch.zeros(shape, device=dev)
    for _ in range(10):
        ss.heat_step(a, b, 1.0, grid, exact=exact)
        a, b = b, a
    out = a.cpu().numpy()
    if exact is False:
        assert oracle.rel_l2_interior(out, ref) <= 1e-5
    else:
        assert np.array_equal(out, ref)
    a, b = to_dev(t0, dev), torch.zeros(shape, device=dev)
    for _ in range(10):
        ss.heat_step(a, b, 1.0, grid, variant="naive")
        a, b = b, a
    assert np.array_equal(a.cpu().numpy(), ref)


@pytest.mark.parametrize("variant,exact", [("naive", None), ("tiled", True), ("tiled", False)])
def test_heat_anisotropic_spacing_vs_oracle(ss, dev, oracle, variant, exact):
    """Heat steps on a grid with three different spacings (x 10, y 12, z 8):
    each axis keeps its own 1/h^2, so a swapped or shared axis weight fails.
    The reference association is bit-exact, the fused form within 1e-5
    relative L2."""
    import torch

    shape = (36, 34, 44)
    t0 = oracle.heat_noise_cube(shape, side=8, hot=10.0, seed=6)
    dte = oracle.heat_dt_eff(np.float32(5.0))
    grid = ss.GridDims(shape[2], shape[1], shape[0], np.float32(10.0), np.float32(12.0), np.float32(8.0), dte)
    ref = oracle.heat_run(t0, np.float32(10.0), dte, np.float32(1.0), 6, dy=np.float32(12.0), dz=np.float32(8.0))
    iso = oracle.heat_run(t0, np.float32(10.0), dte, np.float32(1.0), 6)
    assert oracle.rel_l2_interior(iso, ref) > 1e-3  # the spacings matter at this size
    a, b = to_dev(t0, dev), torch.zeros(shape, device=dev)
    for _ in range(6):
        ss.heat_step(a, b, 1.0, grid, variant=variant, exact=exact)
        a, b = b, a
    out = a.cpu().numpy()
    if exact is False:
        assert oracle.rel_l2_interior(out, ref) <= 1e-5
    else:
        assert np.array_equal(out, ref)


@pytest.mark.parametrize("variant", ["naive", "tiled"])
def test_heat_fp64_is_the_wave_step_with_alpha(ss, dev, oracle, variant):
    """The FP64 heat step equals the reference's definition of it, the wave
    step with vel2 = alpha and p_prev = p_cur (heat_diffusion.cpp:103-122),
    as the oracle's FP64 wave step computes it, bit for bit over 4 steps."""
    import torch

    shape = (30, 27, 41)
    rng = np.random.default_rng(8)
    t = np.zeros(shape)
    t[4:-4, 4:-4, 4:-4] = rng.random((22, 19, 33))
    alpha, dte = 1.5, float(np.sqrt(0.05 * 100.0 / 1.5))
    grid = ss.GridDims(shape[2], shape[1], shape[0], 10.0, 10.0, 10.0, dte)
    ref = t.copy()
    for _ in range(4):
        ref = oracle.wave_step(ref, ref, np.full(shape, alpha), np.zeros(shape), 10.0, dte)
    a, b = to_dev(t, dev), torch.zeros(shape, dtype=torch.float64, device=dev)
    for _ in range(4):
        ss.heat_step(a, b, alpha, grid, variant=variant)
        a, b = b, a
    assert np.array_equal(a.cpu().numpy(), ref) and np.abs(ref - t).max() > 0


def test_heat_every_tile_shape_vs_oracle(ss, dev, oracle):
    """The heat step at every built tile shape (1 to 3 z-chunks, both
    workgroup orders and streaming modes) against the oracle: the reference
    association bit-exact, the fused form within 1e-5 relative L2."""
    import torch

    shape = (33, 29, 56)
    t0 = oracle.heat_noise_cube(shape, side=8, hot=10.0, seed=4)
    dte = oracle.heat_dt_eff(np.float32(5.0))
    grid = ss.GridDims(shape[2], shape[1], shape[0], np.float32(10.0), np.float32(10.0), np.float32(10.0), dte)
    ref = oracle.heat_run(t0, np.float32(10.0), dte, np.float32(1.0), 3)
    tiles = [(16, 16), (16, 32), (32, 8), (32, 16), (64, 4), (64, 8), (64, 12), (64, 16)]
    for j, (tx4, ty) in enumerate(tiles):
        for chunks in (1, 2, 3):
            kw = dict(tx4=tx4, ty=ty, n_chunks=chunks, xcd_remap=(j + chunks) % 2, nontemporal=(j + chunks + 1) % 2)
            for exact in (True, False):
                a, b = to_dev(t0, dev), torch.zeros(shape, device=dev)
                for _ in range(3):
                    ss.heat_step(a, b, 1.0, grid, exact=exact, **kw)
                    a, b = b, a
                out = a.cpu().numpy()
                if exact:
                    assert np.array_equal(out, ref), kw
                else:
                    assert oracle.rel_l2_interior(out, ref) <= 1e-5, kw


def test_tiled_steps_replay_in_a_graph(ss, dev, oracle):
    """Ten tiled steps captured in a HIP graph (torch.cuda.CUDAGraph) and
    replayed three times, with eager steps and synchronisations between the
    replays: every replay equals the oracle's ten steps bit for bit (the
    launchers record no memset and allocate nothing)."""
    import torch

    shape = (40, 36, 48)
    p0, vel2 = _inputs(oracle, shape, 5)
    grid = ss.GridDims(shape[2], shape[1], shape[0], np.float32(10.0), np.float32(12.0), np.float32(8.0),
                       np.float32(8e-4))
    want = oracle.wave_run(p0, vel2, grid.dx, grid.dt, 10, grid.dy, grid.dz)
    v = to_dev(vel2, dev)
    init = [to_dev(p0, dev), to_dev(p0, dev), torch.zeros(shape, device=dev)]

    def steps(buf):  # p^{n+1} into buf[(n+2) % 3]; after 10 steps the newest is buf[2]
        for i in range(10):
            ss.launch_tiled(buf[(i + 1) % 3], buf[i % 3], v, buf[(i + 2) % 3], grid)

    bufs = [t.clone() for t in init]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):  # warm-up outside the capture
        steps(bufs)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        steps(bufs)
    other = [t.clone() for t in init]
    for _ in range(3):
        for k in range(3):
            bufs[k].copy_(init[k])
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(bufs[2].cpu().numpy(), want)
        for k in range(3):  # eager steps between the replays
            other[k].copy_(init[k])
        steps(other)
        torch.cuda.synchronize()
        assert np.array_equal(other[2].cpu().numpy(), want)


@pytest.mark.parametrize("variant", ["naive", "tiled"])
def test_halo_never_written(ss, dev, oracle, variant):
    import torch

    shape = (36, 44, 40)
    p0, vel2 = _inputs(oracle, shape, 14)
    grid = ss.GridDims(40, 44, 36, np.float32(10.0), np.float32(10.0), np.float32(10.0), np.float32(1e-3))
    nxt = torch.full(shape, 7.5, device=dev)
    ss.launch(variant, to_dev(p0, dev), to_dev(p0, dev), to_dev(vel2, dev), nxt, grid)
    out = nxt.cpu().numpy()
    mask = np.ones(shape, bool)
    mask[R:-R, R:-R, R:-R] = False
    assert np.all(out[mask] == np.float32(7.5))
    ref = np.full(shape, np.float32(7.5))
    oracle.wave_step(p0, p0, vel2, ref, grid.dx, grid.dt)
    assert oracle.rel_l2_interior(out, ref) <= 1e-6


def test_runs_on_caller_stream(ss, dev, oracle):
    import torch

    shape = (32, 32, 32)
    p0, vel2 = _inputs(oracle, shape, 15)
    grid = ss.GridDims.cube(32, np.float32(10.0), np.float32(1e-3))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        a, v = to_dev(p0, dev), to_dev(vel2, dev)
        n = torch.zeros_like(a)
        ss.launch_tiled(a, a, v, n, grid, stream=s)
    s.synchronize()
    ref = np.zeros_like(p0)
    oracle.wave_step(p0, p0, vel2, ref, grid.dx, grid.dt)
    assert oracle.rel_l2_interior(n.cpu().numpy(), ref) <= 1e-6


def test_argument_errors(ss, dev):
    import torch

    grid = ss.GridDims.cube(16, 10.0, 1e-3)
    a = torch.zeros((16, 16, 16), device=dev)
    with pytest.raises(TypeError):
        ss.launch_naive(a.cpu(), a, a, torch.zeros_like(a), grid)
    with pytest.raises(ValueError):
        ss.launch_naive(a, a, a, torch.zeros((16, 16, 15), device=dev), ss.GridDims.cube(16))
    with pytest.raises(ss.StencilError):
        ss.launch_naive(a, a, a, a, grid)  # p_next aliases p_cur
    with pytest.raises(TypeError):
        ss.launch_tiled(a.double(), a, a, a, grid)


# ── full-size properties ─────────────────────────────────────────────────────
@pytest.mark.parametrize("n", [512, 1024])
def test_fullsize_tiled_equals_naive(ss, dev, n):
    """At 512^3 and 1024^3: exact-mode tiled == naive bit for bit over 3 steps,
    fused tiled within 1e-5 rel-L2 of naive, halo untouched, x-mirror symmetry
    of a centred pulse in a z-layered medium."""
    import torch

    shape = (n, n, n)
    p0 = to_dev(ss.host_gaussian(shape, 0.5, 0.5, 0.5, 3.0), dev)
    vel2 = to_dev(ss.host_layered_noise_vel2(shape, 16, 1500.0, 4000.0, 0.0, 2), dev)
    grid = ss.GridDims.cube(n, np.float32(10.0), np.float32(1e-3))

    def run(kind, steps):
        bufs = [p0.clone(), p0.clone(), torch.full(shape, 3.25, device=dev)]
        i = 0
        for _ in range(steps):
            args = (bufs[(i + 1) % 3], bufs[i], vel2, bufs[(i + 2) % 3], grid)
            if kind == "naive":
                ss.launch_naive(*args)
            else:
                ss.launch_tiled_ex(*args, exact=(kind == "exact"))
            i = (i + 1) % 3
        torch.cuda.synchronize()
        return bufs[(i + 1) % 3], bufs

    a, _ = run("naive", 3)
    b, _ = run("exact", 3)
    assert torch.equal(a, b)
    del b
    c, _ = run("fused", 3)
    num = torch.linalg.vector_norm((c - a).double())
    assert (num / torch.linalg.vector_norm(a.double())).item() <= 1e-5
    # the buffer that started at 3.25 has been written once (step 3): halo intact
    _, bufs = run("fused", 3)
    w = bufs[2]
    assert torch.all(w[:R] == 3.25) and torch.all(w[:, :, -R:] == 3.25) and torch.all(w[:, :R] == 3.25)
    assert torch.equal(a, torch.flip(a, dims=[2]))  # constant-in-x medium, centred pulse


def test_fullsize_heat_tiled_equals_naive(ss, dev):
    """The 512^3 heat workload (U[0,1] noise + a hot cube of side 64, r = 0.05):
    tiled (exact association, default and tuned shape, chained steps sweeping
    z both ways) == naive bit for bit over 4 steps; the fused form within 1e-6
    rel-L2.  (No maximum principle: the 8th-order weights c2, c4 are negative,
    so the cube's edges overshoot by a few per cent.)"""
    import torch

    n = 512
    shape = (n, n, n)
    t0 = to_dev(ss.host_noise_cube(shape, 64, 10.0, 2), dev)
    grid = ss.GridDims.cube(n, np.float32(10.0), np.float32(np.sqrt(np.float64(5.0))))

    def run(steps, **kw):
        a, b = t0.clone(), torch.zeros_like(t0)
        for _ in range(steps):
            ss.heat_step(a, b, 1.0, grid, **kw)
            a, b = b, a
        torch.cuda.synchronize()
        return a

    ref = run(4, variant="naive")
    assert torch.equal(run(4, exact=True), ref)
    ss.tune_heat(t0.clone(), torch.zeros_like(t0), grid, reps=1)
    assert torch.equal(run(4, exact=True), ref)
    fused = run(4, exact=False)
    rel = torch.linalg.vector_norm((fused - ref).double()) / torch.linalg.vector_norm(ref.double())
    assert rel.item() <= 1e-6
    assert not torch.equal(ref, t0) and 10.0 < ref.max().item() < 10.5  # heat spread, bounded overshoot
    ss.clear_tuning()


# ── autotuner ─────────────────────────────────────────────────────────────────
def test_autotuner_picks_and_caches_shape(ss, dev, oracle, tmp_path):
    import torch

    shape = (40, 44, 136)
    p0, vel2 = _inputs(oracle, shape, 21)
    grid = ss.GridDims(136, 44, 40, np.float32(10.0), np.float32(10.0), np.float32(10.0), np.float32(1e-3))
    ss.clear_tuning()
    before, hit = ss.tuned_shape(grid)
    assert not hit
    scratch = torch.zeros(shape, device=dev)
    csv = tmp_path / "sweep.csv"
    best = ss.tune_tiled(to_dev(p0, dev), to_dev(p0, dev), to_dev(vel2, dev), scratch, grid, reps=2,
                         csv_path=csv)
    after, hit = ss.tuned_shape(grid)
    assert hit and after == best
    lines = csv.read_text().strip().splitlines()
    assert lines[0].startswith("nx,ny,nz,tile_x,tile_y,n_chunks,nontemporal,xcd_remap,naive_ms,tiled_ms,speedup")
    assert len(lines) > 20
    # the tuned launch is still bit-identical to the oracle
    ref = oracle.wave_run(p0, vel2, grid.dx, grid.dt, 6)
    assert np.array_equal(ring_run(ss, "tiled", p0, vel2, grid, 6, dev), ref)
    # heat tuner on the same dims
    t0 = oracle.heat_noise_cube(shape, side=8, hot=10.0, seed=2)
    hb = ss.tune_heat(to_dev(t0, dev), torch.zeros(shape, device=dev), grid, reps=2)
    assert ss.tuned_shape(grid, "heat") == (hb, True)
    ss.clear_tuning()
    assert not ss.tuned_shape(grid)[1]
