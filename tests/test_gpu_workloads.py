"""The benchmark workloads at their full sizes and step counts.

* 512^3 acoustic on the 16-layer +-5 % model: the tiled kernel equals the
  oracle (the reference CPU path's restatement, src/kernels/stencil_naive.cu:136-170)
  bit for bit after 20 and after 1000 steps, and the naive kernel after 1000.
* 512^3 heat (U[0,1] noise + hot cube, r = 0.05): tiled == the oracle's heat
  run (src/solver/heat_diffusion.cpp:103-122) bit for bit after 20 steps.
* 16-bit storage at 512^3 and 1024^3 x 1000 steps against an FP64 GPU run of
  the same inputs: block-FP16 within rel-L2 1e-3; bf16 / fp16 reported.
* 64 shots of 96^3..320^3 x 1500 steps in one batch: every shot bit-identical
  to that shot stepped alone through the single-grid forward step.
* FWI gradient at 256^3 x 2000 steps: FP64 central differences on 16 random
  voxels within 1e-4 relative of the FP64 adjoint gradient; the FP32 gradient
  against the FP64 one; the adjoint dot-product identity at that size.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
R = 4


@pytest.fixture(scope="module")
def ss(dev):
    import stencil_solver_amd

    return stencil_solver_amd


def _say(*a):
    print("[workload]", *a, file=sys.stderr, flush=True)


def _ring(step, bufs, steps):
    """steps steps over a 3-buffer ring (p_prev = bufs[0], p_cur = bufs[1]);
    returns (newest, previous)."""
    i = 0
    for _ in range(steps):
        step(bufs[(i + 1) % 3], bufs[i], bufs[(i + 2) % 3])
        i = (i + 1) % 3
    return bufs[(i + 1) % 3], bufs[i]


# ── 512^3 acoustic and heat against the oracle ────────────────────────────────
@pytest.fixture(scope="module")
def acoustic512(ss, dev, oracle):
    n = 512
    shape = (n, n, n)
    p0 = oracle.gaussian(n, sigma=3.0)
    vel2 = oracle.layered_noise_vel2(shape, n_layers=16, vmin=1500.0, vmax=4000.0, noise=0.05, seed=2)
    # the product's own host generators give the same inputs
    assert np.array_equal(ss.host_gaussian(shape, 0.5, 0.5, 0.5, 3.0), p0)
    assert np.array_equal(ss.host_layered_noise_vel2(shape, 16, 1500.0, 4000.0, 0.05, 2), vel2)
    grid = ss.GridDims.cube(n, np.float32(10.0), np.float32(1e-3))
    return p0, vel2, grid


def _gpu_run(ss, dev, kind, p0, vel2_d, grid, steps):
    import torch

    p = torch.from_numpy(p0).to(dev)
    bufs = [p.clone(), p, torch.zeros_like(p)]
    fn = ss.launch_naive if kind == "naive" else ss.launch_tiled
    out, _ = _ring(lambda c, pp, o: fn(c, pp, vel2_d, o, grid), bufs, steps)
    torch.cuda.synchronize()
    return out


def test_acoustic512_20_steps_equal_oracle(ss, dev, oracle, acoustic512):
    import torch

    p0, vel2, grid = acoustic512
    v = torch.from_numpy(vel2).to(dev)
    got = _gpu_run(ss, dev, "tiled", p0, v, grid, 20).cpu().numpy()
    want = oracle.wave_run(p0, vel2, np.float32(10.0), np.float32(1e-3), 20)
    assert np.abs(want).max() > 0.1
    assert np.array_equal(got, want)


def test_acoustic512_1000_steps_equal_oracle(ss, dev, oracle, acoustic512):
    """The configured length: 1000 steps, tiled and naive kernels and the
    oracle all bit-identical (about a minute of oracle time on 16 cores)."""
    import torch

    p0, vel2, grid = acoustic512
    v = torch.from_numpy(vel2).to(dev)
    tiled = _gpu_run(ss, dev, "tiled", p0, v, grid, 1000)
    naive = _gpu_run(ss, dev, "naive", p0, v, grid, 1000)
    assert torch.equal(tiled, naive)
    del naive
    _say("acoustic 512^3: GPU runs done, oracle 1000 steps ...")
    want = oracle.wave_run(p0, vel2, np.float32(10.0), np.float32(1e-3), 1000)
    got = tiled.cpu().numpy()
    assert np.abs(want).max() > 1e-3
    assert np.array_equal(got, want)
    # halo untouched
    mask = np.ones(got.shape, bool)
    mask[R:-R, R:-R, R:-R] = False
    assert not got[mask].any()


def test_heat512_20_steps_equal_oracle(ss, dev, oracle):
    import torch

    n = 512
    shape = (n, n, n)
    t0 = oracle.heat_noise_cube(shape, side=64, hot=10.0, seed=2)
    assert np.array_equal(ss.host_noise_cube(shape, 64, 10.0, 2), t0)
    dt_eff = np.float32(np.sqrt(np.float64(5.0)))  # r = alpha dt / h^2 = 0.05 at h = 10
    grid = ss.GridDims.cube(n, np.float32(10.0), dt_eff)
    a = torch.from_numpy(t0).to(dev)
    b = torch.zeros_like(a)
    for _ in range(20):
        ss.heat_step(a, b, 1.0, grid)
        a, b = b, a
    got = a.cpu().numpy()
    want = oracle.heat_run(t0, np.float32(10.0), dt_eff, np.float32(1.0), 20)
    assert not np.array_equal(want, t0)
    assert np.array_equal(got, want)


# ── 16-bit storage at 512^3 and 1024^3 x 1000 steps ──────────────────────────
@pytest.mark.parametrize("n", [512, 1024])
def test_16bit_storage_1000_steps_vs_fp64(ss, dev, n):
    """rel-L2 and max-abs of the final p against an FP64 GPU run of the same
    seeded inputs after 1000 steps, and the drift of the FP64 leapfrog energy.
    Block-FP16 must stay within 1e-3 (the headline bar); bf16 and fp16 are
    reported (they do not meet it) and must order bfp16 < fp16 < bf16."""
    import torch

    steps = 1000
    shape = (n, n, n)
    grid = ss.GridDims.cube(n, np.float32(10.0), np.float32(1e-3))
    p0 = torch.from_numpy(ss.host_gaussian(shape, 0.5, 0.5, 0.5, 3.0)).to(dev)
    vel2 = torch.from_numpy(ss.host_layered_noise_vel2(shape, 16, 1500.0, 4000.0, 0.05, 2)).to(dev)
    red = ss.Reducer(dev)

    def energy(prev, last, v):
        k, p = red.energy(prev, last, v, grid)
        return k + p

    v64 = vel2.double()
    b = [p0.double(), p0.double(), torch.zeros(shape, dtype=torch.float64, device=dev)]
    e_first = None
    i = 0
    for s in range(steps):
        ss.launch_naive(b[(i + 1) % 3], b[i], v64, b[(i + 2) % 3], grid)
        i = (i + 1) % 3
        if s == 0:
            e_first = energy(b[i], b[(i + 1) % 3], v64)
    ref = b[(i + 1) % 3]
    drift64 = (energy(b[i], ref, v64) - e_first) / e_first
    del b
    _say(f"16-bit {n}^3: FP64 run done (energy drift {drift64:.1e})")
    res = {}

    def report(name, last, prev, v):
        diff = (last.double() - ref)[R:-R, R:-R, R:-R]
        res[name] = dict(rel=red.rel_l2(last, ref, grid), maxabs=float(diff.abs().max()))
        del diff

    f = [p0.clone(), p0.clone(), torch.zeros_like(p0)]
    last, prev = _ring(lambda c, pp, o: ss.launch_tiled(c, pp, vel2, o, grid), f, steps)
    report("fp32", last, prev, vel2)
    del f, last, prev
    for fmt in ("bf16", "fp16", "bfp16"):
        vs = ss.pow2_scale_for(float(vel2.max())) if fmt == "fp16" else 1.0
        gv = ss.encode16(fmt, vel2, ss.Field16(fmt, shape, dev, vs))
        g = [ss.encode16(fmt, x, ss.Field16(fmt, shape, dev, 1.0)) for x in (p0, p0, torch.zeros_like(p0))]
        last, prev = _ring(lambda c, pp, o: ss.launch_tiled_16(c, pp, gv, o, grid), g, steps)
        report(fmt, ss.decode16(last), None, None)
        del g, last, prev, gv
    torch.cuda.synchronize()
    _say(f"16-bit {n}^3 x {steps}:", {k: f"rel {v['rel']:.2e} max {v['maxabs']:.1e}" for k, v in res.items()})
    assert abs(drift64) < 1e-10
    assert res["fp32"]["rel"] < 1e-3
    assert res["bfp16"]["rel"] <= 1e-3
    assert res["bfp16"]["rel"] < res["fp16"]["rel"] < res["bf16"]["rel"]


# ── 64-shot batch, 96^3 .. 320^3 x 1500 steps ─────────────────────────────────
def test_batch_64_shots_1500_steps_equal_single_grid_runs(ss, dev):
    """The batched modelling workload (sizes drawn from {96..320} with seed 4,
    smooth random models, 15 Hz Ricker at random interior cells, dt = 8e-4):
    after one batch run, every shot's p^1500 equals the same shot stepped
    alone, one global-memory forward step per launch, bit for bit."""
    import torch
    from stencil_solver_amd import synth

    n_shots, steps = 64, 1500
    rng = np.random.default_rng(4)
    sizes = rng.choice([96, 128, 160, 192, 256, 320], size=n_shots)
    w = ss.ricker(steps, 8e-4, 15.0)
    wd = torch.from_numpy(w.astype(np.float32)).to(dev)
    shots = []
    for i, n in enumerate(sizes):
        n = int(n)
        v = synth.smooth_velocity((n, n, n), 8.0, 1500.0, 4500.0, seed=1000 + i, device=dev)
        src = tuple(int(c) for c in rng.integers(4, n - 4, size=3))
        shots.append({"grid": ss.GridDims.cube(n, 10.0, 8e-4), "vel2": (v * v).contiguous(), "wavelet": w,
                      "src": src})
        del v
    b = ss.ShotBatch(shots, steps, dev, tune=True)
    b.run()
    torch.cuda.synchronize()
    _say(f"batch: {n_shots} shots x {steps} steps run; checking each shot alone")
    for i, sh in enumerate(shots):
        g = sh["grid"]
        shape = (g.nz, g.ny, g.nx)
        sz, sy, sx = sh["src"]
        lin = (sz * g.ny + sy) * g.nx + sx
        f = [torch.zeros(shape, device=dev) for _ in range(3)]
        j = 0
        for s in range(steps):
            ss.forward_step(f[(j + 1) % 3], f[j], sh["vel2"], f[(j + 2) % 3], g, src_index=lin, wavelet=wd, n=s,
                            variant="naive")
            j = (j + 1) % 3
        alone = f[(j + 1) % 3]
        got = b.newest(i)
        assert alone.abs().max().item() > 0, i
        assert torch.equal(got, alone), f"shot {i} ({g.nx}^3)"
        del f, alone
    b.close()


# ── FWI gradient at 256^3 x 2000 steps ────────────────────────────────────────
@pytest.fixture(scope="module")
def fwi256(ss, dev):
    import torch
    from stencil_solver_amd import synth

    n, steps = 256, 2000
    shape = (n, n, n)
    v0 = synth.smooth_velocity(shape, 10.0, 1800.0, 4000.0, seed=5, device=dev)
    sphere = synth.sphere_mask(shape, (n // 2, n // 2, n // 2), 20.0, device=dev)
    v_true = torch.where(sphere, v0 * 1.1, v0)
    w = ss.ricker(steps, 8e-4, 12.0)
    return dict(n=n, steps=steps, m0=(v0 * v0).contiguous(), m_true=(v_true * v_true).contiguous(), w=w,
                src=(12, n // 2, n // 2), rec_z=8)


def _fwi_problem(ss, dev, c, dtype):
    import torch

    T = np.float32 if dtype == torch.float32 else np.float64
    grid = ss.GridDims.cube(c["n"], T(10.0), T(8e-4))
    return ss.FwiProblem(grid, c["steps"], c["src"], c["rec_z"], c["w"], dev, dtype)


@pytest.fixture(scope="module")
def fwi256_fp64(ss, dev, fwi256):
    import torch

    c = fwi256
    pr = _fwi_problem(ss, dev, c, torch.float64)
    m0 = c["m0"].double()
    d_obs = pr.traces(c["m_true"].double())
    J0 = pr.misfit(m0, d_obs).item()
    g = pr.gradient(m0)
    return pr, m0, d_obs, J0, g


def test_fwi256_gradient_matches_finite_differences(ss, dev, fwi256, fwi256_fp64):
    """Central differences of the FP64 misfit (FP64 forward runs) on 16
    random interior voxels, eps = 1e-4 m, against the FP64 adjoint gradient:
    within 1e-4 relative.  Voxels are drawn among those whose gradient is at
    least 1e-4 of the largest (the gradient peaks sharply at the source),
    where a relative error is meaningful: the difference quotient carries a
    rounding error of about 1e-14 J / eps, and its truncation error is
    O((eps/m)^2)."""
    import torch

    pr, m0, d_obs, J0, g = fwi256_fp64
    n = fwi256["n"]
    assert J0 > 0
    inner = torch.zeros((n, n, n), dtype=torch.bool, device=dev)
    inner[R:-R, R:-R, R:-R] = True
    gmax = g.abs().max().item()
    cand = torch.nonzero(inner & (g.abs() > 1e-4 * gmax)).cpu().numpy()
    _say(f"FWI 256^3: {len(cand)} interior voxels with |g| > 1e-4 max|g|")
    assert len(cand) > 1000
    rng = np.random.default_rng(11)
    picks = [tuple(int(v) for v in cand[k]) for k in rng.choice(len(cand), 16, replace=False)]
    errs = []
    for idx in picks:
        eps = 1e-4 * m0[idx].item()
        mp = m0.clone()
        mp[idx] += eps
        Jp = pr.misfit(mp, d_obs).item()
        mp[idx] -= 2 * eps
        Jm = pr.misfit(mp, d_obs).item()
        fd = (Jp - Jm) / (2 * eps)
        ga = g[idx].item()
        errs.append(abs(fd - ga) / abs(ga))
        _say(f"  voxel {idx}: |g|/max {abs(ga) / gmax:.1e}, FD rel err {errs[-1]:.1e}")
    _say("FWI 256^3 x 2000 FD relative errors:", ["%.1e" % e for e in errs])
    assert max(errs) < 1e-4


def test_fwi256_fp32_gradient_vs_fp64(ss, dev, fwi256, fwi256_fp64):
    """The FP32 pipeline (tiled kernels, sqrt(n) checkpoints) against the FP64
    one on the same models: relative L2 of the gradient and of the misfit."""
    import torch

    _, _, _, J64, g64 = fwi256_fp64
    c = fwi256
    pr = _fwi_problem(ss, dev, c, torch.float32)
    d_obs = pr.traces(c["m_true"])
    J32 = pr.misfit(c["m0"], d_obs).item()
    g32 = pr.gradient(c["m0"])
    rel = (torch.linalg.vector_norm(g32.double() - g64) / torch.linalg.vector_norm(g64)).item()
    _say(f"FWI 256^3 x 2000: FP32 gradient rel-L2 vs FP64 {rel:.2e}, misfit {J32:.6e} vs {J64:.6e}")
    assert rel < 1e-2
    assert J32 == pytest.approx(J64, rel=1e-3)


@pytest.mark.parametrize("dtype_name,tol", [("float64", 1e-9), ("float32", 1e-3)])
def test_fwi256_dot_product_identity(ss, dev, fwi256, dtype_name, tol):
    """<P b, r> = <b, P^T r> at 256^3 x 2000 steps: P maps the source wavelet
    to the receiver traces, P^T random receiver data (zero on the halo) to
    the adjoint field at the source cell."""
    import torch

    dtype = getattr(torch, dtype_name)
    c = fwi256
    pr = _fwi_problem(ss, dev, c, dtype)
    m0 = c["m0"].to(dtype)
    d = pr.traces(m0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    r = torch.randn(tuple(d.shape), generator=gen, device=dev, dtype=dtype)
    r[:, :R, :] = 0
    r[:, -R:, :] = 0
    r[:, :, :R] = 0
    r[:, :, -R:] = 0
    pr.misfit(m0, d - r)  # residual = d - (d - r) = r (up to one rounding)
    _, lam_src = pr.gradient(m0, src_trace=True)
    resid = d - (d - r)
    lhs = float((d.double() * resid.double()).sum())
    rhs = float((pr.wavelet.double() * lam_src.double()).sum())
    _say(f"FWI 256^3 x 2000 dot test ({dtype_name}): <Pb,r> = {lhs:.12e}, <b,P^T r> = {rhs:.12e}, "
         f"rel {abs(lhs - rhs) / abs(lhs):.1e}")
    assert rhs == pytest.approx(lhs, rel=tol)
