"""Adjoint-state FWI on the GPU: forward traces, misfit and gradient against
the oracle (bit-identical, any checkpoint interval), central finite
differences (FP64, 16 voxels), the adjoint dot-product identity, and the
torch.autograd binding."""
from __future__ import annotations

import numpy as np
import pytest

from test_oracle_fwi import PROB, _problem, smooth_model

pytestmark = pytest.mark.gpu
R = 4


@pytest.fixture(scope="module")
def ss(dev):
    import stencil_solver_amd

    return stencil_solver_amd


def _torch_dtype(dtype):
    import torch

    return torch.float32 if dtype == np.float32 else torch.float64


def _grid(ss, shape, h, dt, dtype):
    nz, ny, nx = shape
    return ss.GridDims(nx, ny, nz, dtype(h), dtype(h), dtype(h), dtype(dt))


def _setup(ss, dev, dtype, ckpt_every=0, prob=PROB, seed=1, variant="tiled"):
    import torch

    m, w = _problem(dtype, seed)
    grid = _grid(ss, prob["shape"], prob["h"], prob["dt"], dtype)
    pr = ss.FwiProblem(grid, prob["n"], prob["src"], prob["rec_z"], w, dev, _torch_dtype(dtype), ckpt_every,
                       variant)
    return m, w, grid, pr, torch.from_numpy(m).to(dev)


@pytest.mark.parametrize("variant", ["naive", "tiled"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_forward_traces_bit_identical_to_oracle(ss, dev, oracle, dtype, variant):
    m, w, grid, pr, md = _setup(ss, dev, dtype, variant=variant)
    tr = pr.traces(md).cpu().numpy()
    want = oracle.fwi_forward(m, dtype(PROB["h"]), dtype(PROB["dt"]), PROB["n"], PROB["src"], PROB["rec_z"], w)
    assert np.abs(want).max() > 0
    assert np.array_equal(tr, want)


@pytest.mark.parametrize("variant", ["naive", "tiled"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ckpt", [0, 1, 5, 36])
def test_gradient_bit_identical_to_oracle(ss, dev, oracle, dtype, ckpt, variant):
    """Checkpoint recomputation reproduces the stored levels exactly, so every
    segment length gives the oracle's gradient (which keeps all levels)."""
    import torch

    m, w, grid, pr, md = _setup(ss, dev, dtype, ckpt, variant=variant)
    h, dt, n, src, rz = PROB["h"], PROB["dt"], PROB["n"], PROB["src"], PROB["rec_z"]
    m_true = m.copy()
    m_true[10:14, 8:12, 10:15] *= dtype(1.1)
    d_obs = oracle.fwi_forward(m_true, dtype(h), dtype(dt), n, src, rz, w)
    J_o, g_o, st_o = oracle.fwi_gradient(m, dtype(h), dtype(dt), n, src, rz, w, d_obs)
    J = pr.misfit(md, torch.from_numpy(d_obs).to(dev)).item()
    g, st = pr.gradient(md, src_trace=True)
    assert J == pytest.approx(J_o, rel=1e-12)
    assert np.array_equal(g.cpu().numpy(), g_o)
    assert np.array_equal(st.cpu().numpy(), st_o)
    assert np.abs(g_o).max() > 0


@pytest.mark.parametrize("shape,tuned", [((40, 44, 72), False), ((52, 36, 136), True), ((30, 70, 20), True)])
def test_tiled_fwi_equals_naive(ss, dev, shape, tuned):
    """FP32 tiled forward / recompute / adjoint kernels against the
    global-memory ones on ragged grids, with the default launch shape or the
    one the wave autotuner caches for the grid: traces, misfit, gradient and
    source trace bit-identical."""
    import torch

    n = 40
    prob = dict(shape=shape, h=10.0, dt=1e-3, n=n, src=(shape[0] // 2, shape[1] // 2, shape[2] // 2), rec_z=5)
    m = smooth_model(shape, 3, 1800.0, 3000.0, np.float32)
    w = (1e3 * ss.ricker(n, 1e-3, 20.0)).astype(np.float32)
    grid = _grid(ss, shape, 10.0, 1e-3, np.float32)
    md = torch.from_numpy(m).to(dev)
    ss.clear_tuning()
    if tuned:
        tb = [torch.rand(shape, device=dev) for _ in range(3)]
        ss.tune_tiled(tb[0], tb[1], md, tb[2], grid, reps=1)
        assert ss.tuned_shape(grid)[1]
    out = {}
    for variant in ("naive", "tiled"):
        pr = ss.FwiProblem(grid, n, prob["src"], prob["rec_z"], w, dev, torch.float32, 7, variant)
        d_obs = pr.traces(md * 1.05)
        J = pr.misfit(md, d_obs).item()
        g, st = pr.gradient(md, src_trace=True)
        out[variant] = (d_obs.cpu().numpy(), J, g.cpu().numpy(), st.cpu().numpy())
    a, b = out["naive"], out["tiled"]
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1]
    assert np.array_equal(a[2], b[2]) and np.abs(a[2]).max() > 0
    assert np.array_equal(a[3], b[3])
    ss.clear_tuning()


@pytest.mark.parametrize("variant", ["naive", "tiled"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_single_steps_match_oracle_step(ss, dev, oracle, dtype, variant):
    """stencil_forward_step / stencil_adjoint_step against the oracle's wave
    step plus the FWI terms formed in numpy in the same precision and order:
    the forward step with and without the source, L p out (the oracle step
    with p_prev = 2p, dt = 0.5 and vel2 = 4 returns L p exactly), and the
    adjoint step with the residual plane, the gradient term and the source
    value."""
    import torch

    shape = (26, 33, 44)
    nz, ny, nx = shape
    h, dt = dtype(10.0), dtype(1e-3)
    rng = np.random.default_rng(9)
    halo0 = np.zeros(shape, dtype)
    halo0[4:-4, 4:-4, 4:-4] = 1
    a, b = (rng.standard_normal(shape).astype(dtype) * halo0 for _ in range(2))
    m = (4e6 * (1 + 0.1 * rng.random(shape))).astype(dtype)
    grid = ss.GridDims(nx, ny, nz, h, h, h, dt)
    src = (13, 16, 21)
    si = (src[0] * ny + src[1]) * nx + src[2]
    w = np.array([0.25, -1.5, 3.0], dtype)
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)

    # forward: oracle step, then the source term
    want = oracle.wave_step(a, b, m, np.zeros(shape, dtype), h, dt)
    out, lap = d(np.zeros(shape, dtype)), d(np.zeros(shape, dtype))
    ss.forward_step(d(a), d(b), d(m), out, grid, variant=variant)
    assert np.array_equal(out.cpu().numpy(), want)
    want_src = want.copy()
    want_src.flat[si] = want.flat[si] + w[2]
    ss.forward_step(d(a), d(b), d(m), out, grid, lap_out=lap, src_index=si, wavelet=d(w), n=2, variant=variant)
    assert np.array_equal(out.cpu().numpy(), want_src)
    # L a, exactly: 2a - 2a = 0, then (0.25 * 4) * L a
    lap_want = oracle.wave_step(a, 2 * a, np.full(shape, 4, dtype), np.zeros(shape, dtype), h, dtype(0.5))
    assert np.array_equal(lap.cpu().numpy(), lap_want) and np.abs(lap_want).max() > 0

    # adjoint: the same step in psi, then + m * res on plane rec_z, the
    # gradient term and psi_out / m at the source
    rec_z = 6
    res = rng.standard_normal((ny, nx)).astype(dtype)
    g0 = rng.standard_normal(shape).astype(dtype)
    base = oracle.wave_step(a, b, m, np.zeros(shape, dtype), h, dt)
    psi_want = base.copy()
    plane = psi_want[rec_z]
    plane[4:-4, 4:-4] = plane[4:-4, 4:-4] + m[rec_z, 4:-4, 4:-4] * res[4:-4, 4:-4]
    g_want = g0.copy()
    g_want[4:-4, 4:-4, 4:-4] = g0[4:-4, 4:-4, 4:-4] + (dt * dt * lap_want[4:-4, 4:-4, 4:-4]) * a[4:-4, 4:-4, 4:-4]
    psi, grad, sv = d(np.zeros(shape, dtype)), d(g0), d(np.zeros(2, dtype))
    ss.adjoint_step(d(a), d(b), d(m), psi, grid, lap=d(lap_want), grad=grad, res_plane=d(res), rec_z=rec_z,
                    src_index=si, src_value=sv, variant=variant)
    assert np.array_equal(psi.cpu().numpy(), psi_want)
    assert np.array_equal(grad.cpu().numpy(), g_want)
    assert sv.cpu().numpy()[0] == psi_want.flat[si] / m.flat[si]
    # no optional inputs: the plain step
    ss.adjoint_step(d(a), d(b), d(m), psi, grid, variant=variant)
    assert np.array_equal(psi.cpu().numpy(), base)


@pytest.mark.parametrize("shape", [(30, 44, 72), (41, 20, 36)])
def test_fwi_every_tile_shape_equals_naive(ss, dev, shape):
    """The FP32 FWI forward, recompute and adjoint kernels at every built tile
    shape, 1 to 3 z-chunks, both workgroup orders and both streaming modes
    (the three kernels given different shapes), against the global-memory
    kernels on ragged grids:
    traces, misfit, gradient and source trace bit-identical.  Four segments,
    so the backward pass overlaps recomputes with adjoint steps."""
    import torch

    n = 24
    src, rec_z = (shape[0] // 2, shape[1] // 2, shape[2] // 2), 5
    m = smooth_model(shape, 11, 1800.0, 3000.0, np.float32)
    w = (1e3 * ss.ricker(n, 1e-3, 20.0)).astype(np.float32)
    grid = _grid(ss, shape, 10.0, 1e-3, np.float32)
    md = torch.from_numpy(m).to(dev)
    ss.clear_tuning()
    pn = ss.FwiProblem(grid, n, src, rec_z, w, dev, torch.float32, 6, "naive")
    d_obs = pn.traces(md * 1.05)
    J_n = pn.misfit(md, d_obs).item()
    g_n, st_n = (t.cpu().numpy() for t in pn.gradient(md, src_trace=True))
    assert np.abs(g_n).max() > 0
    pt = ss.FwiProblem(grid, n, src, rec_z, w, dev, torch.float32, 6, "tiled")
    tiles = [(16, 16), (16, 32), (32, 8), (32, 16), (64, 4), (64, 8), (64, 12), (64, 16)]  # float4 x rows
    combos = [(t, c, r) for t in tiles for c in (1, 2, 3) for r in (0, 1)]
    for i, (t, c, r) in enumerate(combos):
        def shp(j):
            (tx, ty), cj, rj = combos[j % len(combos)]
            return {"tile_x": 4 * tx, "tile_y": ty, "n_chunks": cj, "xcd_remap": rj, "nontemporal": (j // 2) % 2}
        ss.clear_tuning()
        ss.set_tuned_shape(grid, "fwi_forward", shp(i))
        ss.set_tuned_shape(grid, "fwi_adjoint", shp(i + 17))
        ss.set_tuned_shape(grid, "fwi_recompute", shp(i + 31))
        J = pt.misfit(md, d_obs).item()
        g, st = pt.gradient(md, src_trace=True)
        assert J == J_n, (i, shp(i))
        assert np.array_equal(g.cpu().numpy(), g_n), (i, shp(i), shp(i + 17), shp(i + 31))
        assert np.array_equal(st.cpu().numpy(), st_n), i
    assert np.array_equal(pt.traces(md * 1.05).cpu().numpy(), d_obs.cpu().numpy())
    ss.clear_tuning()


def _naive_reference(ss, dev, shape, n, ckpt, seed):
    """(grid, m on the device, d_obs, J, gradient, source trace) of the
    global-memory kernels on a smooth model, data from a 5 % faster model."""
    import torch

    src, rec_z = (shape[0] // 2, shape[1] // 2, shape[2] // 2), 5
    m = smooth_model(shape, seed, 1800.0, 3000.0, np.float32)
    w = (1e3 * ss.ricker(n, 1e-3, 20.0)).astype(np.float32)
    grid = _grid(ss, shape, 10.0, 1e-3, np.float32)
    md = torch.from_numpy(m).to(dev)
    pn = ss.FwiProblem(grid, n, src, rec_z, w, dev, torch.float32, ckpt, "naive")
    d_obs = pn.traces(md * 1.05)
    J = pn.misfit(md, d_obs).item()
    g, st = (t.cpu().numpy() for t in pn.gradient(md, src_trace=True))
    assert np.abs(g).max() > 0
    return grid, md, d_obs, J, g, st, (src, rec_z, w)


@pytest.mark.parametrize("schedule", ["pair", "overlap", "serial"])
@pytest.mark.parametrize("ckpt", [7, 1, 39])
def test_backward_schedules_equal_naive(ss, dev, monkeypatch, schedule, ckpt):
    """The tiled FP32 backward pass in each schedule (STENCIL_FWI_SCHEDULE):
    the step pair (a recompute step and an adjoint step in one launch), the
    recompute on a second stream (the default), and one stream in turn.  All
    give the global-memory kernels' misfit, gradient and source trace bit for
    bit.  40 steps in segments of 7 (the short last segment leaves the first
    paired round with single recompute steps), of 1, or of 39 (two segments,
    the last one step long: one pair, then 38 single recompute steps)."""
    import torch

    shape, n = (36, 40, 104), 40
    grid, md, d_obs, J_n, g_n, st_n, (src, rec_z, w) = _naive_reference(ss, dev, shape, n, ckpt, 5)
    monkeypatch.setenv("STENCIL_FWI_SCHEDULE", schedule)
    ss.clear_tuning()
    pt = ss.FwiProblem(grid, n, src, rec_z, w, dev, torch.float32, ckpt, "tiled")
    assert pt.misfit(md, d_obs).item() == J_n
    g, st = pt.gradient(md, src_trace=True)
    assert np.array_equal(g.cpu().numpy(), g_n)
    assert np.array_equal(st.cpu().numpy(), st_n)


@pytest.mark.parametrize("shape", [(30, 44, 72), (41, 20, 36)])
def test_fwi_step_pair_every_shape_equals_naive(ss, dev, monkeypatch, shape):
    """The step pair at each tile it is built for (128 x 8, 128 x 16, 256 x 4,
    256 x 8 cells), 1 to 3 z-chunks, both workgroup orders and both streaming
    modes, with the single recompute / adjoint steps at other shapes: misfit,
    gradient and source trace bit-identical to the global-memory kernels on
    ragged grids (four segments of 6 steps)."""
    import torch

    n = 24
    grid, md, d_obs, J_n, g_n, st_n, (src, rec_z, w) = _naive_reference(ss, dev, shape, n, 6, 11)
    monkeypatch.setenv("STENCIL_FWI_SCHEDULE", "pair")
    pt = ss.FwiProblem(grid, n, src, rec_z, w, dev, torch.float32, 6, "tiled")
    tiles = [(32, 8), (32, 16), (64, 4), (64, 8)]  # float4 x rows
    combos = [(t, c, r, nt) for t in tiles for c in (1, 2, 3) for r in (0, 1) for nt in (0, 1)]
    for i, ((tx, ty), c, r, nt) in enumerate(combos):
        ss.clear_tuning()
        ss.set_tuned_shape(grid, "fwi_pair",
                           {"tile_x": 4 * tx, "tile_y": ty, "n_chunks": c, "xcd_remap": r, "nontemporal": nt})
        ss.set_tuned_shape(grid, "fwi_recompute", {"tile_x": 64, "tile_y": 16, "n_chunks": 1 + i % 3})
        assert pt.misfit(md, d_obs).item() == J_n
        g, st = pt.gradient(md, src_trace=True)
        assert np.array_equal(g.cpu().numpy(), g_n), combos[i]
        assert np.array_equal(st.cpu().numpy(), st_n), combos[i]
    ss.clear_tuning()


@pytest.mark.parametrize("ckpt", [1, 24])
def test_fwi_tune_without_overlap_timing(ss, dev, ckpt):
    """With one checkpoint per step, or one segment (every L p kept), the FWI
    tuner has no room or no use for the recompute-beside-adjoint timing: the
    recompute shape repeats the forward's and the gradient still equals the
    global-memory kernels' bit for bit."""
    import torch

    shape, n = (36, 40, 104), 24
    m = smooth_model(shape, 4, 1800.0, 3000.0, np.float32)
    w = (1e3 * ss.ricker(n, 1e-3, 20.0)).astype(np.float32)
    grid = _grid(ss, shape, 10.0, 1e-3, np.float32)
    md = torch.from_numpy(m).to(dev)
    ss.clear_tuning()
    pt = ss.FwiProblem(grid, n, (18, 20, 52), 6, w, dev, torch.float32, ckpt, tune=True)
    assert pt.shape["recompute"] == pt.shape["forward"]
    assert not ss.tuned_shape(grid, "fwi_recompute")[1]
    pn = ss.FwiProblem(grid, n, (18, 20, 52), 6, w, dev, torch.float32, ckpt, "naive")
    d_obs = pn.traces(md * 1.05)
    assert pt.misfit(md, d_obs).item() == pn.misfit(md, d_obs).item()
    assert np.array_equal(pt.gradient(md).cpu().numpy(), pn.gradient(md).cpu().numpy())
    ss.clear_tuning()


def test_gradient_replays_in_a_graph(ss, dev):
    """misfit() + gradient() captured in a HIP graph (torch.cuda.CUDAGraph):
    the backward pass's second stream joins the capture through its events,
    and three replays, with eager gradients between them, give the eager
    gradient bit for bit."""
    import torch

    shape, n = (36, 40, 104), 24
    m = smooth_model(shape, 4, 1800.0, 3000.0, np.float32)
    w = (1e3 * ss.ricker(n, 1e-3, 20.0)).astype(np.float32)
    grid = _grid(ss, shape, 10.0, 1e-3, np.float32)
    md = torch.from_numpy(m).to(dev)
    pr = ss.FwiProblem(grid, n, (18, 20, 52), 6, w, dev, torch.float32)
    d_obs = pr.traces(md * 1.05)
    pr.misfit(md, d_obs)
    want = pr.gradient(md).cpu().numpy()
    assert np.abs(want).max() > 0
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):  # warm-up outside the capture
        pr.misfit(md, d_obs)
        pr.gradient(md)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    out = torch.empty_like(md)
    with torch.cuda.graph(g):
        pr.misfit(md, d_obs)
        out.copy_(pr.gradient(md))
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), want)
        pr.misfit(md, d_obs)
        assert np.array_equal(pr.gradient(md).cpu().numpy(), want)


def test_fwi_problem_tune(ss, dev):
    """FwiProblem(tune=True) tunes the FWI kernels for its grid (forward,
    adjoint, recompute beside the adjoint, and the step pair); its traces
    and gradient stay bit-identical to the global-memory kernels' (receiver
    plane written by the tiled step), also when tune() runs between misfit()
    and gradient()."""
    import torch

    shape, n = (36, 40, 104), 24
    m = smooth_model(shape, 3, 1800.0, 3000.0, np.float32)
    w = (1e3 * ss.ricker(n, 1e-3, 20.0)).astype(np.float32)
    grid = _grid(ss, shape, 10.0, 1e-3, np.float32)
    md = torch.from_numpy(m).to(dev)
    ss.clear_tuning()
    pt = ss.FwiProblem(grid, n, (18, 20, 52), 6, w, dev, torch.float32, tune=True)
    assert set(pt.shape) == {"forward", "adjoint", "recompute", "pair"}
    for kind in ("fwi_forward", "fwi_adjoint", "fwi_recompute", "fwi_pair"):
        shp, hit = ss.tuned_shape(grid, kind)
        assert hit and shp["nontemporal"] in (0, 1)
        assert shp == pt.shape[kind[4:]]
    pn = ss.FwiProblem(grid, n, (18, 20, 52), 6, w, dev, torch.float32, variant="naive")
    a, b = pt.traces(md), pn.traces(md)
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy()) and a.abs().max() > 0
    # tuning between misfit and gradient leaves the forward pass's state intact
    d_obs = 0.5 * a
    pt.misfit(md, d_obs)
    pt.tune(reps=1)
    pn.misfit(md, d_obs)
    ga, gb = pt.gradient(md), pn.gradient(md)
    assert np.array_equal(ga.cpu().numpy(), gb.cpu().numpy()) and ga.abs().max() > 0
    ss.clear_tuning()


def test_gradient_matches_finite_differences_fp64(ss, dev):
    """Central differences of J on 16 random interior voxels (FP64), 40^3 x 160
    steps with a velocity anomaly between source and receivers; agreement
    within 1e-4 relative (measured ~1e-7)."""
    import torch

    shape, n = (40, 40, 40), 160
    prob = dict(shape=shape, h=10.0, dt=8e-4, n=n, src=(8, 20, 20), rec_z=6)
    m = smooth_model(shape, 7, 1800.0, 3000.0)
    m_true = m.copy()
    z, y, x = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    m_true[(z - 24) ** 2 + (y - 20) ** 2 + (x - 20) ** 2 <= 36] *= 1.21
    w = 1e3 * ss.ricker(n, prob["dt"], 20.0)
    grid = _grid(ss, shape, prob["h"], prob["dt"], np.float64)
    pr = ss.FwiProblem(grid, n, prob["src"], prob["rec_z"], w, dev, torch.float64)
    d_obs = pr.traces(torch.from_numpy(m_true).to(dev))
    md = torch.from_numpy(m).to(dev)
    J0 = pr.misfit(md, d_obs).item()
    g = pr.gradient(md).cpu().numpy()
    assert J0 > 0
    inner = np.zeros(shape, bool)
    inner[R:-R, R:-R, R:-R] = True
    rng = np.random.default_rng(11)
    cand = np.argwhere(inner & (np.abs(g) > 1e-2 * np.abs(g).max()))
    picks = [tuple(c) for c in cand[rng.choice(len(cand), 16, replace=False)]]
    errs = []
    for idx in picks:
        eps = 1e-6 * m[idx]
        mp, mm = m.copy(), m.copy()
        mp[idx] += eps
        mm[idx] -= eps
        Jp = pr.misfit(torch.from_numpy(mp).to(dev), d_obs).item()
        Jm = pr.misfit(torch.from_numpy(mm).to(dev), d_obs).item()
        fd = (Jp - Jm) / (2 * eps)
        errs.append(abs(fd - g[idx]) / abs(g[idx]))
    print("FD relative errors:", ["%.1e" % e for e in errs])
    assert max(errs) < 1e-4


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-4)])
def test_adjoint_dot_product_identity(ss, dev, dtype, tol):
    """<P b, r> = <b, P^T r>: P maps the source wavelet to the receiver traces,
    P^T the receiver data to the adjoint field at the source."""
    import torch

    m, w, grid, pr, md = _setup(ss, dev, dtype)
    d = pr.traces(md)
    rng = np.random.default_rng(5)
    r = rng.standard_normal(tuple(d.shape)).astype(dtype)
    r[:, :R, :] = r[:, -R:, :] = 0
    r[:, :, :R] = r[:, :, -R:] = 0
    rt = torch.from_numpy(r).to(dev)
    pr.misfit(md, d - rt)  # residual = r
    _, lam_src = pr.gradient(md, src_trace=True)
    lhs = float((d.double() * rt.double()).sum())
    rhs = float((pr.wavelet.double() * lam_src.double()).sum())
    assert rhs == pytest.approx(lhs, rel=tol)


def test_autograd_function(ss, dev):
    import torch

    m, w, grid, pr, md = _setup(ss, dev, np.float64)
    m_true = m.copy()
    m_true[10:14, 8:12, 10:15] *= 1.1
    d_obs = pr.traces(torch.from_numpy(m_true).to(dev))
    v = torch.sqrt(md).clone().requires_grad_(True)
    J = ss.fwi_misfit(v * v, pr, d_obs)
    J.backward()
    pr.misfit(md, d_obs)
    g = pr.gradient(md)
    assert torch.allclose(v.grad, 2 * v.detach() * g, rtol=1e-12, atol=0)
    assert float(J) == pytest.approx(pr.misfit(md, d_obs).item(), rel=1e-14)


def test_autograd_two_misfits_on_one_problem(ss, dev):
    """Two fwi_misfit calls on one FwiProblem summed before backward(), and a
    line-search misfit in between: each vel2 still gets its own gradient
    (the workspace holds only the latest call's checkpoints and residual)."""
    import torch

    m, w, grid, pr, md = _setup(ss, dev, np.float64)
    m_true = m.copy()
    m_true[10:14, 8:12, 10:15] *= 1.1
    d_obs = pr.traces(torch.from_numpy(m_true).to(dev))
    m2 = torch.from_numpy(m * 1.02).to(dev)
    want = []
    for mm in (md, m2):
        pr.misfit(mm, d_obs)
        want.append(pr.gradient(mm).clone())
    assert not torch.equal(want[0], want[1])
    a = md.clone().requires_grad_(True)
    b = m2.clone().requires_grad_(True)
    J = ss.fwi_misfit(a, pr, d_obs) + ss.fwi_misfit(b, pr, d_obs)
    pr.misfit(torch.from_numpy(m * 0.97).to(dev), d_obs)  # e.g. a line-search trial
    J.backward()
    assert torch.equal(a.grad, want[0]) and torch.equal(b.grad, want[1])


def test_gradient_needs_the_misfit_vel2(ss, dev):
    import torch

    m, w, grid, pr, md = _setup(ss, dev, np.float32)
    d_obs = pr.traces(md * 1.05)
    pr.misfit(md, d_obs)
    with pytest.raises(RuntimeError):
        pr.gradient(md.clone())  # another tensor
    md.mul_(1.0)  # same values, but modified in place since the misfit
    with pytest.raises(RuntimeError):
        pr.gradient(md)
    pr.misfit(md, d_obs)
    assert pr.gradient(md).abs().max() > 0


def test_single_steps_reject_bad_auxiliary_arrays(ss, dev):
    """The wavelet, residual plane and source-value arrays are indexed by the
    kernels at caller-chosen offsets: a host tensor, another dtype, too few
    values or an out-of-grid source raise ValueError before any launch."""
    import torch

    grid = _grid(ss, (20, 20, 24), 10.0, 1e-3, np.float32)
    f = [torch.zeros((20, 20, 24), device=dev) for _ in range(5)]
    src = 10 * 24 * 20 + 10 * 24 + 12
    w = torch.ones(8, device=dev)
    torch.cuda.synchronize()
    bad_fwd = [dict(src_index=src, wavelet=w.cpu(), n=0),
               dict(src_index=src, wavelet=w.double(), n=0),
               dict(src_index=src, wavelet=w, n=8),
               dict(src_index=src, wavelet=w, n=-1),
               dict(src_index=src, wavelet=None, n=0),
               dict(src_index=grid.size(), wavelet=w, n=0)]
    for kw in bad_fwd:
        with pytest.raises(ValueError):
            ss.forward_step(f[0], f[1], f[2], f[3], grid, **kw)
    ss.forward_step(f[0], f[1], f[2], f[3], grid, src_index=src, wavelet=w, n=7)  # the last sample is fine
    plane = torch.zeros(20 * 24, device=dev)
    sv = torch.zeros(1, device=dev)
    bad_adj = [dict(res_plane=plane.cpu(), rec_z=6),
               dict(res_plane=plane.double(), rec_z=6),
               dict(res_plane=plane[:-1], rec_z=6),
               dict(src_index=src, src_value=sv.cpu()),
               dict(src_index=src, src_value=sv.double()),
               dict(src_index=src, src_value=sv[:0]),
               dict(src_index=-1, src_value=sv),
               dict(src_index=grid.size(), src_value=sv)]
    for kw in bad_adj:
        with pytest.raises(ValueError):
            ss.adjoint_step(f[0], f[1], f[2], f[3], grid, **kw)
    ss.adjoint_step(f[0], f[1], f[2], f[3], grid, res_plane=plane, rec_z=6, src_index=src, src_value=sv)
    torch.cuda.synchronize()


def test_fwi_rejects_bad_configuration(ss, dev):
    import torch

    grid = _grid(ss, (20, 20, 20), 10.0, 1e-3, np.float32)
    with pytest.raises(ValueError):
        ss.FwiProblem(grid, 10, (2, 10, 10), 6, np.zeros(10), dev)  # source in the halo
    with pytest.raises(ValueError):
        ss.FwiProblem(grid, 10, (10, 10, 10), 17, np.zeros(10), dev)  # receivers in the halo
    with pytest.raises(ValueError):
        ss.FwiProblem(grid, 10, (10, 10, 10), 6, np.zeros(9), dev)  # wavelet length
    pr = ss.FwiProblem(grid, 10, (10, 10, 10), 6, np.zeros(10), dev)
    with pytest.raises(RuntimeError):
        pr.gradient(torch.ones((20, 20, 20), device=dev))
    with pytest.raises(ValueError):
        pr.traces(torch.ones((20, 20, 21), device=dev))
