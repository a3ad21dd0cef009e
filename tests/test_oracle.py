"""Pins the oracle (oracle/stencil_oracle.c) before anything is checked against it.

1. Golden vectors produced by the reference's own CPU path (tests/golden/,
   written by tests/golden/make_golden.py from oracle/_ref): bit-identical to the
   -ffp-contract=off build, within 1e-5 relative L2 of the reference's Release
   (-O3 -march=native, FMA-contracted) build.
2. What the mathematics fixes: exact rational weights, polynomial exactness of
   the 8th-order Laplacian, the discrete dispersion relation of a plane wave,
   fixed points, SplitMix64's published outputs.
3. Each closed-form check is shown to fail on a deliberately wrong restatement
   (a planted mistake in a numpy copy of the step, on the CPU), so the checks
   are not restatements of the formula they test.
4. The relative-L2 function every tolerance test uses (it has no reference
   counterpart) against values its definition fixes (a == b, a == 2b, a == 0,
   halo cells ignored, one perturbed cell on each interior edge, b == 0), with
   planted mistakes (halo counted, off-by-one interior, no square root, wrong
   denominator, RMS, transposed axes, b narrowed to float32) each caught.
"""
from __future__ import annotations

import hashlib
from fractions import Fraction

import numpy as np
import pytest

from conftest import load_golden

R = 4
RATIONALS = [Fraction(-205, 72), Fraction(8, 5), Fraction(-1, 5), Fraction(8, 315), Fraction(-1, 560)]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ── 1. golden vectors from the reference CPU path ──────────────────────────────
def test_fd_coeffs_are_the_rationals(oracle):
    c32 = oracle.fd_coeffs(np.float32)
    c64 = oracle.fd_coeffs(np.float64)
    for i, q in enumerate(RATIONALS):
        assert c64[i] == float(q)
        assert c32[i] == np.float32(float(q))
    # consistency c0 + 2 sum c = 0 holds exactly for the rationals
    assert RATIONALS[0] + 2 * sum(RATIONALS[1:]) == 0


def test_make_grid_dt_matches_reference(oracle, golden_meta):
    for row in golden_meta["make_grid_dt"]:
        assert oracle.make_grid_dt(np.float32(row["dx"]), np.float32(row["velocity"])) == row["dt"]


def test_gaussian_matches_reference(oracle, golden_meta):
    for row in golden_meta["gaussian"]:
        g = oracle.gaussian(row["n"], row["c"], row["c"], row["c"], row["sigma"])
        assert sha(g) == row["sha"]
        n = row["n"]
        assert float(g[n // 2, n // 2, n // 2]) == row["centre_val"]
        assert np.all(g[:R] == 0) and np.all(g[-R:] == 0)
        assert np.all(g[:, :R] == 0) and np.all(g[:, :, -R:] == 0)


def test_kernel_step_matches_reference(oracle, golden_meta):
    gd = load_golden("kernel_24.npz")
    m = golden_meta["kernel_24"]
    g = oracle.gaussian(m["n"], sigma=m["sigma"])
    assert np.array_equal(g, gd["init"])
    vel2 = np.full_like(g, np.float32(1500.0) * np.float32(1500.0))
    nxt = np.zeros_like(g)
    oracle.wave_step(g, g, vel2, nxt, np.float32(10.0), np.float32(1e-3))
    assert np.array_equal(nxt, gd["next_nofma"])
    assert oracle.rel_l2_interior(nxt, gd["next_native"]) <= 1e-6


@pytest.mark.parametrize("steps", [1, 20])
def test_wave_solver_matches_reference(oracle, golden_meta, steps):
    gd = load_golden("solver_wave_32.npz")
    m = golden_meta["solver_wave_32"]
    p0 = oracle.gaussian(32, sigma=m["sigma"])
    vel2 = oracle.constant_velocity(p0.shape, np.float32(1500.0))
    dt = oracle.make_grid_dt(np.float32(10.0), np.float32(1500.0))
    assert dt == m["dt_make_grid"]
    out = oracle.wave_run(p0, vel2, np.float32(10.0), dt, steps)
    assert np.array_equal(out, gd[f"naive{steps}_nofma"])
    # CPU launch_tiled == launch_naive in the reference (stencil_naive.cu:173-176)
    assert np.array_equal(out, gd[f"tiled{steps}_nofma"])
    assert oracle.rel_l2_interior(out, gd[f"naive{steps}_native"]) <= 1e-5


def test_run_warmup_advances_state(oracle):
    """run(5, warmup=2) leaves the state 7 steps in (solver_base.cpp:71-73)."""
    gd = load_golden("solver_wave_32.npz")
    p0 = oracle.gaussian(32, sigma=4.0)
    vel2 = oracle.constant_velocity(p0.shape, np.float32(1500.0))
    dt = oracle.make_grid_dt(np.float32(10.0), np.float32(1500.0))
    assert np.array_equal(oracle.wave_run(p0, vel2, np.float32(10.0), dt, 7), gd["run5_warm2_nofma"])


def test_heat_solver_matches_reference(oracle, golden_meta):
    gd = load_golden("solver_heat_48.npz")
    m = golden_meta["solver_heat_48"]
    t0 = oracle.gaussian(48, sigma=m["sigma"])
    dt = oracle.make_grid_dt(np.float32(1.0), np.float32(500.0))
    dte = oracle.heat_dt_eff(dt)
    assert dte == np.float32(np.sqrt(np.float64(np.float32(dt))))
    out = oracle.heat_run(t0, np.float32(1.0), dte, np.float32(1e-3), m["steps"])
    assert np.array_equal(out, gd["heat20_nofma"])
    assert oracle.rel_l2_interior(out, gd["heat20_native"]) <= 1e-5


@pytest.fixture(scope="module")
def acoustic128(oracle):
    p0 = oracle.gaussian(128, sigma=3.0)
    vel2 = oracle.constant_velocity(p0.shape, np.float32(2000.0))
    return p0, oracle.wave_run(p0, vel2, np.float32(10.0), np.float32(1e-3), 200)


def test_acoustic128_matches_reference(oracle, golden_meta, acoustic128):
    """The 128^3 x 200-step workload (h=10, v=2000, dt=1e-3, sigma=3)."""
    p0, fin = acoustic128
    gd = load_golden("acoustic_128.npz")
    assert sha(p0) == golden_meta["acoustic128_nofma"]["sha_init"]
    assert sha(fin) == golden_meta["acoustic128_nofma"]["sha_final"]
    idx = gd["sample_idx_nofma"]
    assert np.array_equal(fin.reshape(-1)[idx], gd["sample_val_nofma"])
    # The reference's own Release build: mid plane + samples within 1e-5.
    mid = fin[64]
    nat = gd["mid_z_native"]
    assert np.linalg.norm(mid - nat) / np.linalg.norm(nat) <= 1e-5
    sv = gd["sample_val_native"].astype(np.float64)
    assert np.linalg.norm(fin.reshape(-1)[gd["sample_idx_native"]] - sv) <= 1e-5 * np.linalg.norm(sv) + 1e-12
    s = fin.astype(np.float64)
    assert abs(np.sqrt((s * s).sum()) - np.sqrt(golden_meta["acoustic128_native"]["sum_sq"])) < 1e-5


def test_acoustic128_fp64_matches_reference(oracle, golden_meta):
    p0 = oracle.gaussian(128, sigma=3.0, dtype=np.float64)
    vel2 = np.full_like(p0, 2000.0 * 2000.0)
    fin = oracle.wave_run(p0, vel2, 10.0, 1e-3, 200)
    assert sha(p0) == golden_meta["acoustic128_f64"]["sha_init"]
    assert sha(fin) == golden_meta["acoustic128_f64"]["sha_final"]


def test_layered_velocity_matches_reference(oracle, golden_meta):
    """init_layered_velocity (grid.hpp:91-99) against the reference's own
    output: v_top^2 for iz < nz/2, v_bot^2 below, in every cell."""
    for row in golden_meta["layered_velocity"]:
        shape = tuple(row["shape"])
        lv = oracle.layered_velocity(shape, np.float32(row["v_top"]), np.float32(row["v_bot"]))
        assert sha(lv) == row["sha"]
        assert float(lv.reshape(-1)[0]) == row["first"] and float(lv.reshape(-1)[-1]) == row["last"]
        vt2 = np.float32(row["v_top"]) * np.float32(row["v_top"])
        vb2 = np.float32(row["v_bot"]) * np.float32(row["v_bot"])
        nz = shape[0]
        assert np.all(lv[: nz // 2] == vt2) and np.all(lv[nz // 2:] == vb2)


def test_norms_match_reference(oracle, golden_meta):
    """l2_error (absolute RMS over every cell) and max_abs_error against the
    reference's outputs on the same seeded inputs."""
    for row in golden_meta["norms"]:
        rng = np.random.default_rng(row["seed"])
        a = rng.uniform(-1.0, 1.0, tuple(row["shape"])).astype(np.float32)
        b = rng.uniform(-1.0, 1.0, tuple(row["shape"])).astype(np.float32)
        assert oracle.l2_error(a, b) == row["l2_error"]
        assert oracle.max_abs_error(a, b) == row["max_abs_error"]


def test_norms_match_reference_semantics(oracle):
    a = np.zeros((4, 4, 4), np.float32)
    b = np.ones((4, 4, 4), np.float32)
    assert oracle.l2_error(a, b) == pytest.approx(1.0)  # RMS over ALL cells (grid.cpp:86-97)
    b2 = np.zeros_like(a)
    b2[2, 2, 2] = 5.0
    assert oracle.max_abs_error(a, b2) == 5.0


# ── 2. closed forms (float64 oracle where exactness needs headroom) ───────────
def _mk(shape, fn, dtype=np.float64):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.ascontiguousarray(fn(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64))).astype(dtype)


def _interior(a):
    return a[R:-R, R:-R, R:-R]


def _lap_via_step(step, p, dx, dy, dz):
    """Lap(p) recovered from one step with p_prev = p, dt^2 v^2 = 1."""
    nxt = np.zeros_like(p)
    vel2 = np.ones_like(p)
    step(p, p, vel2, nxt, dx, 1.0, dy, dz)
    return _interior(nxt) - _interior(p)


def check_polynomial_exactness(step):
    """The radius-4 central stencil is exact for polynomials of degree <= 9:
    p = X^8 + 2 Y^6 + 3 Z^4 - X Y Z with X = ix*dx etc. (distinct spacings)."""
    dx, dy, dz = 0.5, 0.25, 0.75
    shape = (20, 22, 18)
    p = _mk(shape, lambda x, y, z: (x * dx) ** 8 + 2 * (y * dy) ** 6 + 3 * (z * dz) ** 4 - (x * dx) * (y * dy) * (z * dz))
    exact = _mk(shape, lambda x, y, z: 56 * (x * dx) ** 6 + 60 * (y * dy) ** 4 + 36 * (z * dz) ** 2)
    lap = _lap_via_step(step, p, dx, dy, dz)
    err = np.max(np.abs(lap - _interior(exact)) / (1 + np.abs(_interior(exact))))
    return err


def check_dispersion(step, n_steps=4):
    """A plane wave cos(k.X) is an eigenfunction of the discrete Laplacian,
    lambda = sum_a (c0 + 2 sum_r c_r cos(r k_a h_a)) / h_a^2; with p_prev = p_0 the
    leapfrog gives p_n = cos((n + 1/2) theta) / cos(theta / 2) * p_0,
    cos(theta) = 1 + dt^2 v^2 lambda / 2, in cells more than 4n from every wall."""
    h = (10.0, 12.0, 8.0)
    k = (0.21, 0.13, 0.17)
    v, dt = 1500.0, 1e-3
    shape = (48, 48, 48)
    p0 = _mk(shape, lambda x, y, z: np.cos(k[0] * x * h[0] + k[1] * y * h[1] + k[2] * z * h[2]))
    c = [float(q) for q in RATIONALS]
    lam = sum((c[0] + 2 * sum(c[r] * np.cos(r * k[a] * h[a]) for r in range(1, 5))) / h[a] ** 2 for a in range(3))
    theta = np.arccos(1 + 0.5 * dt * dt * v * v * lam)
    bufs = [p0.copy(), p0.copy(), np.zeros_like(p0)]
    vel2 = np.full_like(p0, v * v)
    i = 0
    for _ in range(n_steps):
        step(bufs[(i + 1) % 3], bufs[i], vel2, bufs[(i + 2) % 3], h[0], dt, h[1], h[2])
        i = (i + 1) % 3
    out = bufs[(i + 1) % 3]
    m = R + 4 * n_steps
    want = np.cos((n_steps + 0.5) * theta) / np.cos(0.5 * theta) * p0
    return np.max(np.abs(out[m:-m, m:-m, m:-m] - want[m:-m, m:-m, m:-m]))


def test_polynomial_exactness(oracle):
    assert check_polynomial_exactness(oracle.wave_step) < 1e-9


def test_plane_wave_dispersion(oracle):
    assert check_dispersion(oracle.wave_step) < 1e-12


def test_flat_and_linear_fields_are_fixed_points(oracle):
    """tests/unit/test_kernels.cpp:100-138 (float32, same tolerances)."""
    shape = (24, 24, 24)
    flat = np.full(shape, 2.0, np.float32)
    nxt = np.zeros_like(flat)
    oracle.wave_step(flat, flat, np.ones_like(flat), nxt, np.float32(10.0), np.float32(1e-3))
    assert np.max(np.abs(_interior(nxt) - 2.0)) <= 1e-5
    lin = _mk(shape, lambda x, y, z: np.float32(0.1) * x + np.float32(0.2) * y + np.float32(0.3) * z, np.float32)
    nxt = np.zeros_like(lin)
    oracle.wave_step(lin, lin, np.ones_like(lin), nxt, np.float32(10.0), np.float32(1e-3))
    assert np.all(np.abs(_interior(nxt) - _interior(lin)) <= np.abs(_interior(lin)) * 1e-4 + 1e-6)


def test_halo_never_written(oracle):
    p = oracle.gaussian(24, sigma=3.0)
    nxt = np.full_like(p, np.float32(7.5))
    oracle.wave_step(p, p, np.ones_like(p), nxt, np.float32(10.0), np.float32(1e-3))
    mask = np.ones(p.shape, bool)
    mask[R:-R, R:-R, R:-R] = False
    assert np.all(nxt[mask] == np.float32(7.5))


def test_splitmix64_published_outputs(oracle):
    """SplitMix64 (Steele, Lea & Flood 2014), seed 1234567: the first outputs of
    the reference generator."""
    g = 0x9E3779B97F4A7C15
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    assert [oracle.splitmix64((1234567 + i * g) % 2**64) for i in range(5)] == want


def test_u01_range_and_moments(oracle):
    u = np.array([oracle.u01(2, 2, i) for i in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    assert np.all(u * 16777216.0 == np.floor(u * 16777216.0))  # 24-bit lattice


def test_layered_noise_model(oracle):
    shape = (64, 8, 8)
    v2 = oracle.layered_noise_vel2(shape, n_layers=16, vmin=1500.0, vmax=4000.0, noise=0.05, seed=2)
    v = np.sqrt(v2.astype(np.float64))
    for layer in range(16):
        band = v[layer * 4:(layer + 1) * 4]
        vl = 1500.0 + 2500.0 * oracle.u01(2, 1, layer)
        assert band.min() >= vl * 0.95 * (1 - 1e-6) and band.max() <= vl * 1.05 * (1 + 1e-6)
    assert v.min() >= 1500 * 0.95 - 1e-3 and v.max() <= 4000 * 1.05 + 1e-3


def test_heat_noise_cube(oracle):
    t = oracle.heat_noise_cube((32, 32, 32), side=8, hot=10.0, seed=2)
    assert np.all(t[12:20, 12:20, 12:20] == 10.0)
    assert t[:R].max() == 0 and t[:, :, -R:].max() == 0
    inner = t[R:-R, R:-R, R:-R].copy()
    inner[8:16, 8:16, 8:16] = 0.5
    assert inner.min() >= 0 and inner.max() < 1


# ── 3. the closed-form checks have teeth: planted mistakes fail them ──────────
def _numpy_step(mutation=None):
    c = [float(q) for q in RATIONALS]
    if mutation == "c3_sign":
        c[3] = -c[3]
    if mutation == "drop_r4":
        c[4] = 0.0

    def step(cur, prev, vel2, nxt, dx, dt, dy=None, dz=None):
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        if mutation == "swap_dx_dz":
            dx, dz = dz, dx
        p = cur
        s = (slice(R, -R),) * 3
        ctr = p[s]

        def sh(axis, r):
            sl = [slice(R, -R)] * 3
            n = p.shape[axis]
            sl[axis] = slice(R + r, n - R + r)
            return p[tuple(sl)]

        lap = np.zeros_like(ctr)
        for axis, h in ((2, dx), (1, dy), (0, dz)):
            acc = c[0] * ctr
            for r in range(1, 5):
                acc = acc + c[r] * (sh(axis, r) + sh(axis, -r))
            lap = lap + acc / (h * h)
        sign = -1.0 if mutation == "prev_sign" else 1.0
        nxt[s] = 2 * ctr - sign * prev[s] + dt * dt * vel2[s] * lap
        return nxt

    return step


@pytest.mark.parametrize("mutation", ["c3_sign", "drop_r4", "swap_dx_dz", "prev_sign"])
def test_planted_mistakes_are_caught(mutation):
    good = _numpy_step(None)
    assert check_polynomial_exactness(good) < 1e-9 and check_dispersion(good) < 1e-12
    bad = _numpy_step(mutation)
    caught = check_polynomial_exactness(bad) > 1e-6 or check_dispersion(bad) > 1e-6
    assert caught, mutation


# ── 4. the relative-L2 checker itself (no reference counterpart: the reference
#       has only absolute RMS, grid.cpp:86-97), pinned by closed forms ───────
def check_rel_l2(fn, f64_ref=False):
    """Largest deviation of fn(a, b) = ||a - b||_2 / ||b||_2 over the interior
    [R, n-R)^3 from values the definition fixes; b is float64 when f64_ref."""
    rng = np.random.default_rng(11)
    shape = (10, 11, 13)  # nz, ny, nx all different: a transposed index shows
    bt = np.float64 if f64_ref else np.float32
    b = rng.uniform(0.5, 2.0, shape).astype(bt)
    n_int = (shape[0] - 2 * R) * (shape[1] - 2 * R) * (shape[2] - 2 * R)
    f32 = lambda x: np.array(x, dtype=np.float32)  # noqa: E731  (always a copy)
    errs = []
    errs.append(abs(fn(f32(b), b) - 0.0) if not f64_ref else 0.0)  # a == b
    b2 = b.astype(np.float32).astype(bt)  # exactly representable in float32
    errs.append(abs(fn(f32(b2), b2) - 0.0))
    errs.append(abs(fn(f32(2 * b2), b2) - 1.0))  # a == 2b
    errs.append(abs(fn(f32(np.zeros(shape)), b2) - 1.0))  # a == 0
    errs.append(abs(fn(f32(-b2), b2) - 2.0))  # a == -b
    # halo cells are ignored: scribble over every halo cell of a and b
    a = f32(2 * b2)
    ah, bh = a.copy(), b2.copy()
    halo = np.ones(shape, bool)
    halo[R:-R, R:-R, R:-R] = False
    ah[halo] = 1e6
    bh[halo] = -3e5
    errs.append(abs(fn(ah, bh) - fn(a, b2)))
    # one interior cell off by d, every interior b = 1: d / sqrt(N_interior);
    # the cells chosen sit on the interior's last row / column / plane
    ones = np.zeros(shape, bt)
    ones[R:-R, R:-R, R:-R] = 1.0
    for cell in ((shape[0] - R - 1, R, R), (R, shape[1] - R - 1, R), (R, R, shape[2] - R - 1)):
        a1 = f32(ones)
        a1[cell] += 0.5
        errs.append(abs(fn(a1, ones) - 0.5 / np.sqrt(n_int)))
    # the first halo cell past each interior edge does not count
    for cell in ((shape[0] - R, R, R), (R, shape[1] - R, R), (R, R, shape[2] - R), (R - 1, R, R)):
        a1 = f32(ones)
        a1[cell] += 0.5
        errs.append(abs(fn(a1, ones)))
    # b == 0 on the interior: the absolute norm sqrt(sum a^2)
    z = np.zeros(shape, bt)
    a1 = np.zeros(shape, np.float32)
    a1[R, R + 1, R + 2] = 3.0
    a1[R + 1, R, R] = 4.0
    errs.append(abs(fn(a1, z) - 5.0))
    if f64_ref:
        # b in double, not representable in float32: |fl32(1/3) - 1/3| / (1/3)
        third = np.full(shape, 1.0 / 3.0)
        want = abs(float(np.float32(1.0 / 3.0)) - 1.0 / 3.0) * 3.0
        errs.append(abs(fn(f32(third), third) - want) / want * 1e-3)
    return max(errs)


def test_rel_l2_interior_closed_forms(oracle):
    assert check_rel_l2(oracle.rel_l2_interior) < 1e-12
    assert check_rel_l2(oracle.rel_l2_interior, f64_ref=True) < 1e-12


def _numpy_rel_l2(mutation=None):
    def fn(a, b):
        lo = R + 1 if mutation == "off_by_one" else R
        s = (slice(lo, -R),) * 3 if mutation != "halo_included" else (slice(None),) * 3
        if mutation == "transposed":
            s = (slice(R, -R), slice(R, -R), slice(R, a.shape[2] - R)) if a.shape[0] == a.shape[2] else \
                (slice(R, a.shape[2] - R), slice(R, -R), slice(R, a.shape[0] - R))
        bb = b[s].astype(np.float32 if mutation == "b_in_f32" else np.float64).astype(np.float64)
        d = a[s].astype(np.float64) - bb
        num = float((d * d).sum())
        den = float((bb * bb).sum()) if mutation != "den_from_a" else float((a[s].astype(np.float64) ** 2).sum())
        if mutation == "rms":
            return np.sqrt(num / d.size)
        if mutation == "no_sqrt":
            return num / den if den > 0 else num
        return np.sqrt(num / den) if den > 0 else np.sqrt(num)
    return fn


@pytest.mark.parametrize("mutation", ["halo_included", "off_by_one", "no_sqrt", "den_from_a", "rms",
                                      "transposed", "b_in_f32"])
def test_rel_l2_planted_mistakes_are_caught(mutation):
    assert check_rel_l2(_numpy_rel_l2(None)) < 1e-12
    assert check_rel_l2(_numpy_rel_l2(None), f64_ref=True) < 1e-12
    bad = _numpy_rel_l2(mutation)
    worst = max(check_rel_l2(bad), check_rel_l2(bad, f64_ref=True))
    assert worst > 1e-9, mutation
