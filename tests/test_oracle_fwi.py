"""Pins the oracle's FWI restatement (oracle/stencil_oracle.c, ORACLE_FWI): the
forward pass against the already-pinned acoustic step plus the source term,
the gradient against central finite differences of the misfit (FP64), and the
adjoint against the dot-product identity <P b, r> = <b, P^T r>."""
from __future__ import annotations

import numpy as np
import pytest

R = 4


def ricker(n, dt, f, t0=None):
    t0 = 1.2 / f if t0 is None else t0
    t = np.arange(n) * dt - t0
    a = (np.pi * f * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def smooth_model(shape, seed, lo, hi, dtype=np.float64):
    """A heterogeneous v^2 (smooth plus a bump), so L(m lam) != m L(lam)."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, 3)
    v = 0.5 + 0.25 * np.sin(0.31 * x + ph[0]) * np.cos(0.23 * y + ph[1]) + 0.25 * np.sin(0.17 * z + ph[2])
    v = lo + (hi - lo) * v
    return (v * v).astype(dtype)


PROB = dict(shape=(22, 20, 24), h=10.0, dt=1e-3, n=36, src=(9, 10, 12), rec_z=6)


def _problem(dtype=np.float64, seed=1):
    m = smooth_model(PROB["shape"], seed, 2000.0, 3200.0, dtype)
    w = (1e3 * ricker(PROB["n"], PROB["dt"], 25.0)).astype(dtype)
    return m, w


def test_forward_is_wave_step_plus_source(oracle):
    """Trace n = p^{n+1} on plane rec_z, p^{n+1} = oracle_wave_step(p^n, p^{n-1}) + w[n] at src."""
    m, w = _problem(np.float32)
    h, dt, n, src, rz = PROB["h"], PROB["dt"], PROB["n"], PROB["src"], PROB["rec_z"]
    tr = oracle.fwi_forward(m, np.float32(h), np.float32(dt), n, src, rz, w)
    p_prev = np.zeros_like(m)
    p = np.zeros_like(m)
    for k in range(n):
        nxt = np.zeros_like(m)
        oracle.wave_step(p, p_prev, m, nxt, np.float32(h), np.float32(dt))
        nxt[src] += w[k]
        assert np.array_equal(tr[k], nxt[rz]), k
        p_prev, p = p, nxt
    assert np.abs(tr).max() > 0
    # the batched-shot restatement ends on the same field
    assert np.array_equal(oracle.shot_run(m, np.float32(h), np.float32(dt), n, src, w), p)
    # and without a source stays zero
    assert not np.any(oracle.shot_run(m, np.float32(h), np.float32(dt), 5, None, None))


def test_gradient_matches_central_differences(oracle):
    """FP64: dJ/dm at single voxels against (J(m+e) - J(m-e)) / 2e."""
    m, w = _problem()
    h, dt, n, src, rz = PROB["h"], PROB["dt"], PROB["n"], PROB["src"], PROB["rec_z"]
    m_true = m.copy()
    m_true[10:14, 8:12, 10:15] *= 1.1
    d_obs = oracle.fwi_forward(m_true, h, dt, n, src, rz, w)
    J, g, _ = oracle.fwi_gradient(m, h, dt, n, src, rz, w, d_obs)
    assert J > 0
    inner = np.zeros(m.shape, bool)
    inner[R:-R, R:-R, R:-R] = True
    assert np.all(g[~inner] == 0)
    rng = np.random.default_rng(3)
    cand = np.argwhere(inner & (np.abs(g) > 1e-3 * np.abs(g).max()))
    picks = [tuple(c) for c in cand[rng.choice(len(cand), 10, replace=False)]]
    picks.append(tuple(np.unravel_index(np.argmax(np.abs(g)), g.shape)))
    for idx in picks:
        eps = 1e-6 * m[idx]
        mp, mm = m.copy(), m.copy()
        mp[idx] += eps
        mm[idx] -= eps
        Jp = oracle.fwi_gradient(mp, h, dt, n, src, rz, w, d_obs)[0]
        Jm = oracle.fwi_gradient(mm, h, dt, n, src, rz, w, d_obs)[0]
        fd = (Jp - Jm) / (2 * eps)
        assert fd == pytest.approx(g[idx], rel=1e-6), idx


def _dot_terms(oracle, dtype, seed=5):
    m, w = _problem(dtype)
    h, dt, n, src, rz = PROB["h"], PROB["dt"], PROB["n"], PROB["src"], PROB["rec_z"]
    d = oracle.fwi_forward(m, dtype(h), dtype(dt), n, src, rz, w)
    r = np.random.default_rng(seed).standard_normal(d.shape).astype(dtype)
    r[:, :R, :] = 0
    r[:, -R:, :] = 0
    r[:, :, :R] = 0
    r[:, :, -R:] = 0
    # residual = d - d_obs = r  ->  lambda solves the adjoint with source r
    _, _, lam_src = oracle.fwi_gradient(m, dtype(h), dtype(dt), n, src, rz, w, (d.astype(np.float64) - r).astype(dtype))
    lhs = float(np.dot(d.astype(np.float64).ravel(), r.astype(np.float64).ravel()))
    return lhs, w.astype(np.float64), lam_src.astype(np.float64)


def test_adjoint_dot_product_identity(oracle):
    lhs, w, lam = _dot_terms(oracle, np.float64)
    rhs = float(np.dot(w, lam))
    assert rhs == pytest.approx(lhs, rel=1e-11)
    # FP32: the same identity to rounding
    lhs32, w32, lam32 = _dot_terms(oracle, np.float32)
    assert float(np.dot(w32, lam32)) == pytest.approx(lhs32, rel=1e-4)


def test_dot_product_catches_a_time_shift(oracle):
    """A planted off-by-one (pairing b^n with lambda^n instead of lambda^{n+1})
    breaks the identity, so the test is not vacuous."""
    lhs, w, lam = _dot_terms(oracle, np.float64)
    shifted = float(np.dot(w[1:], lam[:-1]))
    assert abs(shifted - lhs) > 1e-3 * abs(lhs)
