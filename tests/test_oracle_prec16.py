"""Pins the oracle's 16-bit storage restatements (oracle/stencil_oracle.c) against
independent implementations: numpy's IEEE half conversion, an exact
rational-arithmetic rounding for bfloat16 and for the block format, and the
already-pinned FP32 step for the composed 16-bit step."""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

R = 4


def _samples(rng, n=4000):
    x = np.concatenate([
        rng.standard_normal(n).astype(np.float32),
        (rng.standard_normal(n) * 1e-5).astype(np.float32),       # fp16 subnormal range
        (rng.standard_normal(n) * 3e4).astype(np.float32),        # near fp16 max
        (rng.standard_normal(n) * 1e-40).astype(np.float32),      # fp32 subnormals
        np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65520.0, 1e9, 5.96e-8, 2.98e-8, 1.5 * 2.0**-24], np.float32),
    ])
    return x


def test_fp16_matches_numpy(oracle):
    """Round to nearest even, subnormals kept, overflow to inf: numpy's float16 cast."""
    rng = np.random.default_rng(1)
    x = _samples(rng)
    ours = np.array([oracle.lib.oracle_f32_to_fp16(float(v)) for v in x], dtype=np.uint16)
    ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)
    back = np.array([oracle.lib.oracle_fp16_to_f32(int(h)) for h in ref], dtype=np.float32)
    assert np.array_equal(back, ref.view(np.float16).astype(np.float32))


def _bf16_exact(v: np.float32) -> int:
    """Nearest bfloat16 by exact rational comparison (ties to even)."""
    b = int(np.float32(v).view(np.uint32))
    lo = b & 0xFFFF0000
    hi = (lo + 0x10000) & 0xFFFFFFFF
    f = Fraction(float(np.float32(v)))
    flo = Fraction(float(np.uint32(lo).view(np.float32)))
    fhi = Fraction(float(np.uint32(hi).view(np.float32)))
    dlo, dhi = abs(f - flo), abs(fhi - f)
    if dlo < dhi or (dlo == dhi and ((lo >> 16) & 1) == 0):
        return lo >> 16
    return hi >> 16


def test_bf16_round_to_nearest_even(oracle):
    rng = np.random.default_rng(2)
    x = _samples(rng, 1500)
    x = x[np.isfinite(x) & (np.abs(x) < 3e38)]
    for v in x:
        assert oracle.lib.oracle_f32_to_bf16(float(v)) == _bf16_exact(v), v
    # known values
    assert oracle.lib.oracle_f32_to_bf16(1.0) == 0x3F80
    assert oracle.lib.oracle_f32_to_bf16(float(np.float32(1 + 2**-8))) == 0x3F80  # tie -> even
    assert oracle.lib.oracle_f32_to_bf16(float(np.float32(1 + 3 * 2**-8))) == 0x3F82  # tie -> even (up)
    assert oracle.lib.oracle_bf16_to_f32(0x4049) == np.float32(3.140625)


def test_bfp16_block_format(oracle):
    """Block scale k = e - 15, e the frexp exponent of fl32(max * (1 + 2^-15))
    clamped to [-112, 127] (0 for an all-zero block); each mantissa is the
    nearest integer (ties to even) of v * 2^-k and never needs clamping;
    decode exact."""
    rng = np.random.default_rng(3)
    blocks = [rng.standard_normal(16).astype(np.float32) * s
              for s in (1.0, 1e-3, 1e4, 1e-30, 1e-34, 1e-36, 1e-40)]
    blocks.append(np.zeros(16, np.float32))
    blocks.append(np.array([0.5] + [0.0] * 15, np.float32))
    blocks.append(np.array([-1.0] * 16, np.float32))
    # maxima just below a power of two: rounding v * 2^(15-e) would give 2^15,
    # so the exponent moves up one
    blocks.append(np.array([np.nextafter(np.float32(1.0), np.float32(0.0))] + [0.25] * 15, np.float32))
    blocks.append(np.array([-np.float32(1.0 - 2.0 ** -17)] + [0.1] * 15, np.float32))
    # and just far enough below that it stays (v * 2^15 = 32766.75 -> |m| = 32767)
    blocks.append(np.array([np.float32(1.0 - 5 * 2.0 ** -17)] + [0.0] * 15, np.float32))
    x = np.concatenate(blocks)
    d, e = oracle.encode16("bfp16", x)
    back = oracle.decode16("bfp16", d, e)
    for b in range(len(blocks)):
        blk = x[16 * b:16 * b + 16].astype(np.float64)
        m = np.float32(np.max(np.abs(blk)))
        if m == 0:
            want_k = -15
        else:
            m2 = np.float32(m * np.float32(1.0 + 2.0 ** -15))  # one float32 rounding
            want_k = min(127, max(-112, int(np.frexp(np.float64(m2))[1]))) - 15
        assert int(e[b]) == want_k, b
        for k in range(16):
            t = Fraction(float(blk[k])) * Fraction(2) ** (-want_k)
            q = round(t)  # Python rounds half to even
            assert abs(q) <= 32767
            assert int(d[16 * b + k].astype(np.int16)) == q
            assert Fraction(float(back[16 * b + k])) == Fraction(q) * Fraction(2) ** want_k
        if m > 0:  # half a mantissa step
            assert np.max(np.abs(back[16 * b:16 * b + 16] - blk)) <= 2.0 ** (want_k - 1) * (1 + 1e-12)
    assert int(d[16 * (len(blocks) - 1)].astype(np.int16)) == 32767
    assert int(e[len(blocks) - 3]) == 1 - 15  # 1 - 2^-24 moved up to e = 1


def test_fp16_scaled_roundtrip(oracle):
    x = (np.random.default_rng(4).standard_normal(64) * 3e6).astype(np.float32)
    d, _ = oracle.encode16("fp16", x, scale=2.0**-8)
    back = oracle.decode16("fp16", d, None, scale=2.0**-8)
    assert np.array_equal(back, (x * np.float32(2.0**-8)).astype(np.float16).astype(np.float32) * np.float32(256.0))


@pytest.mark.parametrize("fmt", ["bf16", "fp16", "bfp16"])
def test_step16_is_decode_step_encode(oracle, fmt):
    """The 16-bit step equals: exact decode -> the pinned FP32 step -> encode
    of the interior (halo mantissas and halo-row exponents untouched)."""
    shape = (20, 24, 32)
    p0 = oracle.gaussian(0, 0.5, 0.5, 0.5, 3.0, shape=shape)
    p1 = p0 * np.float32(0.9)
    vel2 = oracle.layered_noise_vel2(shape, n_layers=3, seed=5)
    ps = 1.0
    vs = 2.0 ** -12 if fmt == "fp16" else 1.0
    cur, prev, v = oracle.encode16(fmt, p1, ps), oracle.encode16(fmt, p0, ps), oracle.encode16(fmt, vel2, vs)
    nxt = oracle.encode16(fmt, np.zeros(shape, np.float32), ps)
    d0 = nxt[0].copy()
    oracle.wave_step16(fmt, cur, prev, v, nxt, 10.0, 1e-3, ps, vs)
    fn = np.zeros(shape, np.float32)
    oracle.wave_step(oracle.decode16(fmt, *cur, scale=ps), oracle.decode16(fmt, *prev, scale=ps),
                     oracle.decode16(fmt, *v, scale=vs), fn, np.float32(10.0), np.float32(1e-3))
    got = oracle.decode16(fmt, *nxt, scale=ps)
    inner = (slice(R, -R),) * 3
    want_d, want_e = oracle.encode16(fmt, fn, ps)
    if fmt == "bfp16":
        # blocks re-encoded from the interior values only (halo cells count as 0)
        assert np.array_equal(got[inner], oracle.decode16(fmt, want_d, want_e)[inner])
    else:
        assert np.array_equal(nxt[0][inner], want_d[inner])
    mask = np.ones(shape, bool)
    mask[inner] = False
    assert np.array_equal(nxt[0][mask], d0[mask])
