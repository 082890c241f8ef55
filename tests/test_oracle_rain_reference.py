"""Pins of the oracle's reference rain (glsl/rain.glsl:32-56): the simplex
noise (glsl/simplex_noise.glsl:377-402) against a brute-force lattice sum,
the fBm (:419-456) against its octave sum, and the rain increment and
mountain bonus against their closed forms.  CPU only."""
from __future__ import annotations

import math

import numpy as np
import pytest

import pyoracle as O

f32 = lambda v: float(np.float32(v))  # noqa: E731
CX, CY, CZ, CW = f32(0.211324865405187), f32(0.366025403784439), f32(-0.577350269189626), f32(0.024390243902439)


def rnd(q: int) -> int:
    return ((q * 34 + 1) * q) % 289


def simplex_brute(x: float, y: float) -> float:
    """Sum over every lattice vertex near (x, y) of the radial kernel
    max(0, 0.5 - r^2)^4 times the hashed gradient term, with no corner
    selection: only the three corners of the containing triangle reach."""
    s = x * CY + y * CY
    i0, j0 = math.floor(x + s), math.floor(y + s)
    total = 0.0
    for i in range(i0 - 2, i0 + 4):
        for j in range(j0 - 2, j0 + 4):
            dx = x - (i - (i + j) * CX)
            dy = y - (j - (i + j) * CX)
            t = 0.5 - dx * dx - dy * dy
            if t <= 0:
                continue
            p = rnd(rnd(j % 289) + i % 289)
            gx = 2.0 * ((p * CW) - math.floor(p * CW)) - 1.0
            h = abs(gx) - 0.5
            a0 = gx - math.floor(gx + 0.5)
            norm = f32(1.79284291400159) - f32(0.85373472095314) * (a0 * a0 + h * h)
            total += t ** 4 * norm * (a0 * dx + h * dy)
    return 130.0 * total


def test_reference_simplex_matches_brute_force():
    rng = np.random.default_rng(12)
    pts = np.vstack([rng.uniform(-500, 4000, size=(400, 2)), [[0.0, 0.0], [0.5, 0.25], [288.7, 289.3]]])
    for x, y in pts:
        a = O.ref_simplex(float(x), float(y))
        b = simplex_brute(float(x), float(y))
        assert abs(a - b) < 1e-11, (x, y, a, b)
        assert -1.2 < a < 1.2


def test_reference_simplex_fp32_close_to_fp64_and_zero_at_vertices():
    rng = np.random.default_rng(3)
    for x, y in rng.uniform(0, 300, size=(200, 2)):
        a64 = O.ref_simplex(float(np.float32(x)), float(np.float32(y)))
        a32 = O.ref_simplex(float(x), float(y), np.float32)
        assert abs(a64 - a32) < 2e-4
    for i, j in [(0, 0), (5, 9), (300, 17)]:
        vx, vy = i - (i + j) * CX, j - (i + j) * CX
        assert abs(O.ref_simplex(vx, vy)) < 1e-9


def test_reference_sfbm_is_normalised_octave_sum():
    for (x, y, seed, scale) in [(3.0, 7.0, 0.0, 0.02), (1000.0, 2000.0, 431.5, 0.02), (5.0, 5.0, 1.0, 0.3)]:
        vx, vy = x + seed * 100, y + seed * 100
        acc, amp, freq = 0.0, 1.0, 1.0
        for _ in range(8):
            acc += O.ref_simplex(vx * freq * scale, vy * freq * scale) * amp
            freq *= 2.0
            amp *= 0.5
        assert O.ref_sfbm(x, y, seed, scale, 8) == pytest.approx(acc / (2.0 - 2.0 ** -8), rel=1e-12, abs=1e-15)


def test_reference_rain_increment_and_mountain_bonus():
    n = 12
    rs = O.rain_settings(amount=0.01, mountain_thresh=0.55, mountain_multip=0.05, drops=0.02)
    max_h, time = 256.0, 12.75
    st = O.new_state(n)
    st[O.ROCK] = 10.0
    st[O.ROCK, 4, 7] = 240.0                                   # above 0.55 * 256
    st[O.WATER] = 0.5
    out = O.rain_reference(st.copy(), rs, time, max_h)
    seed = (f32(time) * f32(1.372914227e3)) % 1.0 * 1000.0
    for (y, x) in [(0, 0), (3, 5), (4, 7), (11, 11)]:
        r = max(0.0, O.ref_sfbm(float(x), float(y), seed, f32(0.02), 8))
        incr = f32(0.01) * r
        w = st[O.ROCK, y, x] + st[O.WATER, y, x]
        mountain = w - max_h * f32(0.55)
        if mountain > 0:
            incr += mountain * f32(0.05) * r / ((1 - f32(0.55)) * max_h)
        assert out[O.WATER, y, x] == pytest.approx(0.5 + incr, rel=1e-12, abs=1e-15), (y, x)
    assert (out[O.WATER] >= 0.5).all() and (out[O.WATER] > 0.5).any()
    for p in range(O.NUM_PLANES):
        if p != O.WATER:
            assert np.array_equal(out[p], st[p])
