"""Pins of the oracle's reference heightmap generator (glsl/heightmap.glsl
and the noise of glsl/simplex_noise.glsl:579-676).  CPU only.

The Hugo Elias hash against a vectorised int32 restatement; the gradient
noise against its own analytic derivative (central differences), its zeros
at lattice points and its continuity across cell edges (a wrong corner hash
breaks it); the two fBms against their octave sums; and one cell of the whole
generator, every branch on, against a plain-Python composition."""
from __future__ import annotations

import math

import numpy as np
import pytest

import pyoracle as O

f32 = lambda v: float(np.float32(v))  # noqa: E731


def hash_np(px, py):
    px = np.asarray(px, dtype=np.int64)
    py = np.asarray(py, dtype=np.int64)
    out = []
    for a, b in ((3, 311), (37, 113)):
        n = (px * a + py * b) & 0xFFFFFFFF
        n = ((n << 13) & 0xFFFFFFFF) ^ n
        n = (n * ((n * n * 15731 + 789221) & 0xFFFFFFFF) + 1376312589) & 0xFFFFFFFF
        lo = (n & 0x0FFFFFFF).astype(np.float32).astype(np.float64)
        out.append(f32(-1.0) + (2.0 * lo) / float(np.float32(0x0FFFFFFF)))
    return out


def test_elias_hash_matches_int32_restatement():
    rng = np.random.default_rng(1)
    pts = np.vstack([rng.integers(-200000, 200000, size=(300, 2)), [[0, 0], [1, 0], [0, 1], [-1, -1]]])
    hx, hy = hash_np(pts[:, 0], pts[:, 1])
    for k, (x, y) in enumerate(pts):
        gx, gy = O.elias_hash(int(x), int(y))
        assert gx == hx[k] and gy == hy[k], (x, y)
        assert -1.0 <= gx <= 1.0 and -1.0 <= gy <= 1.0


def test_noised_derivative_zeros_and_continuity():
    rng = np.random.default_rng(2)
    h = 1e-6
    for x, y in rng.uniform(-50, 50, size=(200, 2)):
        v, dx, dy = O.noised(x, y)
        fdx = (O.noised(x + h, y)[0] - O.noised(x - h, y)[0]) / (2 * h)
        fdy = (O.noised(x, y + h)[0] - O.noised(x, y - h)[0]) / (2 * h)
        assert dx == pytest.approx(fdx, abs=1e-6) and dy == pytest.approx(fdy, abs=1e-6)
    for i, j in [(0, 0), (7, -3), (-12, 40)]:
        assert O.noised(float(i), float(j))[0] == 0.0
    e = 1e-9
    for x0, y in rng.uniform(-30, 30, size=(50, 2)):
        k = math.floor(x0)                           # across the vertical edge x = k
        assert O.noised(k - e, y)[0] == pytest.approx(O.noised(k + e, y)[0], abs=1e-7)
        k = math.floor(y)
        assert O.noised(x0, k - e)[0] == pytest.approx(O.noised(x0, k + e)[0], abs=1e-7)


def fbm_py(x, y, m, damped):
    x += f32(m.seed) * 100.0
    y += f32(m.seed) * 100.0
    res, amp, freq, mx = 0.0, 1.0, (1.5 if damped else 1.0), 1.0
    sx = sy = 0.0
    for _ in range(m.octaves):
        v, dx, dy = O.noised(x * freq * f32(m.scale), y * freq * f32(m.scale))
        nv = (v + 1.0) / 2.0
        if damped:
            sx += dx
            sy += dy
            res += nv * amp / (1.0 + sx * sx + sy * sy)
        else:
            res += nv * amp
        freq *= f32(m.lacunarity)
        amp *= f32(m.persistance)
        mx += amp
    return res / mx


def cell_py(x, y, n, m):
    fx, fy = float(x), float(y)
    uvx, uvy = fx / n, fy / n
    dx = dy = 1.0
    if m.domain_warp:
        dx = fbm_py(fx + f32(2.3), fy + f32(2.9), m, False)
        dy = fbm_py(fx - f32(3.1), fy - f32(4.3), m, False)
        if m.domain_warp == 2:
            wx, wy = fx + f32(m.domain_warp_scale) * dx, fy + f32(m.domain_warp_scale) * dy
            dx = fbm_py(wx - f32(5.7), wy + f32(27.9), m, False)
            dy = fbm_py(wx + f32(11.5), wy - f32(23.7), m, False)
    v = fbm_py(fx + f32(m.domain_warp_scale) * dx, fy + f32(m.domain_warp_scale) * dy, m, True)
    hm = f32(m.height_mult)
    if m.uplift:
        sx, sy = fx - f32(7.3) + f32(m.seed) * 100.0, fy + f32(19.9) + f32(m.seed) * 100.0
        sc = f32(m.scale) / f32(m.uplift_scale)
        res, amp, freq, mx = 0.0, 1.0, 1.0, 1.0
        for _ in range(m.octaves):
            res += (1.0 - abs(O.ref_simplex(sx * freq * sc, sy * freq * sc))) * amp
            freq *= f32(m.lacunarity)
            amp *= f32(m.persistance)
            mx += amp
        hm += (2.5 if m.mask_exp else 0.0) + 1.0
        v *= res / mx
    if m.mask_round:
        hm += 16.0 if m.mask_exp else 0.0
        v *= max(0.0, 1.0 - ((uvx - 0.5) ** 2 + (uvy - 0.5) ** 2 + 0.75))
    if m.mask_exp:
        hm += 2.0
        v = v * (math.exp(v) - 1.0) / f32(1.718)
    if m.mask_power:
        hm += 1.25 + (2.0 if m.mask_exp else 0.0)
        v = v * (((v + 0.5) ** 3 - 0.125) / 3.25 * f32(0.55) + f32(0.45))
    if m.mask_slope:
        v = 0.25 * v + 0.75 * v * uvx * uvy
    v = v + 4 * m.mask_round * v
    if m.terrace > 0:
        step = 1.0 / (m.terrace * hm)
        v = f32(m.terrace_scale) * math.floor(v / step) * step + v * (1 - f32(m.terrace_scale))
    rock = min(f32(m.max_height), v * f32(m.max_height) * hm)
    dirt = (fbm_py(fx + f32(13.7), fy + f32(27.1), m, False) + 1.5) * f32(m.max_dirt)
    return rock, dirt


@pytest.mark.parametrize("kw", [{}, dict(domain_warp=2, uplift=1, mask_round=1, mask_slope=1, terrace=4,
                                         seed=17.25, scale=0.01, octaves=5)],
                         ids=["defaults", "every-branch"])
def test_generator_cells_match_python_composition(kw):
    m = O.map_settings(**kw)
    n = 48
    rock, dirt = O.gen_heightmap(n, m)
    for (x, y) in [(0, 0), (5, 17), (31, 2), (47, 47), (24, 24)]:
        r, d = cell_py(x, y, n, m)
        assert rock[y, x] == pytest.approx(r, rel=1e-9, abs=1e-9), (x, y)
        assert dirt[y, x] == pytest.approx(d, rel=1e-12), (x, y)
    assert np.isfinite(rock).all() and (rock <= m.max_height).all()


def test_generator_fp32_exp_within_two_ulp_of_exp():
    """The FP32 exp of the exp mask (the method the device kernel runs:
    Cody-Waite reduction, degree-7 Taylor polynomial, scaling by 2^k)
    against FP64 exp: within 2 ulp over [-30, 30] densely and over the
    normal range at random; exact at 0; overflow and underflow ends."""
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.linspace(-30, 30, 60001, dtype=np.float32),
                         rng.uniform(-86.9, 88.4, 20000).astype(np.float32),
                         np.float32([1e-8, -1e-8, 0.5, -0.5, 1.0, -1.0])])
    worst = 0.0
    for x in xs:
        got = float(np.float32(O.gen_exp_f32(float(x))))
        want = math.exp(float(x))
        worst = max(worst, abs(got - want) / float(np.spacing(np.float32(want))))
    assert worst <= 2.0, worst
    assert O.gen_exp_f32(0.0) == 1.0
    assert math.isinf(O.gen_exp_f32(89.0)) and O.gen_exp_f32(-90.0) == 0.0
    # a dropped polynomial term or a wrong ln2 split is off by far more
    assert abs(O.gen_exp_f32(0.34) - math.exp(0.34)) < 2 * float(np.spacing(np.float32(1.4)))


def test_generator_masks_and_terrace_closed_forms():
    """Each mask and the terrace against the plain noise height of the same
    cells (FP64): the round mask multiplies by 5 max(0, 1 - (|uv - 0.5|^2 +
    0.75)) (zero outside the disc), the exp mask maps v to v (e^v - 1) / 1.718
    and adds 2 to the height multiplier, the power mask maps v to
    v ((((v + 0.5)^3 - 0.125) / 3.25) 0.55 + 0.45) and adds 1.25, the slope
    mask maps v to v (0.25 + 0.75 uvx uvy), and a terrace with scale 1 puts
    every height on a multiple of max_height / terrace
    (glsl/heightmap.glsl:94-135)."""
    n = 48
    base = dict(mask_round=0, mask_exp=0, mask_power=0, mask_slope=0, uplift=0, terrace=0,
                max_height=4096.0, seed=3.5)
    rock0, _ = O.gen_heightmap(n, O.map_settings(**base), np.float64)
    maxh, hm = 4096.0, 1.0
    v = rock0 / (maxh * hm)                      # unclamped: max_height is far above the noise
    assert np.abs(v).max() < 0.9 and np.abs(v).max() > 0.05
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    ux, uy = x / n, y / n

    def run(**kw):
        return O.gen_heightmap(n, O.map_settings(**{**base, **kw}), np.float64)[0]

    w = 1.0 - ((ux - 0.5) ** 2 + (uy - 0.5) ** 2 + 0.75)
    got = run(mask_round=1)
    assert np.allclose(got, 5 * v * np.maximum(0.0, w) * maxh * hm, rtol=1e-12, atol=1e-9)
    assert (got[w <= 0] == 0).all() and (got[w > 0] != 0).mean() > 0.9
    got = run(mask_exp=1)
    assert np.allclose(got, v * (np.exp(v) - 1) / np.float32(1.718) * maxh * (hm + 2), rtol=1e-12, atol=1e-9)
    got = run(mask_power=1)
    b = v + 0.5
    c55, c45 = float(np.float32(0.55)), float(np.float32(0.45))   # the shader's FP32 literals
    assert np.allclose(got, v * (((b ** 3 - 0.125) / 3.25) * c55 + c45) * maxh * (hm + 1.25),
                       rtol=1e-12, atol=1e-9)
    got = run(mask_slope=1)
    assert np.allclose(got, v * (0.25 + 0.75 * ux * uy) * maxh * hm, rtol=1e-12, atol=1e-9)
    for steps in (3, 7):
        got = run(terrace=steps, terrace_scale=1.0)
        levels = got * steps / maxh
        assert np.allclose(levels, np.round(levels), atol=1e-9)
        assert np.array_equal(np.round(levels), np.floor(v * steps * hm))

