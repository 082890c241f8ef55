"""Pins of the CPU oracle (CPU only, no GPU).

The oracle is the checker every GPU parity test compares against, so it is
itself checked here against what fixes the answer independently of its own
code: published known-answer vectors (Philox), closed forms worked by hand,
conservation laws and invariants of each pass, special cases that reduce to a
textbook formula, a brute-force restatement of the simplex noise, and an
array-form NumPy restatement of every GRID pass (tests/np_restatement.py).
Each check is chosen so that a plausible transcription slip (a dropped term,
a wrong sign, index or direction, a transposed operand) fails it.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

import np_restatement as NP
import pyoracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
C0 = dict(Kc=0.05, Ks=0.02, Kd=0.02, Ke=0.01, talus_deg=35.0, slippage=0.1, G=9.81, dt=0.005)
f32 = lambda v: float(np.float32(v))  # noqa: E731


def settings(**kw):
    s = O.config_settings(**{**C0, **kw})
    return s, NP.settings_dict(s)


def one_pass(st, s, p, **kw):
    return O.run_pass(st.copy(), s, p, **kw)


# ---------------------------------------------------------------------------
# Philox4x32-10 and the rain counter stream
# ---------------------------------------------------------------------------
def test_philox_known_answers():
    with open(os.path.join(GOLDEN, "philox4x32_10_kat.json")) as f:
        kat = json.load(f)
    for v in kat["vectors"]:
        ctr = [int(x, 16) for x in v["ctr"]]
        key = [int(x, 16) for x in v["key"]]
        assert O.philox(ctr, key) == [int(x, 16) for x in v["out"]]


def test_rain_uniform_counter_layout_and_moments():
    seed = 0x1234_5678_9abc
    # u(x, y) is word (x & 3) of Philox(ctr = {x >> 2, y, step, terrain}, key = seed) >> 8, / 2^24
    for (x, y, step, terr) in [(0, 0, 0, 0), (5, 3, 7, 1), (4093, 4095, 1999, 63)]:
        o = O.philox([x >> 2, y, step, terr], [seed & 0xffffffff, seed >> 32])
        assert O.rain_u(seed, step, terr, x, y) == (o[x & 3] >> 8) / 2.0 ** 24
    u = np.array([O.rain_u(7, 3, 0, x, y) for y in range(48) for x in range(48)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01
    # a rain pass adds amount * u exactly (FP64), and a constant rain adds the amount
    st = O.new_state(16, ny=8)
    st[O.WATER] = 0.25
    out = one_pass(st, O.config_settings(), O.PASS_RAIN, rain_mode=O.RAIN_UNIFORM, rain_amount=0.02,
                   seed=7, step=3)
    amt = f32(0.02)
    want = np.array([[0.25 + amt * O.rain_u(7, 3, 0, x, y) for x in range(16)] for y in range(8)])
    assert np.array_equal(out[O.WATER], want)
    out = one_pass(st, O.config_settings(), O.PASS_RAIN, rain_mode=O.RAIN_CONSTANT, rain_amount=0.01)
    assert np.array_equal(out[O.WATER], np.full((8, 16), 0.25 + f32(0.01)))


# ---------------------------------------------------------------------------
# fBm init: permutation, simplex noise, min-max rescale
# ---------------------------------------------------------------------------
def test_fbm_permutation_is_a_seeded_permutation():
    a, b = O.fbm_permutation(1), O.fbm_permutation(2)
    assert np.array_equal(np.sort(a), np.arange(256))
    assert np.array_equal(np.sort(b), np.arange(256))
    assert not np.array_equal(a, b)
    assert np.array_equal(a, O.fbm_permutation(1))


def simplex_brute(perm: np.ndarray, x: float, y: float) -> float:
    """2-D simplex noise as a plain sum over every lattice vertex near the
    point of the radial kernel (0.5 - r^2)^4 * (gradient . offset), without the
    simplex-corner selection: only the 3 corners of the containing triangle
    have r^2 < 0.5, so the sums agree."""
    F2 = (math.sqrt(3.0) - 1) / 2
    G2 = (3 - math.sqrt(3.0)) / 6
    grads = [(1, 1), (-1, 1), (1, -1), (-1, -1), (1, 0), (-1, 0), (0, 1), (0, -1)]
    s = (x + y) * F2
    i0, j0 = math.floor(x + s), math.floor(y + s)
    total = 0.0
    for i in range(i0 - 2, i0 + 4):
        for j in range(j0 - 2, j0 + 4):
            vx = i - (i + j) * G2
            vy = j - (i + j) * G2
            dx, dy = x - vx, y - vy
            t = 0.5 - dx * dx - dy * dy
            if t <= 0:
                continue
            h = int(perm[(i + int(perm[j & 255])) & 255]) & 7
            gx, gy = grads[h]
            total += t ** 4 * (gx * dx + gy * dy)
    return 70.0 * total


def test_simplex_matches_brute_force_lattice_sum():
    rng = np.random.default_rng(5)
    for seed in (1, 99):
        perm = O.fbm_permutation(seed)
        pts = rng.uniform(-300, 300, size=(300, 2))
        pts = np.vstack([pts, [[0.0, 0.0], [1.0, 0.0], [0.5, 0.5], [1e-9, 2.0]]])
        for x, y in pts:
            a = O.simplex2(seed, float(x), float(y))
            b = simplex_brute(perm, float(x), float(y))
            assert abs(a - b) < 1e-12, (seed, x, y, a, b)
        # noise vanishes at every lattice vertex (zero offset to the nearest corner)
        for i, j in [(0, 0), (3, -2), (17, 40)]:
            vx, vy = i - (i + j) * (3 - math.sqrt(3.0)) / 6, j - (i + j) * (3 - math.sqrt(3.0)) / 6
            assert abs(O.simplex2(seed, vx, vy)) < 1e-9


def test_fbm_is_octave_sum_rescaled_to_range():
    n, ny = 12, 9
    p = O.fbm_params(seed=3, octaves=4, lacunarity=2.0, gain=0.5, base_freq=1 / 8, lo=-2.0, hi=30.0)
    got = O.fbm_init(n, p, np.float64, ny=ny)
    raw = np.zeros((ny, n))
    for y in range(ny):
        for x in range(n):
            amp, f = 1.0, f32(1 / 8)
            for _ in range(4):
                raw[y, x] += amp * O.simplex2(3, x * f, y * f)
                amp *= 0.5
                f *= 2.0
    want = -2.0 + 32.0 * (raw - raw.min()) / (raw.max() - raw.min())
    assert np.max(np.abs(got - want)) < 1e-9
    assert got.min() == -2.0 and got.max() == 30.0
    g32 = O.fbm_init(n, p, np.float32, ny=ny)
    assert g32.min() == np.float32(-2.0) and g32.max() == np.float32(30.0)
    assert np.max(np.abs(g32 - got)) < 1e-4


# ---------------------------------------------------------------------------
# hydro_flux
# ---------------------------------------------------------------------------
def test_flux_flat_water_is_at_rest():
    s, _ = settings()
    st = O.new_state(10, ny=7)
    st[O.ROCK] = 3.0
    st[O.WATER] = 0.5
    out = one_pass(st, s, O.PASS_FLUX)
    for p in (O.FL, O.FR, O.FT, O.FB, O.VX, O.VY):
        assert not out[p].any()
    assert np.array_equal(out[O.WATER], st[O.WATER])
    assert np.array_equal(out[O.VZ], np.full((7, 10), 1.0))


def test_flux_single_column_by_hand():
    s, k = settings()
    dt, G = k["dt"], k["G"]
    st = O.new_state(5)
    st[O.WATER, 2, 2] = 1.0
    out = one_pass(st, s, O.PASS_FLUX)
    f = dt * G * 1.0                          # dt * A * G * dw / L, old flux 0
    K = min(1.0, 1.0 / (4 * f * dt))
    for p in (O.FL, O.FR, O.FT, O.FB):
        assert out[p, 2, 2] == pytest.approx(f * K, rel=1e-15)
        assert np.count_nonzero(out[p]) == 1
    assert out[O.WATER, 2, 2] == pytest.approx(1.0 - dt * 4 * f * K, rel=1e-15)
    # inflow and velocity use the OLD fluxes (zero here): neighbours stay dry
    assert np.count_nonzero(out[O.WATER]) == 1
    assert not out[O.VX].any() and not out[O.VY].any()
    # second pass: the neighbours now receive the first pass's outflow
    out2 = one_pass(out, s, O.PASS_FLUX)
    for (y, x) in [(2, 1), (2, 3), (1, 2), (3, 2)]:
        assert out2[O.WATER, y, x] == pytest.approx(dt * f * K, rel=1e-12)
    # velocity of the centre from the old fluxes: symmetric -> zero
    assert abs(out2[O.VX, 2, 2]) < 1e-15 and abs(out2[O.VY, 2, 2]) < 1e-15
    # left neighbour: old R of its left (0) - its old L (0) + its old R (0) - old L of the centre (f)
    vz = out[O.WATER, 2, 1] + out2[O.WATER, 2, 1]
    assert out2[O.VX, 2, 1] == pytest.approx(-f * K / vz, rel=1e-12)


def test_flux_velocity_uses_old_fluxes_by_hand():
    """glsl/hydro_flux.glsl:146-158: the velocity reads get_flux() of the pass's
    input image, so both the cell's own outflow and its neighbours' are the OLD
    fluxes.  Asymmetric, non-zero old fluxes whose new values differ make the
    old-flux and new-flux formulas disagree, and only the old one may match."""
    s, k = settings()
    dt, G, E = k["dt"], k["G"], k["E"]
    st = O.new_state(5)
    st[O.ROCK] = np.arange(25, dtype=np.float64).reshape(5, 5) * 0.01   # a slope, no symmetry
    st[O.WATER, 1:4, 1:4] = [[0.30, 0.20, 0.10], [0.25, 0.50, 0.05], [0.15, 0.35, 0.45]]
    old = {O.FL: 0.7, O.FR: 0.3, O.FT: 0.9, O.FB: 0.2}
    rng = np.random.default_rng(5)
    for p, v in old.items():
        st[p, 1:4, 1:4] = v * (1.0 + rng.uniform(-0.5, 0.5, (3, 3)))
    out = one_pass(st, s, O.PASS_FLUX)
    y, x = 2, 2
    F = lambda p, yy, xx: st[p, yy, xx]        # noqa: E731  old flux
    N = lambda p, yy, xx: out[p, yy, xx]       # noqa: E731  new flux
    # the new fluxes of the centre, by hand: E*f + dt*G*dw, clamped at 0, then K
    w = st[O.ROCK] + st[O.DIRT] + st[O.WATER]
    dw = [w[y, x] - w[y, x - 1], w[y, x] - w[y, x + 1], w[y, x] - w[y + 1, x], w[y, x] - w[y - 1, x]]
    f = [max(0.0, E * F(p, y, x) + dt * G * d) for p, d in zip((O.FL, O.FR, O.FT, O.FB), dw)]
    K = min(1.0, st[O.WATER, y, x] / (sum(f) * dt))
    for p, v in zip((O.FL, O.FR, O.FT, O.FB), f):
        assert N(p, y, x) == pytest.approx(v * K, rel=1e-12)
    sum_in = F(O.FR, y, x - 1) + F(O.FL, y, x + 1) + F(O.FB, y + 1, x) + F(O.FT, y - 1, x)
    d1 = st[O.WATER, y, x]
    d2 = max(0.0, d1 + dt * (sum_in - K * sum(f)))
    vz = d1 + d2
    assert out[O.VZ, y, x] == pytest.approx(vz, rel=1e-12)
    vx_old = (F(O.FR, y, x - 1) - F(O.FL, y, x) + F(O.FR, y, x) - F(O.FL, y, x + 1)) / vz
    vy_old = (F(O.FT, y - 1, x) - F(O.FB, y, x) + F(O.FT, y, x) - F(O.FB, y + 1, x)) / vz
    assert out[O.VX, y, x] == pytest.approx(vx_old, rel=1e-12)
    assert out[O.VY, y, x] == pytest.approx(vy_old, rel=1e-12)
    # the plausible slips: the cell's own outflow taken from the new fluxes,
    # or every flux taken from the new fluxes
    vx_own_new = (F(O.FR, y, x - 1) - N(O.FL, y, x) + N(O.FR, y, x) - F(O.FL, y, x + 1)) / vz
    vy_own_new = (F(O.FT, y - 1, x) - N(O.FB, y, x) + N(O.FT, y, x) - F(O.FB, y + 1, x)) / vz
    vx_all_new = (N(O.FR, y, x - 1) - N(O.FL, y, x) + N(O.FR, y, x) - N(O.FL, y, x + 1)) / vz
    for wrong_x, wrong_y in [(vx_own_new, vy_own_new), (vx_all_new, vy_own_new)]:
        assert abs(wrong_x - vx_old) > 1e-3 * abs(vx_old) + 1e-6
        assert abs(wrong_y - vy_old) > 1e-3 * abs(vy_old) + 1e-6


def random_hydro_state(nx, ny, seed):
    rng = np.random.default_rng(seed)
    st = O.new_state(nx, ny=ny)
    st[O.ROCK] = rng.uniform(0, 3, (ny, nx))
    st[O.DIRT] = rng.uniform(0, 0.2, (ny, nx)) * (rng.random((ny, nx)) < 0.7)
    st[O.WATER] = rng.uniform(0, 0.3, (ny, nx)) * (rng.random((ny, nx)) < 0.8)
    for p in (O.FL, O.FR, O.FT, O.FB):
        st[p] = rng.uniform(0, 2, (ny, nx)) * (rng.random((ny, nx)) < 0.5)
    st[O.S0] = rng.uniform(0, 0.01, (ny, nx))
    st[O.S1] = rng.uniform(0, 0.01, (ny, nx))
    st[O.VX] = rng.normal(0, 30, (ny, nx))
    st[O.VY] = rng.normal(0, 30, (ny, nx))
    st[O.VZ] = rng.uniform(0, 2e-3, (ny, nx)) * (rng.random((ny, nx)) < 0.6) + \
        rng.uniform(0, 0.5, (ny, nx)) * (rng.random((ny, nx)) < 0.5)
    return st


def test_flux_bounds_walls_and_water_identity():
    s, k = settings()
    st = random_hydro_state(13, 9, 1)
    out = one_pass(st, s, O.PASS_FLUX)
    f = out[[O.FL, O.FR, O.FT, O.FB]]
    assert (f >= 0).all()
    # K scaling: a cell never sends more than its water
    assert (f.sum(0) * k["dt"] <= st[O.WATER] * (1 + 1e-12) + 1e-300).all()
    # closed walls
    assert not out[O.FL, :, 0].any() and not out[O.FR, :, -1].any()
    assert not out[O.FB, 0, :].any() and not out[O.FT, -1, :].any()
    # water identity with the OLD inflow: d2 = max(0, b + dt (sum_in_old - sum_out_new))
    inflow = np.zeros_like(st[O.WATER])
    inflow[:, 1:] += st[O.FR, :, :-1]
    inflow[:, :-1] += st[O.FL, :, 1:]
    inflow[:-1, :] += st[O.FB, 1:, :]
    inflow[1:, :] += st[O.FT, :-1, :]
    want = np.maximum(0, st[O.WATER] + k["dt"] * (inflow - f.sum(0)))
    assert np.allclose(out[O.WATER], want, rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------------------
# hydro_erosion
# ---------------------------------------------------------------------------
def test_erosion_conserves_material_per_cell():
    s, _ = settings(Kc=0.5, Ks=0.3, Kd=0.2)
    st = random_hydro_state(12, 10, 2)
    out = one_pass(st, s, O.PASS_EROSION)
    tot = lambda a: a[O.ROCK] + a[O.DIRT] + a[O.S0] + a[O.S1]  # noqa: E731
    assert np.allclose(tot(out), tot(st), rtol=0, atol=1e-14)
    assert (out[O.ROCK] >= 0).all() and (out[O.DIRT] >= 0).all()
    assert np.array_equal(out[O.WATER], st[O.WATER])


def test_erosion_sin_slope_closed_form_and_break():
    # deep dirt everywhere: the dirt layer takes the whole capacity and the loop
    # breaks before the rock layer (rock and s0 untouched, Kconv 0)
    s, k = settings(Kc=0.3, Ks=0.2, Kconv=0.0)
    rng = np.random.default_rng(3)
    n = 7
    st = O.new_state(n)
    st[O.ROCK] = rng.uniform(0, 4, (n, n))
    st[O.DIRT] = 10.0
    st[O.VX] = 3.0
    st[O.VY] = -4.0
    st[O.VZ] = 1.0
    out = one_pass(st, s, O.PASS_EROSION)
    assert np.array_equal(out[O.ROCK], st[O.ROCK]) and not out[O.S0].any()
    h = st[O.ROCK] + st[O.DIRT]
    hp = np.pad(h, 1)                      # out-of-grid neighbours read 0
    dx = hp[1:-1, 2:] - hp[1:-1, :-2]
    dz = hp[2:, 1:-1] - hp[:-2, 1:-1]
    sin_a = np.sqrt((dx ** 2 + dz ** 2) / (dx ** 2 + dz ** 2 + 4.0))
    cap = k["Kc"] * np.maximum(f32(0.02), sin_a) * 5.0
    assert np.allclose(out[O.S1], k["dt"] * k["Ks"][1] * cap, rtol=1e-12, atol=0)
    assert np.allclose(out[O.DIRT], 10.0 - out[O.S1], rtol=0, atol=1e-14)


def test_erosion_still_water_deposits_only_and_exhaustion():
    s, k = settings(Kc=0.3, Ks=0.5, Kd=0.25, Kconv=0.0)
    n = 5
    st = O.new_state(n)
    st[O.ROCK] = 1.0
    st[O.S0] = 0.04
    st[O.S1] = 0.08
    st[O.VZ] = 1.0                           # velocity 0 -> capacity 0
    out = one_pass(st, s, O.PASS_EROSION)
    assert np.allclose(out[O.DIRT], k["dt"] * k["Kd"][1] * 0.08, rtol=1e-14)
    assert np.allclose(out[O.ROCK], 1.0 + k["dt"] * k["Kd"][0] * 0.04, rtol=1e-14)
    # thin dirt, strong flow: dirt is exhausted, its volume is carried over to
    # reduce the rock layer's capacity
    st = O.new_state(n)
    st[O.ROCK] = 1.0
    st[O.DIRT] = 1e-6
    st[O.VX] = 400.0
    st[O.VZ] = 1.0
    out = one_pass(st, s, O.PASS_EROSION)
    c = k["Kc"] * f32(0.02) * 400.0        # flat: sin_a below 0.02 in the interior
    inner = (slice(1, -1), slice(1, -1))
    assert not out[O.DIRT][inner].any()
    assert np.allclose(out[O.S1][inner], 1e-6, rtol=1e-12)
    assert np.allclose(out[O.S0][inner], k["dt"] * k["Ks"][0] * (c - 1e-6), rtol=1e-12)


def test_erosion_smoothstep_ramp_for_thin_water():
    s, k = settings(Kc=0.3, Ks=0.2, Kconv=0.0)
    n = 5
    for vz, factor in [(1e-3, 1.0), (5e-4, 0.0), (1e-4, 0.0), (7.5e-4, 0.5)]:
        st = O.new_state(n)
        st[O.ROCK] = 1.0
        st[O.DIRT] = 1.0
        st[O.VX] = 100.0
        st[O.VZ] = vz
        out = one_pass(st, s, O.PASS_EROSION)
        want = k["dt"] * k["Ks"][1] * k["Kc"] * f32(0.02) * 100.0 * factor
        assert out[O.S1, 2, 2] == pytest.approx(want, rel=1e-6, abs=1e-18), vz


# ---------------------------------------------------------------------------
# sediment_transport
# ---------------------------------------------------------------------------
def test_transport_identity_shift_constant_evaporation():
    s, k = settings()
    rng = np.random.default_rng(4)
    nx, ny = 11, 8
    st = O.new_state(nx, ny=ny)
    st[O.S0] = rng.random((ny, nx))
    st[O.S1] = rng.random((ny, nx))
    st[O.WATER] = rng.random((ny, nx))
    out = one_pass(st, s, O.PASS_TRANSPORT)               # zero velocity
    assert np.array_equal(out[O.S0], st[O.S0]) and np.array_equal(out[O.S1], st[O.S1])
    assert np.allclose(out[O.WATER], st[O.WATER] * (1 - k["Ke"] * k["dt"]), rtol=1e-15)
    # velocity of exactly one cell per step to +x: an exact shift, clamped at x = 0
    st2 = st.copy()
    st2[O.VX] = 1.0 / k["dt"]
    out = one_pass(st2, s, O.PASS_TRANSPORT)
    assert np.array_equal(out[O.S0][:, 1:], st[O.S0][:, :-1])
    assert np.array_equal(out[O.S0][:, 0], st[O.S0][:, 0])
    st2[O.VX] = 0.0
    st2[O.VY] = -2.0 / k["dt"]                            # two cells to -y
    out = one_pass(st2, s, O.PASS_TRANSPORT)
    assert np.array_equal(out[O.S1][:-2, :], st[O.S1][2:, :])
    # a constant field stays constant under any velocity
    st3 = O.new_state(nx, ny=ny)
    st3[O.S0] = 0.375
    st3[O.VX] = rng.normal(0, 300, (ny, nx))
    st3[O.VY] = rng.normal(0, 300, (ny, nx))
    out = one_pass(st3, s, O.PASS_TRANSPORT)
    assert np.allclose(out[O.S0], 0.375, rtol=1e-15)
    # n passes of evaporation
    w = st[O.WATER].copy()
    cur = st.copy()
    for _ in range(5):
        cur = one_pass(cur, s, O.PASS_TRANSPORT)
    assert np.allclose(cur[O.WATER], w * (1 - k["Ke"] * k["dt"]) ** 5, rtol=1e-14)


# ---------------------------------------------------------------------------
# thermal erosion + transport
# ---------------------------------------------------------------------------
def thermal_pair(st, s, layer):
    p = O.PASS_TFLUX0 if layer == 0 else O.PASS_TFLUX1
    return one_pass(one_pass(st, s, p), s, p + 1)


def test_thermal_conserves_each_layer():
    s, _ = settings(slippage=3.0)
    rng = np.random.default_rng(6)
    nx, ny = 14, 11
    st = O.new_state(nx, ny=ny)
    st[O.ROCK] = rng.uniform(0, 6, (ny, nx))
    st[O.DIRT] = rng.uniform(0, 2, (ny, nx))
    a = thermal_pair(st, s, 0)
    assert a[O.ROCK].sum() == pytest.approx(st[O.ROCK].sum(), rel=1e-14)
    assert np.array_equal(a[O.DIRT], st[O.DIRT])
    assert not np.array_equal(a[O.ROCK], st[O.ROCK])
    b = thermal_pair(a, s, 1)
    assert b[O.DIRT].sum() == pytest.approx(a[O.DIRT].sum(), rel=1e-14)
    assert np.array_equal(b[O.ROCK], a[O.ROCK])
    assert (b[O.DIRT] >= -1e-12).all()


def test_thermal_below_talus_is_still():
    s, _ = settings()                                     # talus 35 deg, tan = 0.700
    nx, ny = 9, 9
    st = O.new_state(nx, ny=ny)
    x = np.arange(nx)[None, :]
    y = np.arange(ny)[:, None]
    st[O.ROCK] = 20 + 0.3 * x + 0.2 * y                   # steepest diagonal 0.5/sqrt2
    out = thermal_pair(st, s, 0)
    assert np.array_equal(out[O.ROCK][1:-1, 1:-1], st[O.ROCK][1:-1, 1:-1])
    for p in O.THERM:
        assert not out[p][1:-1, 1:-1].any()


@pytest.mark.parametrize("layer", [0, 1])
def test_thermal_single_spike_analytic_split(layer):
    s, k = settings(slippage=0.7)
    n = 7
    h = 2.0
    st = O.new_state(n)
    st[O.ROCK] = 10.0
    if layer == 0:
        st[O.ROCK, 3, 3] += h
    else:
        st[O.DIRT] = 0.0
        st[O.DIRT, 3, 3] = h
    out = thermal_pair(st, s, layer)
    Ka = k["Kalpha"][layer]
    sharp = max(1.0, 1.0 + math.atan(h) - Ka)              # all 8 marked; cross is steepest
    assert math.atan(h / f32(math.sqrt(2))) > Ka
    S = k["dt"] * k["Kspeed"][layer] * sharp ** 3 * h / 2
    plane = O.ROCK if layer == 0 else O.DIRT
    base = 10.0 if layer == 0 else 0.0
    assert out[plane, 3, 3] == pytest.approx(base + h - S, rel=1e-13)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dx or dy:
                assert out[plane, 3 + dy, 3 + dx] == pytest.approx(base + S / 8, rel=1e-13, abs=1e-15)


# ---------------------------------------------------------------------------
# smoothing
# ---------------------------------------------------------------------------
def test_smoothing_ramp_spike_border():
    s, k = settings(slippage=40.0)          # m = clamp(40 * 0.005) = 0.2
    m = min(1.0, f32(40.0) * f32(0.005))
    nx, ny = 8, 7
    st = O.new_state(nx, ny=ny)
    st[O.ROCK] = 1.0 + 0.5 * np.arange(nx)[None, :] - 0.25 * np.arange(ny)[:, None]
    st[O.DIRT] = 0.1
    out = one_pass(st, s, O.PASS_SMOOTH)
    assert np.allclose(out[O.ROCK], st[O.ROCK], rtol=1e-15) and np.allclose(out[O.DIRT], 0.1, rtol=1e-15)
    # a pit whose left neighbour towers over it: -d_l = 10 > |mean d| = 3.25 and
    # d_l * d_r > 0, so it takes the 5-point mean, blended by m
    st = O.new_state(nx, ny=ny)
    st[O.ROCK] = 2.0
    st[O.ROCK, 3, 4] = 1.0
    st[O.ROCK, 3, 3] = 11.0
    st[O.ROCK, 0, 2] = 9.0                  # on the border: copied as is
    st[O.DIRT] = 0.1
    out = one_pass(st, s, O.PASS_SMOOTH)
    mean = (1.0 + 11.0 + 2.0 + 2.0 + 2.0) / 5
    assert out[O.ROCK, 3, 4] == pytest.approx(m * mean + (1 - m) * 1.0, rel=1e-14)
    assert out[O.ROCK, 0, 2] == 9.0
    changed = np.argwhere(out[O.ROCK] != st[O.ROCK])
    assert changed.tolist() == [[3, 4]]     # the peak and every other cell stay
    assert np.allclose(out[O.DIRT], 0.1, rtol=1e-15)
    # strict tests: a symmetric pit (-d = |mean d| exactly) and a spike stay
    st = O.new_state(nx, ny=ny)
    st[O.ROCK] = 2.0
    st[O.ROCK, 3, 4] = 1.0
    st[O.ROCK, 2, 2] = 6.0
    out = one_pass(st, s, O.PASS_SMOOTH)
    assert np.array_equal(out[O.ROCK], st[O.ROCK])
    assert not O.smooth_flags(st).any()


def test_smoothing_flags_entry_follows_the_pass():
    """orc_smooth_flags reports the extremum tests that pass_smooth applies:
    the hand-built pit alone, and on a rough random state exactly the cells
    that the pass moves to the 5-point mean (FP64, same blend)."""
    s, k = settings(slippage=40.0)
    m = min(1.0, f32(40.0) * f32(0.005))
    nx, ny = 8, 7
    st = O.new_state(nx, ny=ny)
    st[O.ROCK] = 2.0
    st[O.ROCK, 3, 4] = 1.0
    st[O.ROCK, 3, 3] = 11.0
    st[O.DIRT] = 0.1
    f = O.smooth_flags(st)
    assert np.argwhere(f & 1).tolist() == [[3, 4]]
    # surface = rock + dirt with constant dirt: the same single pit
    assert np.argwhere(f & 2).tolist() == [[3, 4]]
    rng = np.random.default_rng(5)
    nx = ny = 24
    st = O.new_state(nx, ny=ny)
    st[O.ROCK] = rng.uniform(0, 5, (ny, nx))
    st[O.DIRT] = rng.uniform(0, 1.5, (ny, nx))
    f = O.smooth_flags(st)
    out = one_pass(st, s, O.PASS_SMOOTH)
    for plane, bit in ((O.ROCK, 1), (O.DIRT, 2)):
        x = st[plane]
        mean = (x[1:-1, 1:-1] + x[1:-1, :-2] + x[1:-1, 2:] + x[2:, 1:-1] + x[:-2, 1:-1]) / 5
        flagged = (f[1:-1, 1:-1] & bit) != 0
        want = np.where(flagged, m * mean + (1 - m) * x[1:-1, 1:-1], m * x[1:-1, 1:-1] + (1 - m) * x[1:-1, 1:-1])
        assert 0.1 < flagged.mean() < 0.9
        assert np.allclose(out[plane][1:-1, 1:-1], want, rtol=1e-14, atol=0)
    assert not f[0].any() and not f[-1].any() and not f[:, 0].any() and not f[:, -1].any()


# ---------------------------------------------------------------------------
# every pass and whole steps against the array-form restatement
# ---------------------------------------------------------------------------
def rich_state(nx, ny, seed):
    st = random_hydro_state(nx, ny, seed)
    rng = np.random.default_rng(seed + 100)
    st[O.ROCK] = rng.uniform(0, 5, (ny, nx))              # steep: thermal marks many directions
    st[O.DIRT] = rng.uniform(0, 1.5, (ny, nx)) * (rng.random((ny, nx)) < 0.8)
    for p in O.THERM:
        st[p] = rng.uniform(0, 0.05, (ny, nx)) * (rng.random((ny, nx)) < 0.5)
    return st


PASSES = [
    (O.PASS_FLUX, lambda S, k: NP.flux(S, k)),
    (O.PASS_EROSION, lambda S, k: NP.erosion(S, k)),
    (O.PASS_TRANSPORT, lambda S, k: NP.transport(S, k)),
    (O.PASS_TFLUX0, lambda S, k: NP.thermal_flux(S, k, 0)),
    (O.PASS_TTRANS0, lambda S, k: NP.thermal_transport(S, 0)),
    (O.PASS_TFLUX1, lambda S, k: NP.thermal_flux(S, k, 1)),
    (O.PASS_TTRANS1, lambda S, k: NP.thermal_transport(S, 1)),
    (O.PASS_SMOOTH, lambda S, k: NP.smooth(S, k)),
]


@pytest.mark.parametrize("pass_id,fn", PASSES, ids=[str(p) for p, _ in PASSES])
@pytest.mark.parametrize("defaults", [False, True], ids=["config", "reference-defaults"])
def test_pass_matches_array_restatement(pass_id, fn, defaults):
    s = O.reference_grid_defaults() if defaults else O.config_settings(**{**C0, "slippage": 2.0})
    k = NP.settings_dict(s)
    st = rich_state(23, 17, 10 + pass_id)                 # non-square: catches x/y swaps
    got = one_pass(st, s, pass_id)
    want = fn(st, k)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14), \
        [p for p in range(O.NUM_PLANES) if not np.allclose(got[p], want[p], rtol=1e-12, atol=1e-14)]


def test_grid_steps_match_array_restatement():
    s = O.config_settings(**{**C0, "slippage": 1.0})
    k = NP.settings_dict(s)
    st = rich_state(20, 15, 7)
    st[O.VX] = st[O.VY] = st[O.VZ] = 0.0
    got = st.copy()
    want = st.copy()
    # a few steps only: the model has genuine discontinuities (the talus mark,
    # the dirt-exhaustion branch of erosion, the smoothing extremum test), so
    # rounding-level differences between the two can flip a branch later on
    for i in range(6):
        O.run(got, s, 1, O.STEP_THERMAL | O.STEP_SMOOTH, step0=i)
        want = NP.grid_step(want, k)
    for p in range(9):
        assert np.allclose(got[p], want[p], rtol=1e-12, atol=1e-14), p


def test_golden_regression_fixture():
    """A 16 x 12 C0-coefficient run recorded by tests/golden/make_fixtures.py:
    the oracle must keep reproducing it (a change to the oracle shows here)."""
    z = np.load(os.path.join(GOLDEN, "grid16x12_c0.npz"))
    st = z["initial"].copy()
    O.run(st, O.config_settings(**C0), int(z["steps"]), O.STEP_DEFAULT, O.RAIN_UNIFORM, 0.02,
          seed=int(z["seed"]))
    assert np.allclose(st[:9], z["final"], rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------------------
# droplets
# ---------------------------------------------------------------------------
def droplets_python(rock, dirt, n, ds, iteration):
    """Serial droplet model in plain Python floats (brute-force restatement of
    the oracle's documented model)."""
    taps = []
    r = ds.brush_radius
    for by in range(-r, r + 1):
        for bx in range(-r, r + 1):
            d = math.sqrt(bx * bx + by * by)
            if d < r:
                taps.append((bx, by, 1 - d / r))
    tot = sum(w for _, _, w in taps)
    taps = [(bx, by, w / tot) for bx, by, w in taps]
    Hh = lambda x, y: rock[y, x] + dirt[y, x]  # noqa: E731
    cell = lambda q: min(int(q), n - 2)  # noqa: E731

    def bil(px, py):
        x, y = cell(px), cell(py)
        u, v = px - x, py - y
        return (Hh(x, y) * (1 - u) * (1 - v) + Hh(x + 1, y) * u * (1 - v) + Hh(x, y + 1) * (1 - u) * v +
                Hh(x + 1, y + 1) * u * v)

    def deposit(px, py, a):
        x, y = cell(px), cell(py)
        u, v = px - x, py - y
        dirt[y, x] += a * (1 - u) * (1 - v)
        dirt[y, x + 1] += a * u * (1 - v)
        dirt[y + 1, x] += a * (1 - u) * v
        dirt[y + 1, x + 1] += a * u * v

    exported = 0.0
    for i in range(ds.count):
        o = O.philox([i, iteration, 0x64726f70, 0], [ds.seed & 0xffffffff, ds.seed >> 32])
        px = (o[0] >> 8) / 2 ** 24 * (n - 1)
        py = (o[1] >> 8) / 2 ** 24 * (n - 1)
        dx = dy = sed = 0.0
        speed, water = f32(ds.init_speed), f32(ds.init_water)
        alive = True
        inertia = f32(ds.inertia)
        for _ in range(ds.max_lifetime):
            cx, cy = cell(px), cell(py)
            u, v = px - cx, py - cy
            h00, h10, h01, h11 = Hh(cx, cy), Hh(cx + 1, cy), Hh(cx, cy + 1), Hh(cx + 1, cy + 1)
            gx = (h10 - h00) * (1 - v) + (h11 - h01) * v
            gy = (h01 - h00) * (1 - u) + (h11 - h10) * u
            h = bil(px, py)
            dx = dx * inertia - gx * (1 - inertia)
            dy = dy * inertia - gy * (1 - inertia)
            ln = math.sqrt(dx * dx + dy * dy)
            if ln > 0:
                dx, dy = dx / ln, dy / ln
            ox, oy = px, py
            if dx == 0 and dy == 0:
                deposit(px, py, sed)
                sed = 0.0
                alive = False
                break
            px, py = px + dx, py + dy
            if not (0 <= px < n - 1 and 0 <= py < n - 1):
                exported += sed
                sed = 0.0
                alive = False
                break
            dh = bil(px, py) - h
            cap = max(-dh, f32(ds.min_slope)) * speed * water * f32(ds.capacity)
            if sed > cap or dh > 0:
                a = min(dh, sed) if dh > 0 else (sed - cap) * f32(ds.deposition)
                sed -= a
                deposit(ox, oy, a)
            else:
                a = min((cap - sed) * f32(ds.erosion), -dh)
                inside = [(bx, by, w) for bx, by, w in taps if 0 <= cx + bx < n and 0 <= cy + by < n]
                interior = r <= cx < n - r and r <= cy < n - r
                scale = a if interior else a / sum(w for _, _, w in inside)
                for bx, by, w in inside:
                    want = scale * w
                    td = min(max(dirt[cy + by, cx + bx], 0.0), want)
                    tr = min(max(rock[cy + by, cx + bx], 0.0), want - td)
                    dirt[cy + by, cx + bx] -= td
                    rock[cy + by, cx + bx] -= tr
                    sed += td + tr
            speed = math.sqrt(max(0.0, speed * speed - dh * f32(ds.gravity)))
            water = water * (1 - f32(ds.evaporation))
        if alive:
            deposit(px, py, sed)
    return exported


def test_droplets_match_python_restatement_and_conserve_mass():
    n = 40
    p = O.fbm_params(seed=4, base_freq=1 / 16, hi=20.0)
    rock = O.fbm_init(n, p, np.float64)
    dirt = np.full((n, n), 0.05)
    ds = O.droplet_settings(count=60, max_lifetime=48, seed=11, brush_radius=3)
    r1, d1 = rock.copy(), dirt.copy()
    ex1 = O.droplets(r1, d1, ds, 2)
    r2, d2 = rock.copy(), dirt.copy()
    ex2 = droplets_python(r2, d2, n, ds, 2)
    assert ex1 == pytest.approx(ex2, rel=1e-9, abs=1e-12)
    assert np.allclose(r1, r2, rtol=0, atol=1e-10) and np.allclose(d1, d2, rtol=0, atol=1e-10)
    assert not np.allclose(r1, rock)          # the droplets did erode
    before = rock.sum() + dirt.sum()
    assert r1.sum() + d1.sum() + ex1 == pytest.approx(before, rel=1e-13)
    assert (r1 >= 0).all() and (d1 >= -1e-12).all()
