"""Pins of the oracle's restatement of the reference particle mode
(glsl/particle.glsl:64-136, glsl/particle_erosion.glsl:22-128,
glsl/smoothing.glsl:76-95): closed forms on flat and tilted planes, the kill
rules, the corner deposit sums, the dissolve cases (including the sediment
overwrite of successive corners), the particle branch of smoothing, and an
independent plain-Python restatement of movement + erosion.  CPU only."""
from __future__ import annotations

import math

import numpy as np
import pytest

import pyoracle as O

N = 32


def f32(v):
    """A setting as the FP32 field the oracle reads it from."""
    return float(np.float32(v))


def planes(n=N, dtype=np.float64):
    return np.zeros((O.NUM_PLANES, n, n), dtype=dtype)


def particle(x, y, iters=1, volume=1.0, vx=0.0, vy=0.0, sed=(0.0, 0.0), to_kill=0, sc=0.0):
    p = O.RefParticle()
    p.position[0], p.position[1] = x, y
    p.velocity[0], p.velocity[1] = vx, vy
    p.iters, p.volume, p.to_kill, p.sc = iters, volume, to_kill, sc
    p.sediment[0], p.sediment[1] = sed
    return p


def settings(**kw):
    s = O.reference_particle_defaults(1)
    for k, v in kw.items():
        if k in ("Ks", "Kd"):
            getattr(s, k)[0], getattr(s, k)[1] = v
        else:
            setattr(s, k, v)
    return s


def run(st, parts, s, iters=1, rain=False, flags=0, time0=0.0):
    arr = (O.RefParticle * len(parts))(*parts)
    O.ref_particle_run(st, arr, s, iters, rain, time0, 0.0, flags)
    return arr


def test_rand_in_unit_interval():
    xs = np.linspace(0.0, 999.0, 37)
    for x in xs:
        for y in (0.0, 1.0, 77.0, 1e5, 3.14, 1e6 + 3.14):
            r = O.ref_particle_rand(float(x), float(y), np.float64)
            assert 0.0 <= r < 1.0


def test_spawn_hash_fp32_sin_is_rounded_fp64_sin():
    """The spawn hash's FP32 sin (the method the device kernel runs: FP64
    evaluation with a three-part pi/2 reduction, one rounding) against FP64
    sin rounded to FP32, over the arguments the hash produces (17 x + 0.1 y
    up to ~2e4 and 13 y + x up to 13 * 2^26) and near the quadrant
    boundaries: equal, which a missing reduction part or polynomial term
    would break for large arguments."""
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-2e4, 2e4, 20000), rng.uniform(0.0, 13.0 * 2 ** 26, 20000),
                         np.arange(-40, 41) * (math.pi / 2), np.linspace(-10.0, 10.0, 2001)]).astype(np.float32)
    for x in xs:
        assert np.float32(O.hash_sin_f32(float(x))) == np.float32(math.sin(float(x))), x
    # and the hash in FP32 stays in [0, 1)
    for x in np.linspace(0.0, 999.0, 37):
        for y in (0.0, 1.0, 77.0, 1e5, 3.14, 1e6 + 3.14, 6.7e7):
            r = O.ref_particle_rand(float(x), float(y), np.float32)
            assert 0.0 <= r < 1.0


def test_flat_terrain_particle_stops_and_dies():
    st = planes()
    st[O.ROCK] = 10.0
    s = settings()
    p = run(st, [particle(10.3, 12.6)], s)[0]
    # normal (0, -1, 0): no acceleration, no momentum -> speed 0 < min_velocity
    assert p.velocity[0] == 0.0 and p.velocity[1] == 0.0
    assert p.position[0] == pytest.approx(10.3) and p.position[1] == pytest.approx(12.6)
    assert p.volume == pytest.approx(1.0 - 0.25 * 0.03)
    assert p.sc == 0.0 and p.iters == 2 and p.to_kill == 1


@pytest.mark.parametrize("a,b", [(0.3, 0.0), (0.0, -0.2), (0.15, 0.1)])
def test_tilted_plane_closed_form(a, b):
    """rock = a x + b y: the bilinear samples are exact, dx = 2a, dz = 2b, the
    normal is (a, -1, b)/sqrt(1 + a^2 + b^2); from rest one step gives
    v = -dt G (a, b) / L, then the friction factor 1 - dt f ny."""
    st = planes()
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    st[O.ROCK] = 50.0 + a * x + b * y
    s = settings(G=9.81, d_t=0.25, friction=0.2, Kc=0.2)
    x0, y0 = 15.25, 16.5
    p = run(st, [particle(x0, y0)], s)[0]
    L = math.sqrt(1 + a * a + b * b)
    v = np.array([-0.25 * 9.81 * a / L, -0.25 * 9.81 * b / L])
    if np.hypot(*v) > 1.0:
        v = v / np.hypot(*v)
    pos = np.array([x0, y0]) + 0.25 * v
    fr = 1.0 + 0.25 * 0.2 / L
    vf = v * fr
    assert p.position[0] == pytest.approx(pos[0], abs=1e-5) and p.position[1] == pytest.approx(pos[1], abs=1e-5)
    assert p.velocity[0] == pytest.approx(vf[0], rel=1e-5, abs=1e-7)
    assert p.velocity[1] == pytest.approx(vf[1], rel=1e-5, abs=1e-7)
    sin_a = math.sqrt((a * a + b * b) / (1 + a * a + b * b))
    vol = 1.0 - 0.25 * 0.03
    assert p.sc == pytest.approx(0.2 * vol * np.hypot(*vf) * max(0.02, sin_a), rel=1e-5)


def test_leaving_the_map_restores_position_and_kills():
    st = planes()
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    st[O.ROCK] = 100.0 + 20.0 * x      # steep in +x: the particle runs to -x
    s = settings()
    p = run(st, [particle(1.1, 10.0)], s)[0]
    assert p.position[0] == pytest.approx(1.1) and p.velocity[0] == 0.0 and p.to_kill == 1


def test_ttl_and_spawn_rules():
    st = planes()
    st[O.ROCK] = 5.0
    s = settings(ttl=3, min_velocity=0.0)
    p = run(st, [particle(10.0, 10.0, iters=2)], s)[0]
    assert p.iters == 3 and p.to_kill == 1
    # not raining: an unspawned particle stays as it is, erosion skips it
    q = run(st, [particle(0.0, 0.0, iters=0, volume=0.0)], s, rain=False)[0]
    assert q.iters == 0 and q.position[0] == 0.0
    # raining: spawned inside [2, n-2) with the initial volume, iters 1 -> 2
    r = run(st, [particle(0.0, 0.0, iters=0, volume=0.0)], s, rain=True, time0=3.7)[0]
    assert 2.0 <= r.position[0] < N - 2 and 2.0 <= r.position[1] < N - 2
    assert r.iters == 2


def test_killed_particle_deposits_its_sediment_over_the_quad():
    st = planes()
    st[O.ROCK] = 7.0
    st[O.DIRT] = 1.0
    s = settings()
    before = st[[O.ROCK, O.DIRT]].sum(axis=(1, 2))
    p = run(st, [particle(10.3, 11.7, sed=(0.4, 0.25), to_kill=1, iters=5)], s)[0]
    # movement re-spawns nothing (not raining, to_kill: returns unwritten);
    # erosion deposits old_sediment * weight at each corner (weights sum to 1)
    after = st[[O.ROCK, O.DIRT]].sum(axis=(1, 2))
    assert after[0] - before[0] == pytest.approx(f32(0.4), abs=1e-12)
    assert after[1] - before[1] == pytest.approx(f32(0.25), abs=1e-12)
    # the rock->dirt conversion still runs after each corner: what is left of
    # the rock sediment moves to dirt, so the particle ends at (-c, +c)
    assert p.sediment[0] < 0 and abs(p.sediment[0] + p.sediment[1]) < 1e-7
    assert -p.sediment[0] < 4 * f32(0.4) * f32(0.001) * 0.25


def test_dissolve_overwrites_sediment_at_each_corner():
    """On a tilted plane with dirt, a fresh particle dissolves dirt at every
    corner; each corner starts from the particle's sediment before the pass,
    so the sediment after the pass is what the last corner (0,1) left,
    converted 4 times, not the sum over corners (particle_erosion.glsl:49,64)."""
    st = planes()
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    st[O.ROCK] = 50.0 + 0.4 * x
    st[O.DIRT] = 2.0
    s = settings(Kconv=0.0)
    p = run(st, [particle(15.4, 16.2)], s)[0]
    sc = p.sc
    assert sc > 0
    xs, ys = p.position[0], p.position[1]
    fx, fy = xs - math.floor(xs), ys - math.floor(ys)
    w3 = (1 - fx) * fy                      # corner 3 = (x, y + 1)
    k1 = 0.25 * f32(0.09)
    eroded3 = w3 * k1 * sc                  # multipl * dt * Ks[dirt] * (c - 0)
    assert p.sediment[1] == pytest.approx(eroded3, rel=1e-6)
    total = sum(w * k1 * sc for w in ((1 - fx) * (1 - fy), fx * (1 - fy), fx * fy, w3))
    assert 2.0 * N * N - st[O.DIRT].sum() == pytest.approx(total, rel=1e-6)
    assert np.array_equal(st[O.ROCK], 50.0 + 0.4 * x)      # break after the dirt layer


def test_exhausted_dirt_passes_capacity_to_rock():
    st = planes()
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    st[O.ROCK] = 50.0 + 0.4 * x
    st[O.DIRT] = 0.0
    s = settings(Kconv=0.0)
    p = run(st, [particle(15.4, 16.2)], s)[0]
    sc = p.sc
    xs, ys = p.position[0], p.position[1]
    fx, fy = xs - math.floor(xs), ys - math.floor(ys)
    w3 = (1 - fx) * fy
    # dirt: s1 = 0 + eroded + (0 - eroded) = 0; cap += 0; rock: c = sc
    assert p.sediment[1] == 0.0
    assert p.sediment[0] == pytest.approx(w3 * 0.25 * f32(0.03) * sc, rel=1e-6)
    assert not st[O.DIRT].any()


def test_water_display_and_momentum_accumulation():
    st = planes()
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    st[O.ROCK] = 50.0 + 0.3 * y
    s = settings()
    p = run(st, [particle(12.5, 14.25)], s)[0]
    # (the particle comes back rounded to FP32; the planes are FP64)
    assert st[O.WATER].sum() == pytest.approx(1e-5 * p.volume, rel=1e-6)
    assert st[O.FT].sum() == pytest.approx(p.volume * p.velocity[0], rel=1e-6, abs=1e-300)
    assert st[O.FB].sum() == pytest.approx(p.volume * p.velocity[1], rel=1e-6)


def test_smoothing_particle_branch_closed_form():
    rng = np.random.default_rng(3)
    st = planes()
    st[O.WATER] = rng.uniform(0, 1e-5, (N, N))
    for k in (O.FL, O.FR, O.FT, O.FB):
        st[k] = rng.normal(0, 1, (N, N))
    ref = st.copy()
    cnt = 262144
    s = settings(particle_count=cnt)
    run(st, [particle(0.0, 0.0, iters=0)], s, flags=O.STEP_SMOOTH)
    c = float(np.float32(cnt))
    inner = (slice(1, -1), slice(1, -1))
    fm = min(max(1 - 1e-12 * c, 0), 1)
    mx = ref[O.FL][inner] * fm + 1e-12 * c * ref[O.FT][inner]
    assert np.allclose(st[O.FL][inner], mx, rtol=1e-12, atol=0)
    assert not st[O.FT][inner].any() and not st[O.FB][inner].any()
    b = ref[O.WATER][inner] * min(max(1 - 8e-8 * c, 0), 1)
    b[b < 1e-6] = 0
    assert np.allclose(st[O.WATER][inner], b, rtol=1e-12, atol=0)
    # border cells keep water and momentum
    for k in (O.WATER, O.FL, O.FT):
        assert np.array_equal(st[k][0], ref[k][0]) and np.array_equal(st[k][:, -1], ref[k][:, -1])


# ---- an independent plain-Python restatement of movement + erosion ----------

def _bil(img, x, y):
    n = img.shape[0]
    ix, iy = int(x), int(y)
    fx, fy = x - math.floor(x), y - math.floor(y)

    def f(a, b):
        return img[b, a] if 0 <= a < n and 0 <= b < n else 0.0
    v1 = f(ix, iy) + fx * (f(ix + 1, iy) - f(ix, iy))
    v2 = f(ix, iy + 1) + fx * (f(ix + 1, iy + 1) - f(ix, iy + 1))
    return v1 + fy * (v2 - v1)


def _py_dispatch(st, ps, s):
    n = st.shape[-1]
    R, D, B, MX, MY, MZ, MW = (st[k] for k in (O.ROCK, O.DIRT, O.WATER, O.FL, O.FR, O.FT, O.FB))
    dt, G = s.d_t, s.G
    for p in ps:                        # movement (no spawning in this case)
        x, y = p["pos"]
        h = lambda u, v: _bil(R, u, v) + _bil(D, u, v)
        gx = h(x + 1, y) - h(x - 1, y)
        gz = h(x, y + 1) - h(x, y - 1)
        L = math.sqrt(4 * gx * gx + 16 + 4 * gz * gz)
        nx, ny, nz = 2 * gx / L, -4 / L, 2 * gz / L
        vx, vy = p["vel"]
        vx -= dt * nx / p["vol"] * G
        vy -= dt * nz / p["vol"] * G
        mx, my = _bil(MX, x, y), _bil(MY, x, y)
        ml, vl = math.hypot(mx, my), math.hypot(vx, vy)
        if ml > 0 and vl > 0:
            t = s.inertia * (mx * vx + my * vy) / (ml * vl) / (p["vol"] + 1e5 * _bil(B, x, y))
            vx += t * mx
            vy += t * my
        vl = math.hypot(vx, vy)
        if vl > 1:
            vx, vy = vx / vl, vy / vl
        qx, qy = x + dt * vx, y + dt * vy
        kill = False
        if qx <= 1 or qy <= 1 or qx >= n - 2 or qy >= n - 2:
            qx, qy, vx, vy, kill = x, y, 0.0, 0.0, True
        fr = 1 - dt * s.friction * ny
        vx, vy = vx * fr, vy * fr
        p["vol"] -= dt * s.Ke
        vl = math.hypot(vx, vy)
        p["sc"] = max(0.0, s.Kc * p["vol"] * vl * max(0.02, math.sqrt(1 - ny * ny)))
        p["iters"] += 1
        if p["vol"] <= s.min_volume or vl < s.min_velocity or p["iters"] >= s.ttl:
            kill = True
        p["pos"], p["vel"], p["kill"] = (qx, qy), (vx, vy), kill
    for p in ps:                        # erosion, id order
        x, y = p["pos"]
        bx, by = int(x), int(y)
        fx, fy = x - bx, y - by
        old = list(p["sed"])
        for cx, cy, w in ((bx, by, (1 - fx) * (1 - fy)), (bx + 1, by, fx * (1 - fy)),
                          (bx + 1, by + 1, fx * fy), (bx, by + 1, (1 - fx) * fy)):
            terr = [R[cy, cx], D[cy, cx]]
            cap = 0.0
            for l in (1, 0):
                if p["kill"]:
                    terr[l] += old[l] * w
                    p["sed"][l] -= old[l] * w
                    continue
                c = max(0.0, p["sc"] - cap)
                s1 = old[l]
                if c > s1:
                    e = w * dt * s.Ks[l] * (c - s1)
                    s1 += e
                    t = terr[l] - e
                    if t < 0:
                        s1 += t
                        cap += terr[l]
                        terr[l] = 0.0
                    else:
                        terr[l] = t
                        p["sed"][l] = s1
                        break
                else:
                    dep = w * dt * s.Kd[l] * (s1 - c)
                    s1 -= dep
                    terr[l] += dep
                p["sed"][l] = s1
            conv = p["sed"][0] * s.Kconv * dt
            p["sed"][1] += conv
            p["sed"][0] -= conv
            R[cy, cx], D[cy, cx] = terr
            B[cy, cx] += 1e-5 * p["vol"] * w
            MZ[cy, cx] += p["vol"] * p["vel"][0] * w
            MW[cy, cx] += p["vol"] * p["vel"][1] * w


def test_independent_python_restatement():
    rng = np.random.default_rng(11)
    n = 24
    st = np.zeros((O.NUM_PLANES, n, n))
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    st[O.ROCK] = 30 + 3 * np.sin(x / 4.0) * np.cos(y / 5.0) + 0.2 * x
    st[O.DIRT] = rng.uniform(0, 0.3, (n, n))
    st[O.WATER] = rng.uniform(0, 1e-4, (n, n))
    st[O.FL] = rng.normal(0, 0.1, (n, n))
    st[O.FR] = rng.normal(0, 0.1, (n, n))
    s = settings(Kconv=0.01)
    init = [dict(pos=(float(rng.uniform(3, n - 4)), float(rng.uniform(3, n - 4))),
                 vel=(float(rng.normal(0, 0.2)), float(rng.normal(0, 0.2))), vol=1.0,
                 sed=[float(rng.uniform(0, 0.05)), float(rng.uniform(0, 0.05))], iters=3, kill=False, sc=0.0)
            for _ in range(6)]
    ps = [particle(*q["pos"], iters=q["iters"], vx=q["vel"][0], vy=q["vel"][1], sed=tuple(q["sed"])) for q in init]
    # the C API rounds particles to FP32: start the Python model from the same values
    for q, p in zip(init, ps):
        q["pos"] = (p.position[0], p.position[1]); q["vel"] = (p.velocity[0], p.velocity[1])
        q["sed"] = [p.sediment[0], p.sediment[1]]
    st_c = st.copy()
    out = run(st_c, ps, s)
    st_p = st.copy()
    _py_dispatch(st_p, init, s)
    for k in (O.ROCK, O.DIRT, O.WATER, O.FT, O.FB):
        assert np.allclose(st_c[k], st_p[k], rtol=1e-11, atol=1e-13), k
    for p, q in zip(out, init):
        assert p.position[0] == pytest.approx(q["pos"][0], rel=1e-6)
        assert p.velocity[1] == pytest.approx(q["vel"][1], rel=1e-6, abs=1e-8)
        assert p.sediment[0] == pytest.approx(q["sed"][0], rel=1e-6, abs=1e-12)
        assert p.sediment[1] == pytest.approx(q["sed"][1], rel=1e-6, abs=1e-12)
        assert p.sc == pytest.approx(q["sc"], rel=1e-6)
