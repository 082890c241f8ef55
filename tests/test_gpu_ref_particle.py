"""GPU: the reference particle mode (hg_step_particle_reference) against the
oracle's restatement of glsl/particle.glsl, glsl/particle_erosion.glsl and
the particle branch of glsl/smoothing.glsl.

Particles that never share a pixel in one dispatch follow the FP32 oracle bit
for bit (same operations in the same order, correctly rounded division and
square root).  Many particles update shared pixels atomically in an order the
hardware picks, and re-spawns go through sinf, whose last bit may differ from
the host's, so those runs are compared by distribution: heights and water
against the serial FP64 oracle, next to the FP32 oracle's own distance."""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O

pytestmark = pytest.mark.gpu


def _terrain(n, seed=2, hi=48.0, freq=3.0):
    t = hg.Terrain(n)
    t.init_fbm(hg.fbm_params(seed=seed, base_freq=freq / n, hi=hi))
    return t


def _smooth_field(n, seed, scale):
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(3, 9, 2)
    return (scale * np.sin(x / a) * np.cos(y / b)).astype(np.float32)


def _particles(specs):
    arr = (hg.RefParticle * len(specs))()
    for p, (x, y, vx, vy) in zip(arr, specs):
        p.position[0], p.position[1] = x, y
        p.velocity[0], p.velocity[1] = vx, vy
        p.iters, p.volume = 1, 1.0
    return arr


def _to_oracle(arr):
    out = (O.RefParticle * len(arr))()
    for a, b in zip(out, arr):
        ctypes_copy(a, b)
    return out


def ctypes_copy(dst, src):
    import ctypes
    ctypes.memmove(ctypes.addressof(dst), ctypes.addressof(src), ctypes.sizeof(src))


def _fields(p):
    return (p.sc, p.iters, p.position[0], p.position[1], p.velocity[0], p.velocity[1], p.volume,
            p.sediment[0], p.sediment[1], p.to_kill)


def _setup(n, specs, s, momentum=True):
    t = _terrain(n)
    t.set_erosion(s)
    t.ref_particles_init(len(specs))
    if momentum:      # a non-zero momentum map exercises the inertia term
        t.write_plane(hg.P_FLUX_L, _smooth_field(n, 5, 0.3))
        t.write_plane(hg.P_FLUX_R, _smooth_field(n, 6, 0.3))
    t.write_plane(hg.P_DIRT, np.abs(_smooth_field(n, 7, 0.5)))
    t.write_plane(hg.P_WATER, np.abs(_smooth_field(n, 8, 1e-5)))
    parts = _particles(specs)
    t.write_ref_particles(parts)
    return t, parts


def _oracle_state(t):
    return t.read_state(hg.NUM_PLANES)


def _settings(count, **kw):
    s = hg.reference_particle_settings(count)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _os(s):
    """The same settings block as the oracle's ctypes type."""
    o = O.ErosionSettings()
    ctypes_copy(o, s)
    return o


@pytest.mark.parametrize("start", [(20.3, 31.7, 0.0, 0.0), (40.6, 12.2, 0.3, -0.2), (9.9, 50.1, -0.1, 0.05)])
def test_single_particle_matches_fp32_oracle_bitwise(start):
    n = 64
    s = _settings(1)
    print(f"single particle from {start}", flush=True)
    t, parts = _setup(n, [start], s)
    st = _oracle_state(t)
    iters = 40
    t.dispatch_particle_reference(iters, should_rain=False, time0=0.5, time_step=0.25, flags=0)
    got = t.read_ref_particles()
    gst = t.read_state(hg.NUM_PLANES)
    t.close()
    op = _to_oracle(parts)
    O.ref_particle_run(st, op, _os(s), iters, False, 0.5, 0.25, 0)
    assert _fields(got[0]) == _fields(op[0])
    for k in (hg.P_ROCK, hg.P_DIRT, hg.P_WATER, hg.P_FLUX_T, hg.P_FLUX_B):
        assert np.array_equal(gst[k], st[k]), k
    assert got[0].iters > 1


def test_separated_particles_match_fp32_oracle_bitwise():
    n = 128
    specs = [(12.0 + 26.0 * (i % 4) + 0.37 * i, 12.0 + 26.0 * (i // 4) + 0.21 * i, 0.0, 0.0) for i in range(16)]
    s = _settings(16, d_t=0.05)      # short steps: the particles stay far apart
    t, parts = _setup(n, specs, s)
    st = _oracle_state(t)
    t.dispatch_particle_reference(12, should_rain=False, time0=1.0, time_step=0.5, flags=0)
    got = t.read_ref_particles()
    gst = t.read_state(hg.NUM_PLANES)
    assert t.ref_particle_stats() == 0
    t.close()
    op = _to_oracle(parts)
    O.ref_particle_run(st, op, _os(s), 12, False, 1.0, 0.5, 0)
    for a, b in zip(got, op):
        assert _fields(a) == _fields(b)
    for k in (hg.P_ROCK, hg.P_DIRT, hg.P_WATER, hg.P_FLUX_T, hg.P_FLUX_B):
        assert np.array_equal(gst[k], st[k]), k


def test_spawn_positions_follow_the_hash():
    """Unspawned particles with should_rain: positions from rand() of
    glsl/particle.glsl:41-44, whose sin the device evaluates with the same
    FP64 operations as the FP32 oracle (a 1e4 gain would show any last-bit
    difference): every spawn agrees bit for bit."""
    n, count = 128, 2048
    s = _settings(count)
    t = _terrain(n)
    t.set_erosion(s)
    t.ref_particles_init(count)
    st = _oracle_state(t)
    t.dispatch_particle_reference(1, should_rain=True, time0=12.34, time_step=0.0, flags=0)
    got = t.read_ref_particles()
    t.close()
    op = (O.RefParticle * count)()
    O.ref_particle_run(st, op, _os(s), 1, True, 12.34, 0.0, 0)
    gx = np.array([[p.position[0], p.position[1]] for p in got])
    ox = np.array([[p.position[0], p.position[1]] for p in op])
    same = np.all(gx == ox, axis=1).mean()
    err = np.abs(gx - ox).max()
    print(f"spawn: {same:.3f} identical, max |diff| {err:.3e} cells", flush=True)
    assert (gx >= 1.0).all() and (gx <= n - 2.0).all()
    assert same == 1.0 and err == 0.0


@pytest.mark.parametrize("dims", [(1024, 1024), (72, 88)])
def test_hmap_dims_quirk_matches_fp32_oracle_bitwise(dims):
    """Map_settings_data.hmap_dims bounds the particles (glsl/particle.glsl:111-112)
    whatever the grid: (1024, 1024) on a 96^2 grid, the reference's own
    setting, lets particles move on outside the grid (heights read 0 there and
    their corner updates are dropped); (72, 88) kills them at x >= 70 or
    y >= 86 inside it.  Separated particles follow the FP32 oracle run with
    the same dims bit for bit, and differ from a run with dims = n."""
    n = 96
    specs = [(10.3, 12.1, 0.2, 0.1), (80.4, 30.2, 0.4, 0.0), (30.6, 91.7, 0.0, 0.3), (93.9, 93.2, 0.3, 0.3),
             (120.2, 40.5, -0.2, 0.1), (50.1, 60.9, 0.1, -0.2), (69.5, 20.3, 0.5, 0.0), (20.2, 85.6, 0.0, 0.5)]
    s = _settings(len(specs), d_t=0.05)
    results = {}
    for d in (dims, (0, 0)):
        t, parts = _setup(n, specs, s)
        t.set_ref_particle_dims(*d)
        st = _oracle_state(t)
        t.dispatch_particle_reference(10, should_rain=False, time0=1.0, time_step=0.5, flags=0)
        got = t.read_ref_particles()
        gst = t.read_state(hg.NUM_PLANES)
        t.close()
        op = _to_oracle(parts)
        O.ref_particle_run(st, op, _os(s), 10, False, 1.0, 0.5, 0, dims=d)
        for a, b in zip(got, op):
            assert _fields(a) == _fields(b), d
        for k in (hg.P_ROCK, hg.P_DIRT, hg.P_WATER, hg.P_FLUX_T, hg.P_FLUX_B):
            assert np.array_equal(gst[k], st[k]), (d, k)
        results[d] = [_fields(a) for a in got]
    assert results[dims] != results[(0, 0)]


def test_hmap_dims_spawn_range():
    """With the reference's hmap_dims (1024, 1024) on a 128^2 grid, spawns fall
    in [2, 1022) (glsl/particle.glsl:78-79), most of them outside the grid."""
    n, count = 128, 4096
    s = _settings(count)
    with _terrain(n) as t:
        t.set_erosion(s)
        t.ref_particles_init(count)
        t.set_ref_particle_dims(1024, 1024)
        t.dispatch_particle_reference(1, should_rain=True, time0=3.21, time_step=0.0, flags=0)
        got = t.read_ref_particles()
        with pytest.raises(hg.HGError):
            t.set_ref_particle_dims(1024, 0)
        with pytest.raises(hg.HGError):
            t.set_ref_particle_dims(4, 4)
    xy = np.array([[p.position[0], p.position[1]] for p in got])
    print(f"spawn range x {xy[:, 0].min():.2f}..{xy[:, 0].max():.2f}, "
          f"inside the grid {np.mean((xy < n).all(axis=1)):.3f}", flush=True)
    assert (xy >= 1.0).all() and (xy < 1022.0).all()
    assert xy.max() > 900.0 and np.mean((xy < n).all(axis=1)) < 0.1


def test_many_particles_distribution_vs_oracle():
    """4096 particles on 128^2 for 10 dispatches with thermal erosion and
    smoothing, from the same random start: heights and water against the
    serial FP64 oracle, next to the FP32 oracle's distance to it."""
    n, count, iters = 128, 4096, 10
    rng = np.random.default_rng(4)
    specs = [(float(x), float(y), 0.0, 0.0) for x, y in rng.uniform(3.0, n - 4.0, (count, 2))]
    s = _settings(count)
    t, parts = _setup(n, specs, s, momentum=False)
    st = _oracle_state(t)
    print("many particles: GPU", flush=True)
    t.dispatch_particle_reference(iters, should_rain=True, time0=2.0, time_step=0.5,
                                  flags=hg.STEP_THERMAL | hg.STEP_SMOOTH)
    gst = t.read_state(hg.NUM_PLANES)
    giveups = t.ref_particle_stats()
    t.close()
    st64, st32 = st.astype(np.float64), st.copy()
    p64, p32 = _to_oracle(parts), _to_oracle(parts)
    print("many particles: oracles", flush=True)
    O.ref_particle_run(st64, p64, _os(s), iters, True, 2.0, 0.5, O.STEP_THERMAL | O.STEP_SMOOTH)
    O.ref_particle_run(st32, p32, _os(s), iters, True, 2.0, 0.5, O.STEP_THERMAL | O.STEP_SMOOTH)
    h = gst[hg.P_ROCK].astype(np.float64) + gst[hg.P_DIRT]
    h64 = st64[O.ROCK] + st64[O.DIRT]
    h32 = st32[O.ROCK].astype(np.float64) + st32[O.DIRT]
    h0 = st[O.ROCK].astype(np.float64) + st[O.DIRT]

    def rl2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    change = rl2(h0, h64)
    e_gpu, e_32 = rl2(h, h64), rl2(h32, h64)
    w_gpu, w_64 = float(gst[hg.P_WATER].sum()), float(st64[O.WATER].sum())
    print(f"heights: change {change:.3e}, gpu {e_gpu:.3e}, fp32 oracle {e_32:.3e}; water gpu {w_gpu:.6e} "
          f"fp64 {w_64:.6e}; giveups {giveups}", flush=True)
    assert giveups == 0
    assert change > 0
    assert e_gpu <= max(5 * e_32, 0.05 * change)
    assert abs(w_gpu - w_64) <= 0.05 * w_64


def test_dense_particles_atomics_vs_lock_orders():
    """The reference serialises each particle's update of a corner with a
    per-pixel lock (glsl/particle_erosion.glsl:89-99); here each layer is
    updated atomically on its own, so on a dense case (4 particles per cell,
    64^2) updates of one pixel interleave.  The distance of the GPU heights to
    the serial FP64 oracle (id order) is set beside the distance between two
    lock-consistent orders (FP64 in id and in descending id order) and beside
    the GPU's own run-to-run distance."""
    n, count, iters = 64, 16384, 4
    rng = np.random.default_rng(11)
    specs = [(float(x), float(y), 0.0, 0.0) for x, y in rng.uniform(3.0, n - 4.0, (count, 2))]
    s = _settings(count)
    runs = []
    for _ in range(2):
        t, parts = _setup(n, specs, s, momentum=False)
        st = _oracle_state(t)
        t.dispatch_particle_reference(iters, should_rain=False, time0=1.0, time_step=0.5, flags=0)
        runs.append(t.read_state(hg.NUM_PLANES))
        t.close()
    fwd, rev, f32 = st.astype(np.float64), st.astype(np.float64), st.copy()
    O.ref_particle_run(fwd, _to_oracle(parts), _os(s), iters, False, 1.0, 0.5, 0)
    O.ref_particle_run(rev, _to_oracle(parts), _os(s), iters, False, 1.0, 0.5, O.REFP_REVERSE_ORDER)
    O.ref_particle_run(f32, _to_oracle(parts), _os(s), iters, False, 1.0, 0.5, 0)

    def height(x):
        return x[O.ROCK].astype(np.float64) + x[O.DIRT]

    def rl2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    h0, hf, hr, h32 = height(st), height(fwd), height(rev), height(f32)
    g1, g2 = height(runs[0]), height(runs[1])
    change, order, fp32 = rl2(h0, hf), rl2(hr, hf), rl2(h32, hf)
    gpu, gpu_rev, gpu_gpu = rl2(g1, hf), rl2(g1, hr), rl2(g1, g2)
    print(f"dense particles ({count} on {n}^2, {iters} dispatches): change {change:.3e}; lock orders "
          f"{order:.3e}; fp32 oracle {fp32:.3e}; gpu vs id order {gpu:.3e}, vs descending {gpu_rev:.3e}; "
          f"gpu run to run {gpu_gpu:.3e}", flush=True)
    assert change > 0 and order > 0
    assert gpu <= 3.0 * max(order, fp32) and gpu_rev <= 3.0 * max(order, fp32)


def test_reference_particle_api_errors():
    with hg.Terrain(32) as t:
        with pytest.raises(hg.HGError):
            t.dispatch_particle_reference(1)
        t.ref_particles_init(8)
        with pytest.raises(hg.HGError):
            t.dispatch_particle_reference(1, flags=hg.STEP_RAIN)
        t._refp_count = 9
        with pytest.raises(hg.HGError):
            t.read_ref_particles()
