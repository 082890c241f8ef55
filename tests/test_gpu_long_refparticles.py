"""GPU: the reference's own particle mode at the benchmark's size.

2048 x 2048 fBm heightmap in [0, 256] (seed 3, frequency 1/512), 262,144
particles (the reference's default count) with the reference particle
settings, 10 dispatches of hg_step_particle_reference: movement, per-corner
erosion, thermal erosion of both layers and smoothing with its particle
branch (src/erosion.cpp:132-156).  Particles start at seeded positions and
re-spawn through the reference's hash when they die.  Concurrent particles
update shared pixels in an order the hardware picks, so the run is compared
by distribution: heights and water against the serial FP64 oracle, next to
the serial FP32 oracle's distance to it."""
from __future__ import annotations

import ctypes
import time

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N, COUNT, ITERS = 2048, 262144, 10


def _to_oracle(arr):
    out = (O.RefParticle * len(arr))()
    ctypes.memmove(ctypes.addressof(out), ctypes.addressof(arr), ctypes.sizeof(arr))
    return out


def test_reference_particle_mode_2048_full_size(capsys):
    rng = np.random.default_rng(17)
    xy = rng.uniform(3.0, N - 4.0, (COUNT, 2))
    parts = (hg.RefParticle * COUNT)()
    for p, (x, y) in zip(parts, xy):
        p.position[0], p.position[1] = float(x), float(y)
        p.iters, p.volume = 1, 1.0
    s = hg.reference_particle_settings(COUNT)
    flags = hg.STEP_THERMAL | hg.STEP_SMOOTH
    with hg.Terrain(N) as t:
        t.init_fbm(hg.fbm_params(seed=3, base_freq=1.0 / 512, hi=256.0))
        t.set_erosion(s)
        t.ref_particles_init(COUNT)
        t.write_ref_particles(parts)
        st = t.read_state(hg.NUM_PLANES)
        t.dispatch_particle_reference(ITERS, should_rain=True, time0=2.0, time_step=0.5, flags=flags)
        gst = t.read_state(hg.NUM_PLANES)
        giveups = t.ref_particle_stats()
    os_ = O.ErosionSettings()
    ctypes.memmove(ctypes.addressof(os_), ctypes.addressof(s), ctypes.sizeof(s))
    st64, st32 = st.astype(np.float64), st.copy()
    t0 = time.perf_counter()
    O.ref_particle_run(st64, _to_oracle(parts), os_, ITERS, True, 2.0, 0.5, O.STEP_THERMAL | O.STEP_SMOOTH)
    O.ref_particle_run(st32, _to_oracle(parts), os_, ITERS, True, 2.0, 0.5, O.STEP_THERMAL | O.STEP_SMOOTH)
    dt = time.perf_counter() - t0
    h = gst[hg.P_ROCK].astype(np.float64) + gst[hg.P_DIRT]
    h64 = st64[O.ROCK] + st64[O.DIRT]
    h32 = st32[O.ROCK].astype(np.float64) + st32[O.DIRT]
    h0 = st[O.ROCK].astype(np.float64) + st[O.DIRT]

    def rl2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    change, e_gpu, e_32 = rl2(h0, h64), rl2(h, h64), rl2(h32, h64)
    w_gpu, w_64 = float(gst[hg.P_WATER].astype(np.float64).sum()), float(st64[O.WATER].sum())
    with capsys.disabled():
        print(f"\nreference particles {N}^2 x {COUNT}, {ITERS} dispatches: heights changed {change:.3e}; "
              f"gpu vs fp64 {e_gpu:.3e}, fp32 oracle vs fp64 {e_32:.3e}; water gpu {w_gpu:.6e} fp64 {w_64:.6e}; "
              f"CAS give-ups {giveups}; oracles {dt:.0f} s", flush=True)
    assert giveups == 0
    assert change > 0
    assert e_gpu <= max(5 * e_32, 0.05 * change)
    assert abs(w_gpu - w_64) <= 0.05 * w_64
