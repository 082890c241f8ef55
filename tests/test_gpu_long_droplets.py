"""GPU: the particle benchmark's droplet workload at full size.

2048 x 2048 fBm heightmap in [0, 256] (seed 3, frequency 1/512), 2^20
droplets per iteration, at most 64 lifetime steps, inertia 0.05, capacity
4, deposition and erosion 0.3, evaporation 0.01, gravity 4, min slope 0.01,
brush radius 3.  Droplets of one iteration run concurrently and change the
heights with float atomics in an order the hardware picks, so the run is
judged as the particle benchmark states: by mass balance over 100
iterations, and by the L1 distance between height histograms after one
iteration against the serial FP64 oracle, next to the serial FP32 oracle's
distance (the rounding floor of the model).  No layer may end below 0 by
more than FP32 rounding: concurrent droplets eroding one cell take at most
what it holds, and a take that found less than it expected gives the rest
back with one more float add, which can leave a residue of a few ulp."""
from __future__ import annotations

import time

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N, COUNT = 2048, 1 << 20


def hist_l1(a, b, bins=256):
    lo, hi = min(a.min(), b.min()), max(a.max(), b.max())
    ha, _ = np.histogram(a, bins=bins, range=(lo, hi))
    hb, _ = np.histogram(b, bins=bins, range=(lo, hi))
    return float(np.abs(ha / ha.sum() - hb / hb.sum()).sum())


def test_droplets_2048_full_size(capsys):
    d = hg.droplet_settings(count=COUNT, seed=42)
    with hg.Terrain(N) as t:
        t.init_fbm(hg.fbm_params(seed=3, base_freq=1.0 / 512, hi=256.0))
        h0 = t.read_plane(hg.P_ROCK)
        t.set_droplets(d)
        t.dispatch_particle(1)
        r1, d1 = t.read_plane(hg.P_ROCK), t.read_plane(hg.P_DIRT)
        ex1, _ = t.particle_stats()
        t.dispatch_particle(99)
        r, dd = t.read_plane(hg.P_ROCK), t.read_plane(hg.P_DIRT)
        ex, steps = t.particle_stats()
    m0 = float(h0.astype(np.float64).sum())
    h1 = r1.astype(np.float64) + d1
    h = r.astype(np.float64) + dd
    res1, res = (float(h1.sum()) + ex1 - m0) / m0, (float(h.sum()) + ex - m0) / m0
    od = O.droplet_settings(count=COUNT, seed=42)
    r64, d64 = h0.astype(np.float64), np.zeros((N, N))
    r32, d32 = h0.copy(), np.zeros((N, N), np.float32)
    t0 = time.perf_counter()
    oex = O.droplets(r64, d64, od, 0)
    O.droplets(r32, d32, od, 0)
    h64, h32 = r64 + d64, r32.astype(np.float64) + d32
    l1, floor, init = hist_l1(h1, h64), hist_l1(h32, h64), hist_l1(h0, h64)
    with capsys.disabled():
        print(f"\ndroplets {N}^2 x 2^20: mass residual {res1:.2e} after 1, {res:.2e} after 100 iterations "
              f"({steps} droplet-steps); min rock {r.min():.3e} dirt {dd.min():.3e} "
              f"({np.count_nonzero(r < 0) + np.count_nonzero(dd < 0)} cells < 0); histogram L1 vs FP64 "
              f"after 1: gpu {l1:.4f} fp32 oracle {floor:.4f} initial {init:.4f} "
              f"(oracles {time.perf_counter() - t0:.0f} s)", flush=True)
    assert abs(res1) <= 1e-6 and abs(res) <= 1e-6
    assert min(r.min(), dd.min(), r1.min(), d1.min()) >= -1e-5
    assert abs(h64.sum() + oex - m0) <= 1e-9 * m0
    assert l1 <= 3 * floor + 0.01 and l1 < 0.75 * init
