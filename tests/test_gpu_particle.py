"""GPU droplet erosion against the serial FP64 oracle.

Droplets of one iteration run concurrently and add to the heights with float
atomics in an order the hardware chooses, so a many-droplet run is compared
with the serial oracle by mass balance and by the L1 distance of the height
histograms; a single droplet has no concurrency and is compared cell by cell.
"""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O

pytestmark = pytest.mark.gpu


def terrain(n, seed=1, hi=256.0, base_freq=None):
    p = hg.fbm_params(seed=seed, base_freq=base_freq or 4.0 / n, hi=hi)
    t = hg.Terrain(n)
    t.init_fbm(p)
    return t


@pytest.mark.parametrize("slope", [(0.3, 0.1), (-0.2, 0.25), (0.05, -0.4)])
def test_droplets_on_a_plane_match_fp64_oracle(slope):
    """On a tilted plane every droplet runs straight downhill, a path that is
    stable under rounding, so the GPU matches the FP64 oracle cell by cell
    (droplets spaced apart: 16 of them, paths of 64 cells)."""
    n = 192
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    plane = (100.0 + slope[0] * (x - n / 2) + slope[1] * (y - n / 2)).astype(np.float32)
    d = hg.droplet_settings(count=1, seed=7)
    with hg.Terrain(n) as t:
        t.write_plane(hg.P_ROCK, plane)
        t.set_droplets(d)
        t.dispatch_particle(2)
        h = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
        ex, steps = t.particle_stats()
    r64, d64 = plane.astype(np.float64), np.zeros((n, n))
    oex = sum(O.droplets(r64, d64, O.droplet_settings(count=1, seed=7), it) for it in range(2))
    h64 = r64 + d64
    change = np.abs(h64 - plane).max()
    err = np.abs(h - h64).max()
    print(f"plane {slope}: steps {steps}, max change {change:.3e}, |gpu-fp64| {err:.3e}, exported {ex:.4e} vs {oex:.4e}")
    assert change > 1e-3
    assert err <= 1e-3 * change
    assert abs(ex - oex) <= 1e-3 * max(abs(oex), 1e-3)


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
def test_single_droplet_matches_fp32_oracle(seed):
    """One droplet has no concurrency: over a 16-step lifetime the GPU follows
    the FP32 oracle (the same operations in the same order) to the last bit or
    close to it.  Over longer lifetimes any last-bit difference (device and
    host library functions need not round alike) is amplified by the model's
    erode/deposit switching until the paths part, for the FP32 and FP64
    oracles as much as for the GPU."""
    n = 128
    d = hg.droplet_settings(count=1, seed=seed, max_lifetime=16)
    with terrain(n, seed=seed, hi=64.0, base_freq=1.0 / n) as t:
        r0 = t.read_plane(hg.P_ROCK)
        t.set_droplets(d)
        t.dispatch_particle(1)
        h = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
        ex, steps = t.particle_stats()
    od = O.droplet_settings(count=1, seed=seed, max_lifetime=16)
    r32, d32 = r0.copy(), np.zeros_like(r0)
    ex32 = O.droplets(r32, d32, od, 0)
    h32 = r32.astype(np.float64) + d32
    change = np.abs(h32 - r0).max()
    err = np.abs(h - h32).max()
    print(f"seed {seed}: steps {steps}, max change {change:.3e}, |gpu-fp32| {err:.3e}")
    assert steps > 0 and change > 0
    assert err <= 1e-5 * change
    assert abs(ex - ex32) <= 1e-5 * max(abs(ex32), 1.0)


@pytest.mark.parametrize("count", [4096, 3000])
def test_droplet_order_modes_run_every_droplet_once(count):
    """Both thread-to-droplet orders (hg_set_droplet_order) run the same
    droplets.  One-step droplets on a steep rock plane (slope 10 per cell)
    under 1.0 of dirt: min_slope 20 fixes each droplet's capacity, so each
    takes 0.02 of dirt from the brush at its start and deposits it one cell
    downhill, whatever the others did (their changes tilt its direction by
    under 1e-4).  The dirt field is then the sum of every droplet's own
    footprint up to FP32 rounding of the atomic adds (values near 1: ~1e-7),
    while a droplet run twice or skipped would change it by at least 8.8e-5
    (the smallest brush weight times 0.02).  count 3000 leaves the last
    workgroup part-filled."""
    n = 256
    rock = np.broadcast_to((10.0 * np.arange(n)).astype(np.float32)[None, :], (n, n)).copy()
    d = hg.droplet_settings(count=count, max_lifetime=1, min_slope=20.0, erosion=0.00025, seed=11)
    fields, stats = [], []
    for mode in (0, 1):
        with hg.Terrain(n) as t:
            t.write_plane(hg.P_ROCK, rock)
            t.write_plane(hg.P_DIRT, np.ones((n, n), np.float32))
            t.set_droplets(d)
            t.set_droplet_order(mode)
            assert t.get_droplet_order() == mode
            t.dispatch_particle(2)
            assert np.array_equal(t.read_plane(hg.P_ROCK), rock)
            fields.append(t.read_plane(hg.P_DIRT).astype(np.float64))
            stats.append(t.particle_stats())
    change = np.abs(fields[0] - 1.0).max()
    diff = float(np.abs(fields[1] - fields[0]).max())
    print(f"droplet orders, {count} one-step droplets x 2: max change {change:.3e}, |mode1 - mode0| {diff:.2e}, "
          f"steps {[s[1] for s in stats]}, exported {[s[0] for s in stats]}", flush=True)
    assert change > 1e-3
    assert all(s[1] == 2 * count for s in stats)
    assert diff <= 1e-5
    assert all(abs(s[0] - stats[0][0]) <= 1e-6 for s in stats)


def hist_l1(a, b, bins=128):
    lo, hi = min(a.min(), b.min()), max(a.max(), b.max())
    ha, _ = np.histogram(a, bins=bins, range=(lo, hi))
    hb, _ = np.histogram(b, bins=bins, range=(lo, hi))
    return float(np.abs(ha / ha.sum() - hb / hb.sum()).sum())


@pytest.mark.parametrize("hi,freq,order", [(256.0, 4.0, 0), (64.0, 1.0, 0), (256.0, 4.0, 1)])
def test_mass_balance_and_histogram_vs_oracle(hi, freq, order):
    """8192 droplets x 4 iterations on 256^2.  Mass: sum(height) + exported is
    conserved to 1e-6.  Distribution: the GPU's histogram L1 distance to the
    serial FP64 oracle stays within 3x the distance between the FP32 and FP64
    serial oracles (the rounding floor of the model), and well below the
    initial terrain's distance (concurrent droplets see each other's changes
    later than serial ones do, which shows up as a slightly larger distance
    than the rounding floor).  order 1 (droplets sorted by start tile) runs
    neighbouring droplets in lockstep, which moves the spread of heights
    further from the serial model."""
    n, count, iters = 256, 8192, 4
    d = hg.droplet_settings(count=count, seed=11)
    with terrain(n, seed=5, hi=hi, base_freq=freq / n) as t:
        h0 = t.read_plane(hg.P_ROCK)
        t.set_droplets(d)
        t.set_droplet_order(order)
        t.dispatch_particle(iters)
        h = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
        ex, steps = t.particle_stats()
    m0, m1 = h0.astype(np.float64).sum(), h.sum()
    print(f"mass in {m0:.6e} out {m1:.6e} exported {ex:.6e} residual {(m1 + ex - m0) / m0:.3e}, steps {steps}")
    assert abs(m1 + ex - m0) <= 1e-6 * m0
    od = O.droplet_settings(count=count, seed=11)
    r64, d64 = h0.astype(np.float64), np.zeros((n, n))
    r32, d32 = h0.copy(), np.zeros_like(h0)
    oex = 0.0
    for it in range(iters):
        oex += O.droplets(r64, d64, od, it)
        O.droplets(r32, d32, od, it)
    h64, h32 = r64 + d64, r32.astype(np.float64) + d32
    assert abs(h64.sum() + oex - m0) <= 1e-9 * m0       # the oracle conserves mass exactly
    l1, floor, init = hist_l1(h, h64), hist_l1(h32, h64), hist_l1(h0, h64)
    sd, sd64, sd32 = np.std(h), np.std(h64), np.std(h32)
    print(f"histogram L1 vs FP64 (order {order}): gpu {l1:.4f}, fp32 oracle {floor:.4f}, initial {init:.4f}; "
          f"std gpu {sd:.4f} fp64 {sd64:.4f} fp32 {sd32:.4f} ({abs(sd - sd64) / sd64:.2%} off)")
    assert l1 <= 3 * floor + 0.01 and l1 < 0.75 * init
    # The spread of the heights: concurrent droplets read heights that other
    # droplets of the same iteration have already changed, a systematic
    # difference from the serial oracle that varies with the hardware's
    # ordering (seen up to 2.6 % of the spread in order 0), beyond the
    # rounding floor.
    assert abs(sd - sd64) <= 3 * abs(sd32 - sd64) + (0.025 if order == 0 else 0.1) * sd64


def test_droplet_settings_validation():
    with hg.Terrain(64) as t:
        bad = hg.droplet_settings(brush_radius=9)
        with pytest.raises(hg.HGError):
            t.set_droplets(bad)
        bad = hg.droplet_settings(inertia=float("nan"))
        with pytest.raises(hg.HGError):
            t.set_droplets(bad)
        with pytest.raises(hg.HGError):
            t.dispatch_particle(1, hg.STEP_RAIN)


def test_particle_iteration_then_thermal_runs():
    with terrain(128) as t:
        t.set_erosion(hg.erosion_settings())
        t.set_droplets(hg.droplet_settings(count=4096, seed=3))
        t.dispatch_particle(2, hg.STEP_THERMAL | hg.STEP_SMOOTH)
        h = t.read_plane(hg.P_ROCK)
        assert np.isfinite(h).all()
