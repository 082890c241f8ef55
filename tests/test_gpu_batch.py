"""GPU: batches of uneven terrains (persistent launches, hipGraph replay)
against per-terrain launches and the FP64 oracle."""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O
from helpers import rel_l2, spun_up_state

pytestmark = pytest.mark.gpu

SIZES = [64, 96, 128, 200, 72]


def coeffs(i):
    rng = np.random.default_rng(100 + i)
    return dict(Kc=float(rng.uniform(0.02, 0.1)), Ks=float(rng.uniform(0.01, 0.05)),
                Kd=float(rng.uniform(0.01, 0.05)), Ke=float(rng.uniform(0.005, 0.02)),
                talus_deg=float(rng.uniform(30, 45)), slippage=0.1, G=9.81, dt=0.005)


def make_batch(states, sizes=SIZES, seed=5):
    b = hg.Batch(sizes)
    b.set_rain(hg.RAIN_UNIFORM, amount=0.02, seed=seed)
    for i, n in enumerate(sizes):
        b.set_erosion(hg.erosion_settings(**coeffs(i)), terrain=i)
        b.write_state(i, states[i][: hg.NUM_STATE_PLANES])
    return b


@pytest.fixture(scope="module")
def states():
    return [spun_up_state(n, 10, seed=20 + i, base_freq=1 / 32) for i, n in enumerate(SIZES)]


def run_mode(states, steps, flags):
    with make_batch(states) as b:
        b.step(steps, flags)
        return [b.read_state(i) for i in range(len(SIZES))]


def test_persistent_grid_equals_flat(states):
    ref = run_mode(states, 4, hg.STEP_DEFAULT)
    with make_batch(states) as b:
        b.set_persistent(True)
        b.step(4)
        for i in range(len(SIZES)):
            st = b.read_state(i)
            assert all(np.array_equal(st[p], ref[i][p]) for p in range(hg.NUM_STATE_PLANES)), i


@pytest.mark.parametrize("persistent", [False, True])
def test_xcd_chunked_order_equals_natural_order(states, persistent):
    """The XCD-chunked work-list (the default) and the natural one give the
    same terrains bit for bit: tiles are independent within a launch."""
    ref = run_mode(states, 4, hg.STEP_DEFAULT)
    with make_batch(states) as b:
        b.set_xcd_order(False)
        b.set_persistent(persistent)
        b.step(4)
        for i in range(len(SIZES)):
            st = b.read_state(i)
            assert all(np.array_equal(st[p], ref[i][p]) for p in range(hg.NUM_STATE_PLANES)), i


def test_batched_equals_per_terrain_and_per_pass(states):
    a = run_mode(states, 5, hg.STEP_DEFAULT)
    b = run_mode(states, 5, hg.STEP_DEFAULT | hg.STEP_PER_TERRAIN)
    c = run_mode(states, 5, hg.STEP_DEFAULT | hg.STEP_PER_TERRAIN | hg.STEP_PER_PASS)
    for i in range(len(SIZES)):
        for p in range(hg.NUM_STATE_PLANES):
            assert np.array_equal(a[i][p], b[i][p]), (i, p, "batched vs per-terrain fused")
            assert np.array_equal(a[i][p], c[i][p]), (i, p, "batched vs per-terrain per-pass")


def test_batched_tile_variants_agree(states):
    ref = run_mode(states, 3, hg.STEP_DEFAULT)
    nh, nt = len(hg.tile_variants(hg.KERNEL_HYDRO)), len(hg.tile_variants(hg.KERNEL_THERMAL))
    for hv, tv in [(v, 0) for v in range(nh)] + [(0, v) for v in range(nt)]:
        with make_batch(states) as b:
            b.set_tiles(hv, tv)
            b.step(3)
            for i in range(len(SIZES)):
                st = b.read_state(i)
                assert all(np.array_equal(st[p], ref[i][p]) for p in range(hg.NUM_STATE_PLANES)), (hv, tv, i)


def test_batch_autotune_restores_state(states):
    ref = run_mode(states, 3, hg.STEP_DEFAULT)
    nh, nt = len(hg.tile_variants(hg.KERNEL_HYDRO)), len(hg.tile_variants(hg.KERNEL_THERMAL))
    with make_batch(states) as b:
        b.capture(2)                    # the tuner drops a captured graph
        hv, tv = b.autotune(2)
        assert 0 <= hv < nh and 0 <= tv < nt
        assert b.step_count == 0
        with pytest.raises(hg.HGError):
            b.run()
        b.step(3)
        for i in range(len(SIZES)):
            st = b.read_state(i)
            assert all(np.array_equal(st[p], ref[i][p]) for p in range(hg.NUM_STATE_PLANES)), (hv, tv, i)


@pytest.mark.parametrize("persistent,per_terrain", [(False, False), (True, False), (False, True)])
def test_graph_replay_matches_stepping(states, persistent, per_terrain):
    eager = run_mode(states, 8, hg.STEP_DEFAULT)
    with make_batch(states) as b:
        b.set_persistent(persistent)
        b.capture(4, hg.STEP_DEFAULT | (hg.STEP_PER_TERRAIN if per_terrain else 0))
        b.run()
        b.sync()
        assert b.step_count == 4
        mid = [b.read_state(i) for i in range(len(SIZES))]   # eager work between replays
        b.run()
        b.sync()
        assert b.step_count == 8
        got = [b.read_state(i) for i in range(len(SIZES))]
    four = run_mode(states, 4, hg.STEP_DEFAULT)
    for i in range(len(SIZES)):
        for p in range(hg.NUM_STATE_PLANES):
            assert np.array_equal(mid[i][p], four[i][p]), (i, p, "first replay")
            assert np.array_equal(got[i][p], eager[i][p]), (i, p, "second replay")


def test_graph_replay_after_eager_steps(states):
    """The captured launches name the buffers current at capture: after an odd
    number of eager steps the replay is refused, after an even number it
    continues the step sequence as eager stepping does."""
    eager = run_mode(states, 4, hg.STEP_DEFAULT)
    with make_batch(states) as b:
        b.capture(2)
        b.step(1)
        with pytest.raises(hg.HGError) as e:
            b.run()
        assert e.value.code == hg.HG_ERR_STATE
        b.step(1)
        b.run()
        b.sync()
        assert b.step_count == 4
        got = [b.read_state(i) for i in range(len(SIZES))]
    for i in range(len(SIZES)):
        for p in range(hg.NUM_STATE_PLANES):
            assert np.array_equal(got[i][p], eager[i][p]), (i, p)


def test_batched_terrains_match_fp64_oracle(states):
    steps = 6
    got = run_mode(states, steps, hg.STEP_DEFAULT)
    for i, n in enumerate(SIZES):
        ref = states[i].astype(np.float64)
        O.run(ref, O.config_settings(**coeffs(i)), steps, O.STEP_DEFAULT, O.RAIN_UNIFORM, 0.02,
              seed=5, terrain=i, step0=0)
        h = got[i][hg.P_ROCK].astype(np.float64) + got[i][hg.P_DIRT]
        assert rel_l2(h, ref[O.ROCK] + ref[O.DIRT]) < 1e-5, i
        assert rel_l2(got[i][hg.P_WATER], ref[O.WATER]) < 1e-4, i
        s = got[i][hg.P_SED_ROCK].astype(np.float64) + got[i][hg.P_SED_DIRT]
        assert rel_l2(s, ref[O.S0] + ref[O.S1]) < 1e-4, i


def test_batch_init_and_errors():
    with hg.Batch([64, 128]) as b:
        b.init_fbm(1, hg.fbm_params(seed=3, base_freq=1 / 32, hi=10.0))
        r = b.read_plane(1, hg.P_ROCK)
        assert r.min() == 0.0 and r.max() == 10.0
        assert not b.read_plane(0, hg.P_ROCK).any()
        with pytest.raises(hg.HGError):
            b.capture(3)
        with pytest.raises(hg.HGError):
            b.run()
    with pytest.raises(hg.HGError):
        hg.Batch([64, 100])
