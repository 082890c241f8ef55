"""GPU reference heightmap generator (hg_gen_heightmap, glsl/heightmap.glsl)
against the oracle: bit for bit with its FP32 restatement, the exp mask
included (both run the same FP32 exp method, which the oracle pins against
FP64 exp), and close to the FP64 restatement."""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O

pytestmark = pytest.mark.gpu

NO_EXP = dict(domain_warp=2, uplift=1, mask_round=1, mask_exp=0, mask_power=1, mask_slope=1, terrace=3,
              seed=17.25, scale=0.01, octaves=6)


def _gpu(n, kw):
    with hg.Terrain(n, device=0) as t:
        t.step = 5
        t.gen_heightmap(hg.map_settings(**kw))
        st = t.read_state(hg.NUM_STATE_PLANES)
        assert t.step == 0
    return st


@pytest.mark.parametrize("n", [64, 200])
def test_heightmap_bitwise_without_exp(n):
    print(f"heightmap n={n}", flush=True)
    st = _gpu(n, NO_EXP)
    r32, d32 = O.gen_heightmap(n, O.map_settings(**NO_EXP), np.float32)
    assert np.array_equal(st[hg.P_ROCK], r32), f"{np.count_nonzero(st[hg.P_ROCK] != r32)} rock cells differ"
    assert np.array_equal(st[hg.P_DIRT], d32)
    assert not st[2:].any()
    assert r32.std() > 0 and d32.std() > 0


@pytest.mark.parametrize("n", [128, 200])
def test_heightmap_reference_defaults(n):
    """The reference's default map settings, exp mask on."""
    st = _gpu(n, {})
    r32, d32 = O.gen_heightmap(n, O.map_settings(), np.float32)
    r64, d64 = O.gen_heightmap(n, O.map_settings(seed=0.0), np.float64)
    assert O.map_settings().mask_exp != 0
    assert np.array_equal(st[hg.P_DIRT], d32)
    assert np.array_equal(st[hg.P_ROCK], r32), f"{np.count_nonzero(st[hg.P_ROCK] != r32)} rock cells differ"
    st0 = _gpu(n, dict(seed=0.0))
    assert np.abs(st0[hg.P_ROCK] - r64).max() <= 1e-4 * np.abs(r64).max()
    assert np.abs(st0[hg.P_DIRT] - d64).max() <= 1e-5 * np.abs(d64).max()


def test_heightmap_settings_rejected():
    with hg.Terrain(32, device=0) as t:
        for bad in (dict(octaves=31), dict(domain_warp=3), dict(uplift=1, uplift_scale=0.0),
                    dict(scale=float("nan"))):
            with pytest.raises(hg.HGError):
                t.gen_heightmap(hg.map_settings(**bad))


def test_height_file_round_trip(tmp_path):
    n = 64
    path = str(tmp_path / "h.npy")
    with hg.Terrain(n, device=0) as t:
        t.gen_heightmap(hg.map_settings(seed=3.0))
        t.dispatch_grid(5)
        want = t.read(hg.CH_HEIGHT)
        t.save_height(path)
    a = np.load(path)
    assert a.shape == (n, n, 4) and a.dtype == np.float32
    assert np.array_equal(a, want)
    with hg.Terrain(n, device=0) as t2:
        t2.load_height(path)
        assert np.array_equal(t2.read(hg.CH_HEIGHT), want)
    with hg.Terrain(32, device=0) as t3:
        with pytest.raises(hg.HGError):
            t3.load_height(path)                      # wrong size
        with pytest.raises(hg.HGError):
            t3.load_height(str(tmp_path / "missing.npy"))
