"""GPU reference rain (HG_RAIN_REFERENCE, glsl/rain.glsl:32-56) against the
oracle: bit for bit with the FP32 restatement (same operation order, FP32
literals), close to the FP64 one, and the reference's every-`period`-steps
cadence inside hg_step_grid."""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O
from helpers import C0, spun_up_state

pytestmark = pytest.mark.gpu


def _terrain(n, rs, time, max_h):
    t = hg.Terrain(n, device=0)
    t.set_erosion(hg.erosion_settings(**C0))
    t.set_rain(hg.RAIN_REFERENCE, settings=rs)
    t.set_rain_time(time, max_h)
    return t


@pytest.mark.parametrize("n,time", [(96, 0.0), (96, 37.25), (512, 1234.5678)])
def test_reference_rain_pass_matches_oracle(n, time):
    st = spun_up_state(n, 4, seed=3)
    max_h = 64.0
    st[hg.P_ROCK, : n // 4, : n // 4] += 40.0                  # a mountain above 0.55 * max_h
    rs = hg.rain_settings(amount=0.01, mountain_thresh=0.55, mountain_multip=0.05, period=512, drops=0.02)
    print(f"reference rain n={n} time={time}", flush=True)
    with _terrain(n, rs, time, max_h) as t:
        t.write_state(st[: hg.NUM_STATE_PLANES])
        t.dispatch_grid_rain()
        got = t.read_plane(hg.P_WATER)
    ors = O.rain_settings(amount=0.01, mountain_thresh=0.55, mountain_multip=0.05, period=512, drops=0.02)
    ref32 = O.rain_reference(st.copy(), ors, time, max_h)[O.WATER]
    assert np.array_equal(got, ref32), f"{np.count_nonzero(got != ref32)} cells differ from the FP32 oracle"
    assert (got > st[hg.P_WATER]).mean() > 0.2                # it rained on a good share of the map
    if time == 0.0:         # seed fract(0) * 1000 = 0: small noise coordinates, FP64 agrees too
        ref64 = O.rain_reference(st.astype(np.float64), ors, time, max_h)[O.WATER]
        assert np.abs(got - ref64).max() <= 1e-6 * max(1.0, np.abs(ref64).max())


def test_reference_rain_cadence_in_stepping():
    n = 64
    st = spun_up_state(n, 4, seed=5)
    rs = hg.rain_settings(amount=0.5, period=3, drops=0.05)
    with _terrain(n, rs, 3.5, 64.0) as t:
        t.write_state(st[: hg.NUM_STATE_PLANES])
        t.step = 0
        t.dispatch_grid(7, hg.STEP_DEFAULT)
        auto = t.read_state(hg.NUM_STATE_PLANES)
        t.write_state(st[: hg.NUM_STATE_PLANES])
        t.step = 0
        for s in range(7):
            if (s + 1) % 3 == 0:                              # src/main.cpp:313-319
                t.dispatch_grid_rain()
            t.dispatch_grid(1, hg.STEP_THERMAL | hg.STEP_SMOOTH)
        manual = t.read_state(hg.NUM_STATE_PLANES)
        t.write_state(st[: hg.NUM_STATE_PLANES])
        t.step = 0
        t.dispatch_grid(7, hg.STEP_THERMAL | hg.STEP_SMOOTH)
        dry = t.read_state(hg.NUM_STATE_PLANES)
    for p in range(hg.NUM_STATE_PLANES):
        assert np.array_equal(auto[p], manual[p]), p
    assert not np.array_equal(auto[hg.P_WATER], dry[hg.P_WATER])


def test_reference_rain_rejected_where_unsupported():
    rs = hg.rain_settings()
    with _terrain(32, rs, 1.0, 64.0) as t:
        with pytest.raises(hg.HGError):
            t.adjoint_rollout(2, 1, np.zeros((32, 32), np.float32), hg.STEP_DEFAULT)
        with pytest.raises(hg.HGError):
            t.set_rain_time(1.0, 0.0)
        with pytest.raises(hg.HGError):
            t.set_rain(hg.RAIN_REFERENCE, settings=hg.rain_settings(mountain_thresh=1.0))
    with hg.Batch([32, 64]) as b:
        with pytest.raises(hg.HGError):
            b.set_rain(hg.RAIN_REFERENCE)
