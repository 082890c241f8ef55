"""GPU: the batched benchmark's workload at full size.

64 terrains with sides drawn uniformly from {256, 384, ..., 2048} (the
benchmark's draw, 1.02e8 cells), each with its own fBm seed and coefficients
Kc ~ U(0.02, 0.1), Ks, Kd ~ U(0.01, 0.05), Ke ~ U(0.005, 0.02), talus ~
U(30, 45) deg; uniform Philox rain 0.02 with the terrain index in the
counter; thermal erosion and smoothing every step.  One launch per fused
kernel per step covers every terrain.  Checks: 4 batched steps (a captured
graph replayed twice) equal the single-terrain fused kernels launched
terrain by terrain, bit for bit on all 64 terrains; three terrains (the
smallest, a middle one and the largest) stay within the per-step tolerances
of the FP64 oracle over 6 steps."""
from __future__ import annotations

import sys
import os
import time

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O
from helpers import rel_l2

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from bench import batched_workload  # noqa: E402

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def make(sides, coeffs):
    b = hg.Batch(sides)
    b.set_rain(hg.RAIN_UNIFORM, amount=0.02, seed=7)
    for i, n in enumerate(sides):
        b.set_erosion(hg.erosion_settings(**coeffs[i]), terrain=i)
        b.init_fbm(i, hg.fbm_params(seed=1000 + i, base_freq=1.0 / 256, hi=float(n) / 8))
    return b


def test_batch_64_terrains_full_size(capsys):
    sides, coeffs = batched_workload()
    assert len(sides) == 64 and min(sides) >= 256 and max(sides) <= 2048
    with make(sides, coeffs) as b:
        start = [b.read_state(i) for i in (sides.index(min(sides)), 32, sides.index(max(sides)))]
        b.capture(2)
        b.run()
        b.run()
        b.sync()
        graph = [b.read_state(i) for i in range(64)]
    with make(sides, coeffs) as b:
        b.step(4, hg.STEP_DEFAULT | hg.STEP_PER_TERRAIN)
        per = [b.read_state(i) for i in range(64)]
    for i in range(64):
        for p in range(hg.NUM_STATE_PLANES):
            assert np.array_equal(graph[i][p], per[i][p]), (i, p)
    with make(sides, coeffs) as b:
        b.step(6)
        picks = (sides.index(min(sides)), 32, sides.index(max(sides)))
        got = [b.read_state(i) for i in picks]
    t0 = time.perf_counter()
    for st0, g, i in zip(start, got, picks):
        ref = O.new_state(sides[i])
        ref[: hg.NUM_STATE_PLANES] = st0
        O.run(ref, O.config_settings(**coeffs[i]), 6, O.STEP_DEFAULT, O.RAIN_UNIFORM, 0.02, seed=7,
              terrain=i, step0=0)
        h = g[hg.P_ROCK].astype(np.float64) + g[hg.P_DIRT]
        e = (rel_l2(h, ref[O.ROCK] + ref[O.DIRT]), rel_l2(g[hg.P_WATER], ref[O.WATER]),
             rel_l2(g[hg.P_SED_ROCK].astype(np.float64) + g[hg.P_SED_DIRT], ref[O.S0] + ref[O.S1]))
        with capsys.disabled():
            print(f"\nbatch terrain {i} ({sides[i]}^2) vs FP64 after 6 steps: height {e[0]:.2e} "
                  f"water {e[1]:.2e} sediment {e[2]:.2e}", flush=True)
        assert e[0] < 1e-5 and e[1] < 1e-4 and e[2] < 1e-4, (i, e)
    assert time.perf_counter() - t0 < 300
