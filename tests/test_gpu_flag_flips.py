"""GPU: how often the GPU's thermal and smoothing arithmetic flips a branch
flag relative to the FP32 oracle.

The GPU kernels restate three of the reference's decisions in cheaper but
not bitwise-equal FP32 forms: the talus test as D > tan(alpha) (diagonal:
D > sqrt(2) tan(alpha)) instead of atan(D / d) > alpha
(glsl/thermal_erosion.glsl:63-91), the smoothing mean as a pairwise sum times
0.2 instead of a left-to-right sum over 5.0, and the blend as
fma(m, mean - x, x) instead of m mean + (1 - m) x (glsl/smoothing.glsl:62-73,
97-103).  The oracle keeps the reference forms.  Near a threshold the two
can disagree; this test counts, per pass and on the same FP32 input, the
cells whose talus marks (outflow > 0 in a direction) or extremum flags
differ, next to the FP32-vs-FP64 oracle count."""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O
from helpers import C0, spun_up_state

pytestmark = pytest.mark.gpu

N = 256


@pytest.fixture(scope="module")
def state():
    return spun_up_state(N, 40, seed=3, base_freq=1 / 48)


def _gpu_pass(state, pass_id):
    with hg.Terrain(N) as t:
        t.set_erosion(hg.erosion_settings(**C0))
        t.write_state(state)
        t.run_pass(pass_id)
        return t.read_state()


def _oracle_pass(state, pass_id, dtype):
    st = state.astype(dtype)
    O.run_pass(st, O.config_settings(**C0), pass_id)
    return st


@pytest.mark.parametrize("pass_id", [hg.PASS_THERMAL_FLUX_0, hg.PASS_THERMAL_FLUX_1])
def test_talus_mark_flips(state, pass_id):
    g = _gpu_pass(state, pass_id)
    o32, o64 = _oracle_pass(state, pass_id, np.float32), _oracle_pass(state, pass_id, np.float64)
    planes = list(range(hg.P_THERM_L, hg.P_THERM_RB + 1))
    mg, m32, m64 = (np.stack([x[p] > 0 for p in planes]) for x in (g, o32, o64))
    flips, flips_64 = int((mg != m32).sum()), int((m32 != m64).sum())
    marked = int(m32.sum())
    print(f"pass {pass_id}: {marked} marked directions of {m32.size}; gpu vs fp32 oracle {flips} flips, "
          f"fp32 vs fp64 oracle {flips_64}", flush=True)
    assert marked > 0.05 * m32.size          # the test terrain has many talus-marked cells
    assert flips <= max(10, 2e-4 * m32.size)


def test_smoothing_flag_flips(state):
    """GPU flags are the cells the GPU pass changed (its blend returns the
    cell unchanged, bit for bit, when the flag is off); the oracle's come from
    its own extremum tests (orc_smooth_flags) on the FP32 and FP64 input.  A
    flagged cell whose blend m (mean - x) is below half an FP32 ulp of x is
    left unchanged by any FP32 blend, so it cannot show its flag: such cells
    are counted apart and not as flips."""
    g = _gpu_pass(state, hg.PASS_SMOOTH)
    f32, f64 = O.smooth_flags(state), O.smooth_flags(state.astype(np.float64))
    m = min(1.0, np.float32(C0["slippage"]) * np.float32(C0["dt"]))
    out = []
    for p, bit in ((hg.P_ROCK, 1), (hg.P_DIRT, 2)):
        x = state[p].astype(np.float64)
        mean = np.zeros_like(x)
        mean[1:-1, 1:-1] = (x[1:-1, 1:-1] + x[1:-1, :-2] + x[1:-1, 2:] + x[2:, 1:-1] + x[:-2, 1:-1]) / 5
        hidden = np.abs(m * (mean - x)) < 0.5 * np.spacing(state[p])
        fg, o32, o64 = g[p] != state[p], (f32 & bit) != 0, (f64 & bit) != 0
        flips = (fg != o32) & ~(o32 & hidden)
        out.append((int(o32.sum()), int(flips.sum()), int((o32 & hidden).sum()), int((o32 != o64).sum())))
    print(f"smoothing (rock, surface): flagged cells {[o[0] for o in out]} of {N * N}; gpu vs fp32 oracle "
          f"flips {[o[1] for o in out]} (flagged below FP32 resolution {[o[2] for o in out]}); fp32 vs fp64 "
          f"oracle {[o[3] for o in out]}", flush=True)
    for flagged, flips, _, _ in out:
        assert flagged > 0
        assert flips <= max(10, 1e-3 * N * N)
