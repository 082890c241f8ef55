"""GPU parity at the headline size (4096 x 4096) through size-independent
properties: the FP64 oracle cannot run many steps of a 16.7 M-cell grid in
test time, so the checks here are laws every pass obeys whatever the grid:
fused == per-pass bit for bit, the rain counter stream bit for bit, the water
identity and wall/K-scaling bounds of the flux pass, per-cell material
conservation of erosion, per-layer mass conservation of thermal erosion, and
the transport pass's constant-field and evaporation behaviour.  The state is
the benchmark's: simplex fBm [0, 512] (seed 1, base frequency 1/1024), uniform
rain 0.02, 20 GRID steps of spin-up on the GPU.
"""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O
from helpers import C0

pytestmark = pytest.mark.gpu

N = 4096
EPS = float(np.finfo(np.float32).eps)


@pytest.fixture(scope="module")
def c1():
    print("4096^2 spin-up", flush=True)
    t = hg.Terrain(N, device=0)
    t.set_erosion(hg.erosion_settings(**C0))
    t.set_rain(hg.RAIN_UNIFORM, amount=0.02, seed=1)
    t.init_fbm(hg.fbm_params(seed=1, base_freq=1.0 / 1024, lo=0.0, hi=512.0))
    t.dispatch_grid(20)
    state = t.read_state(hg.NUM_STATE_PLANES)
    yield t, state, t.step
    t.close()


def _reset(c1):
    t, state, step = c1
    t.write_state(state)
    t.step = step
    return t, state


@pytest.mark.parametrize("sweep", [-1] + list(range(len(hg.tile_variants(hg.KERNEL_SWEEP)))))
def test_fused_equals_per_pass_4096(c1, sweep):
    """At this size the fused kernels (the two-kernel step, sweep = -1, and the
    one-kernel step's variants) write their outputs with non-temporal stores
    and the per-pass kernels with ordinary ones: same bits."""
    t, state = _reset(c1)
    print(f"fused 2 steps, sweep {sweep}", flush=True)
    t.set_sweep(sweep)
    t.dispatch_grid(2)
    fused = t.read_state(9)
    t, _ = _reset(c1)
    print("per-pass 2 steps", flush=True)
    t.dispatch_grid(2, hg.STEP_DEFAULT | hg.STEP_PER_PASS)
    plain = t.read_state(9)
    for p in range(9):
        diff = np.count_nonzero(fused[p] != plain[p])
        assert diff == 0, f"plane {p}: {diff} cells differ"
    assert not np.array_equal(fused[hg.P_ROCK], state[hg.P_ROCK])


def test_rain_stream_bit_exact_4096(c1):
    t, state = _reset(c1)
    step = t.step
    t.run_pass(hg.PASS_RAIN)
    got = t.read_plane(hg.P_WATER)
    amount = np.float32(0.02)
    for y in (0, 1, 2047, 4095):
        u = np.array([O.rain_u(1, step, 0, x, y) for x in range(N)], dtype=np.float32)
        want = state[hg.P_WATER][y] + amount * u         # FP32 product, then FP32 sum
        assert np.array_equal(got[y], want), f"row {y}"
    assert got.sum(dtype=np.float64) > state[hg.P_WATER].sum(dtype=np.float64)


def test_flux_identity_and_bounds_4096(c1):
    t, state = _reset(c1)
    t.run_pass(hg.PASS_FLUX)
    out = t.read_state(hg.NUM_STATE_PLANES)
    dt = np.float32(0.005)
    f = out[[hg.P_FLUX_L, hg.P_FLUX_R, hg.P_FLUX_T, hg.P_FLUX_B]].astype(np.float64)
    b = state[hg.P_WATER].astype(np.float64)
    assert (f >= 0).all()
    assert not f[0][:, 0].any() and not f[1][:, -1].any() and not f[3][0, :].any() and not f[2][-1, :].any()
    out_sum = f.sum(0)
    assert (out_sum * float(dt) <= b * (1 + 8 * EPS) + 1e-30).all()
    old = state[[hg.P_FLUX_L, hg.P_FLUX_R, hg.P_FLUX_T, hg.P_FLUX_B]].astype(np.float64)
    inflow = np.zeros_like(b)
    inflow[:, 1:] += old[1][:, :-1]
    inflow[:, :-1] += old[0][:, 1:]
    inflow[:-1, :] += old[3][1:, :]
    inflow[1:, :] += old[2][:-1, :]
    want = np.maximum(0.0, b + float(dt) * (inflow - out_sum))
    err = np.abs(out[hg.P_WATER] - want)
    tol = 16 * EPS * (b + float(dt) * (inflow + out_sum)) + 1e-30
    assert (err <= tol).mean() > 1 - 1e-6, f"{(err > tol).sum()} cells break the water identity"


def test_erosion_conserves_material_4096(c1):
    t, state = _reset(c1)
    t.run_pass(hg.PASS_FLUX)
    mid = t.read_state(hg.NUM_STATE_PLANES)
    t.run_pass(hg.PASS_EROSION)
    out = t.read_state(hg.NUM_STATE_PLANES)
    planes = [hg.P_ROCK, hg.P_DIRT, hg.P_SED_ROCK, hg.P_SED_DIRT]
    before = sum(mid[p].astype(np.float64) for p in planes)
    after = sum(out[p].astype(np.float64) for p in planes)
    mag = sum(np.abs(mid[p]).astype(np.float64) for p in planes)
    assert (np.abs(after - before) <= 8 * EPS * mag).all()
    assert (out[hg.P_ROCK] >= 0).all() and (out[hg.P_DIRT] >= 0).all()
    assert np.count_nonzero(out[hg.P_SED_DIRT] != mid[hg.P_SED_DIRT]) > 0


@pytest.mark.parametrize("layer", [0, 1])
def test_thermal_conserves_each_layer_4096(c1, layer):
    t, state = _reset(c1)
    state = state.copy()
    if layer == 1:      # the workload has no dirt yet: lay a smooth dirt blanket
        x = np.arange(N, dtype=np.float32)
        state[hg.P_DIRT] = 2.0 * (1.0 + np.sin(x / 37.0)[None, :] * np.cos(x / 53.0)[:, None])
        t.write_plane(hg.P_DIRT, state[hg.P_DIRT])
    first = hg.PASS_THERMAL_FLUX_0 if layer == 0 else hg.PASS_THERMAL_FLUX_1
    t.run_pass(first)
    t.run_pass(first + 1)
    out = t.read_state(2)
    plane, other = (hg.P_ROCK, hg.P_DIRT) if layer == 0 else (hg.P_DIRT, hg.P_ROCK)
    s0 = state[plane].sum(dtype=np.float64)
    s1 = out[plane].sum(dtype=np.float64)
    moved = np.abs(out[plane].astype(np.float64) - state[plane]).sum()
    print(f"layer {layer}: moved {moved:.4g}, sum change {s1 - s0:.4g} of {s0:.4g}", flush=True)
    assert moved > 0
    # only FP32 rounding of each cell's new height changes the total
    assert abs(s1 - s0) <= 2e-7 * abs(s0)
    assert np.array_equal(out[other], state[other])


def test_transport_constant_field_and_evaporation_4096(c1):
    t, state = _reset(c1)
    t.run_pass(hg.PASS_FLUX)
    mid = t.read_state(hg.NUM_PLANES)
    mid[hg.P_SED_ROCK] = 0.25
    t.write_plane(hg.P_SED_ROCK, mid[hg.P_SED_ROCK])
    t.run_pass(hg.PASS_TRANSPORT)
    out = t.read_state(hg.NUM_STATE_PLANES)
    assert np.abs(out[hg.P_SED_ROCK] - np.float32(0.25)).max() <= 4 * EPS * 0.25
    keep = np.float32(1.0) - np.float32(0.01) * np.float32(0.005)
    assert np.array_equal(out[hg.P_WATER], mid[hg.P_WATER] * keep)
    assert np.abs(mid[hg.P_VEL_U]).max() > 1.0          # the back-traces do move
