"""GPU: the one-kernel step (hg_set_sweep) against the two-kernel step and the
per-pass kernels, bit for bit.

The sweep runs every pass of a GRID step in one kernel, each pass a few rows
behind the one it reads; its per-cell arithmetic is the shared inline code,
so it must reproduce the reference pass order exactly.  Covered: grid sides
from 8 (a single partial strip) to 1000 (partial strips and segments), every
compiled sweep variant, every flag subset, back-traces that leave the
kernel's LDS window (the exact recompute path), and the library's choice
between the paths."""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg
from helpers import C0, spun_up_state

pytestmark = pytest.mark.gpu

SWEEPS = list(range(len(hg.tile_variants(hg.KERNEL_SWEEP))))


def _terrain(n):
    t = hg.Terrain(n, device=0)
    t.set_erosion(hg.erosion_settings(**C0))
    t.set_rain(hg.RAIN_UNIFORM, amount=0.02, seed=1)
    return t


def _run(state, steps, flags, sweep=None, per_pass=False):
    n = state.shape[-1]
    with _terrain(n) as t:
        t.set_sweep(-1 if sweep is None else sweep)
        t.write_state(state)
        t.step = 3
        t.dispatch_grid(steps, flags | (hg.STEP_PER_PASS if per_pass else 0))
        return t.read_state(hg.NUM_STATE_PLANES)


def _same(a, b, what):
    for p in range(hg.NUM_STATE_PLANES):
        diff = np.flatnonzero(a[p] != b[p])
        n = a.shape[-1]
        assert diff.size == 0, f"{what}: plane {p}: {diff.size} cells differ, first at {divmod(int(diff[0]), n)}"


@pytest.mark.parametrize("n", [8, 24, 64, 136, 200, 256, 520, 1000])
def test_sweep_equals_two_kernels_and_per_pass(n):
    st = spun_up_state(n, 20, seed=7, base_freq=1 / 64)
    print(f"sweep n={n}", flush=True)
    plain = _run(st, 4, hg.STEP_DEFAULT, per_pass=True)
    two = _run(st, 4, hg.STEP_DEFAULT)
    _same(two, plain, "two kernels vs per-pass")
    for v in SWEEPS:
        _same(_run(st, 4, hg.STEP_DEFAULT, sweep=v), plain, f"sweep {v} vs per-pass")


@pytest.mark.parametrize("flags", [0, hg.STEP_RAIN, hg.STEP_RAIN | hg.STEP_THERMAL, hg.STEP_RAIN | hg.STEP_SMOOTH,
                                   hg.STEP_THERMAL | hg.STEP_SMOOTH])
def test_sweep_flag_subsets(flags):
    st = spun_up_state(264, 10, seed=4, base_freq=1 / 64)
    plain = _run(st, 3, flags, per_pass=True)
    _same(_run(st, 3, flags, sweep=0), plain, f"flags {flags}")


def test_sweep_long_backtraces():
    """Thin water under large old fluxes: most back-traces leave the sweep's
    window of +-2 rows and take the exact recompute of the taps."""
    n = 256
    st = spun_up_state(n, 5, seed=9)
    rng = np.random.default_rng(3)
    st[hg.P_WATER] = rng.uniform(0.0, 0.02, (n, n)).astype(np.float32)
    for p in (hg.P_FLUX_L, hg.P_FLUX_R, hg.P_FLUX_T, hg.P_FLUX_B):
        st[p] = rng.uniform(0.0, 60.0, (n, n)).astype(np.float32)
    st[hg.P_SED_ROCK] = rng.uniform(0, 1, (n, n)).astype(np.float32)
    st[hg.P_SED_DIRT] = rng.uniform(0, 1, (n, n)).astype(np.float32)
    plain = _run(st, 1, hg.STEP_DEFAULT, per_pass=True)
    for v in SWEEPS:
        _same(_run(st, 1, hg.STEP_DEFAULT, sweep=v), plain, f"sweep {v}")


def test_sweep_selection_and_timing():
    with _terrain(256) as t:
        t.init_fbm(hg.fbm_params(base_freq=1 / 64))
        t.set_timing(True)
        t.set_sweep(0)
        assert t.get_sweep() == 0
        t.dispatch_grid(3)
        t.set_sweep(-1)
        t.dispatch_grid(2)
        # velocity kept: always the two kernels
        t.set_sweep(0)
        t.dispatch_grid(1, hg.STEP_DEFAULT | hg.STEP_KEEP_VELOCITY)
        st = t.kernel_stats()
        assert st["step_sweep"][0] == 3
        assert st["fused_hydro"][0] == 3 and st["fused_thermal"][0] == 3
        with pytest.raises(hg.HGError):
            t.set_sweep(len(SWEEPS))
        with pytest.raises(hg.HGError):
            t.set_sweep(-2)
