"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Pass-level: each reference pass from the same FP32 state, GPU vs the FP64
oracle.  Step-level: fused vs per-pass kernels bit for bit, and multi-step
runs against the FP64 oracle in relative L2 (FP32 rounding makes single cells
drift chaotically over hundreds of steps even for the FP32 build of the
oracle itself; the max-abs error is reported, not gated).
"""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O
from helpers import C0, close_fraction, rel_l2, spun_up_state

pytestmark = pytest.mark.gpu

PASS_PLANES = {
    hg.PASS_RAIN: [hg.P_WATER],
    hg.PASS_FLUX: [hg.P_WATER, hg.P_FLUX_L, hg.P_FLUX_R, hg.P_FLUX_T, hg.P_FLUX_B, hg.P_VEL_U,
                   hg.P_VEL_V, hg.P_VEL_DEPTH],
    hg.PASS_EROSION: [hg.P_ROCK, hg.P_DIRT, hg.P_SED_ROCK, hg.P_SED_DIRT],
    hg.PASS_TRANSPORT: [hg.P_SED_ROCK, hg.P_SED_DIRT, hg.P_WATER],
    hg.PASS_THERMAL_FLUX_0: list(range(hg.P_THERM_L, hg.P_THERM_RB + 1)),
    hg.PASS_THERMAL_TRANSPORT_0: [hg.P_ROCK],
    hg.PASS_THERMAL_FLUX_1: list(range(hg.P_THERM_L, hg.P_THERM_RB + 1)),
    hg.PASS_THERMAL_TRANSPORT_1: [hg.P_DIRT],
    hg.PASS_SMOOTH: [hg.P_ROCK, hg.P_DIRT],
}


@pytest.fixture(scope="module")
def state96():
    return spun_up_state(96, 40)


def _terrain(n, settings=None, rain_mode=hg.RAIN_UNIFORM, amount=0.02, seed=1):
    t = hg.Terrain(n, device=0)
    t.set_erosion(settings if settings is not None else hg.erosion_settings(**C0))
    t.set_rain(rain_mode, amount=amount, seed=seed)
    return t


@pytest.mark.parametrize("pass_id", list(range(hg.NUM_PASSES)))
def test_pass_matches_oracle(state96, pass_id):
    n = state96.shape[-1]
    print(f"pass {pass_id} n={n}", flush=True)
    with _terrain(n) as t:
        t.write_state(state96)
        t.step = 40
        t.run_pass(pass_id)
        got = t.read_state()
    ref = state96.astype(np.float64)
    O.run_pass(ref, O.config_settings(**C0), pass_id, O.RAIN_UNIFORM, 0.02, seed=1, step=40)
    for p in PASS_PLANES[pass_id]:
        scale = max(float(np.abs(ref[p]).max()), 1e-6)
        frac = close_fraction(got[p], ref[p], rtol=2e-5, atol=2e-6 * scale)
        assert frac <= 1e-3, f"pass {pass_id} plane {p}: {frac:.2e} of cells outside tolerance"
        assert rel_l2(got[p], ref[p]) < 1e-5 or scale < 1e-5, f"pass {pass_id} plane {p}"
    # planes a pass does not produce are left untouched
    for p in range(hg.NUM_STATE_PLANES):
        if p not in PASS_PLANES[pass_id]:
            assert np.array_equal(got[p], state96[p]), f"pass {pass_id} touched plane {p}"


def test_init_fbm_matches_oracle():
    n = 256
    p = hg.fbm_params(seed=11, base_freq=1.0 / 64, hi=64.0)
    with hg.Terrain(n) as t:
        t.init_fbm(p)
        rock = t.read_plane(hg.P_ROCK)
        rest = t.read_state(hg.NUM_STATE_PLANES)[1:]
    op = O.fbm_params(seed=11, base_freq=1.0 / 64, hi=64.0)
    r32 = O.fbm_init(n, op, np.float32)
    r64 = O.fbm_init(n, op, np.float64)
    ulp = np.abs(rock.view(np.int32).astype(np.int64) - r32.view(np.int32).astype(np.int64))
    print(f"fbm init vs FP32 oracle: max {ulp.max()} ulp, {np.count_nonzero(ulp)} cells differ", flush=True)
    assert ulp.max() == 0, f"init differs from the FP32 oracle by up to {ulp.max()} ulp"
    assert np.abs(rock - r64).max() < 64 * 2e-5
    assert rock.min() == 0.0 and rock.max() == 64.0
    assert not rest.any()


def _run_pair(state, steps, flags, rain_mode=hg.RAIN_UNIFORM, settings=None):
    n = state.shape[-1]
    out = []
    for extra in (0, hg.STEP_PER_PASS):
        with _terrain(n, settings=settings, rain_mode=rain_mode) as t:
            t.write_state(state)
            t.step = 3
            t.dispatch_grid(steps, flags | extra)
            out.append(t.read_state(hg.NUM_STATE_PLANES))
    return out


@pytest.mark.parametrize("n", [8, 24, 64, 200, 256])
def test_fused_equals_per_pass_bitwise(n):
    st = spun_up_state(n, 30, seed=5)
    print(f"fused vs per-pass n={n}", flush=True)
    fused, plain = _run_pair(st, 6, hg.STEP_DEFAULT)
    for p in range(hg.NUM_STATE_PLANES):
        diff = np.flatnonzero(fused[p] != plain[p])
        assert diff.size == 0, f"plane {p}: {diff.size} cells differ, first at {divmod(int(diff[0]), n)}"


def test_fused_equals_per_pass_long_backtrace():
    """Thin water under large old fluxes gives |v|*dt of many cells, so most
    back-traces leave the kernel's halo and take the exact recompute path."""
    n = 128
    st = spun_up_state(n, 5, seed=9)
    rng = np.random.default_rng(3)
    st[hg.P_WATER] = rng.uniform(0.0, 0.02, (n, n)).astype(np.float32)
    for p in (hg.P_FLUX_L, hg.P_FLUX_R, hg.P_FLUX_T, hg.P_FLUX_B):
        st[p] = rng.uniform(0.0, 60.0, (n, n)).astype(np.float32)
    st[hg.P_SED_ROCK] = rng.uniform(0, 1, (n, n)).astype(np.float32)
    st[hg.P_SED_DIRT] = rng.uniform(0, 1, (n, n)).astype(np.float32)
    fused, plain = _run_pair(st, 1, hg.STEP_DEFAULT)
    for p in range(hg.NUM_STATE_PLANES):
        assert np.array_equal(fused[p], plain[p]), f"plane {p}"
    # and the one step matches the FP64 oracle
    ref = st.astype(np.float64)
    O.run(ref, O.config_settings(**C0), 1, O.STEP_DEFAULT, O.RAIN_UNIFORM, 0.02, seed=1, step0=3)
    for p in (hg.P_SED_ROCK, hg.P_SED_DIRT, hg.P_ROCK, hg.P_DIRT):
        assert rel_l2(fused[p], ref[p]) < 2e-5, p


@pytest.mark.parametrize("n", [200, 256])
def test_fused_tile_variants_equal_per_pass(n):
    """Every compiled tile variant of both fused kernels (including the
    double-buffered thermal ones) gives the per-pass result bit for bit, with
    thermal erosion on and off."""
    st = spun_up_state(n, 20, seed=7)
    nh, nt = len(hg.tile_variants(hg.KERNEL_HYDRO)), len(hg.tile_variants(hg.KERNEL_THERMAL))
    for flags in (hg.STEP_DEFAULT, hg.STEP_RAIN | hg.STEP_SMOOTH):
        with _terrain(n) as t:
            t.write_state(st)
            t.step = 3
            t.dispatch_grid(3, flags | hg.STEP_PER_PASS)
            plain = t.read_state(hg.NUM_STATE_PLANES)
        for hv, tv in [(v, 0) for v in range(nh)] + [(0, v) for v in range(1, nt)]:
            with _terrain(n) as t:
                t.set_sweep(-1)
                t.set_tiles(hv, tv)
                t.write_state(st)
                t.step = 3
                t.dispatch_grid(3, flags)
                got = t.read_state(hg.NUM_STATE_PLANES)
            for p in range(hg.NUM_STATE_PLANES):
                assert np.array_equal(got[p], plain[p]), (n, flags, hv, tv, p)


@pytest.mark.parametrize("flags", [hg.STEP_RAIN, hg.STEP_RAIN | hg.STEP_THERMAL,
                                   hg.STEP_RAIN | hg.STEP_SMOOTH, hg.STEP_THERMAL | hg.STEP_SMOOTH])
def test_fused_equals_per_pass_flag_subsets(flags):
    st = spun_up_state(128, 20, seed=2)
    fused, plain = _run_pair(st, 3, flags)
    for p in range(hg.NUM_STATE_PLANES):
        assert np.array_equal(fused[p], plain[p]), f"flags {flags} plane {p}"


def test_multistep_vs_fp64_oracle_128():
    n, steps = 128, 200
    p = hg.fbm_params(seed=4, base_freq=1.0 / 64, hi=64.0)
    with _terrain(n, rain_mode=hg.RAIN_CONSTANT, amount=0.01) as t:
        t.init_fbm(p)
        init = t.read_state()
        t.dispatch_grid(steps, hg.STEP_DEFAULT)
        got = t.read_state(hg.NUM_STATE_PLANES)
    ref = init.astype(np.float64)
    O.run(ref, O.config_settings(**C0), steps, O.STEP_DEFAULT, O.RAIN_CONSTANT, 0.01)
    h_ref = ref[O.ROCK] + ref[O.DIRT]
    h_got = got[hg.P_ROCK].astype(np.float64) + got[hg.P_DIRT]
    r = rel_l2(h_got, h_ref)
    print(f"128^2 x {steps}: height rel-L2 {r:.2e}, max abs {np.abs(h_got - h_ref).max():.2e}")
    assert r < 1e-4
    assert rel_l2(got[hg.P_WATER], ref[O.WATER]) < 1e-3
    sed_got = got[hg.P_SED_ROCK].astype(np.float64) + got[hg.P_SED_DIRT]
    assert rel_l2(sed_got, ref[O.S0] + ref[O.S1]) < 1e-3


@pytest.mark.slow
def test_config0_512_500_steps_vs_fp64_oracle():
    """Config 0: 512^2, fBm (6 octaves, f=1/256) to [0,64], rain 0.01/step,
    500 steps, heights/water/sediment compared every 100 steps.

    Free-running gates (GPU FP32 vs FP64 oracle from the same initial state):
    height rel-L2 < 1e-4, water < 1e-3, sediment < 1e-2.  Sediment gets the
    loose bound because cells with thin water have velocities ~ flux/depth, so
    their semi-Lagrangian back-trace point is ill-conditioned; rounding alone
    moves it by cells (the FP32 build of the oracle shows the same excursions).
    Shadowed gates isolate per-step arithmetic from that amplification: at every
    checkpoint the oracle restarts from the GPU's own state and both run 25
    steps, which must agree to 1e-5 (heights), 1e-4 (water, sediment)."""
    n = 512
    s = O.config_settings(**C0)
    with _terrain(n, rain_mode=hg.RAIN_CONSTANT, amount=0.01) as t:
        t.init_fbm(hg.fbm_params(seed=1, base_freq=1.0 / 256, hi=64.0))
        ref = t.read_state().astype(np.float64)
        for k in range(5):
            t.dispatch_grid(100, hg.STEP_DEFAULT)
            got = t.read_state(hg.NUM_STATE_PLANES)
            O.run(ref, s, 100, O.STEP_DEFAULT, O.RAIN_CONSTANT, 0.01, step0=100 * k)
            h_ref = ref[O.ROCK] + ref[O.DIRT]
            h_got = got[hg.P_ROCK].astype(np.float64) + got[hg.P_DIRT]
            r = rel_l2(h_got, h_ref)
            rw = rel_l2(got[hg.P_WATER], ref[O.WATER])
            rs = rel_l2(got[hg.P_SED_ROCK].astype(np.float64) + got[hg.P_SED_DIRT], ref[O.S0] + ref[O.S1])
            # shadow: oracle from the GPU state, 25 more steps on both
            sh = O.new_state(n)
            sh[: hg.NUM_STATE_PLANES] = got
            O.run(sh, s, 25, O.STEP_DEFAULT, O.RAIN_CONSTANT, 0.01, step0=100 * (k + 1))
            with _terrain(n, rain_mode=hg.RAIN_CONSTANT, amount=0.01) as t2:
                t2.write_state(got)
                t2.step = 100 * (k + 1)
                t2.dispatch_grid(25, hg.STEP_DEFAULT)
                g2 = t2.read_state(hg.NUM_STATE_PLANES)
            sr = rel_l2(g2[hg.P_ROCK].astype(np.float64) + g2[hg.P_DIRT], sh[O.ROCK] + sh[O.DIRT])
            sw = rel_l2(g2[hg.P_WATER], sh[O.WATER])
            ss = rel_l2(g2[hg.P_SED_ROCK].astype(np.float64) + g2[hg.P_SED_DIRT], sh[O.S0] + sh[O.S1])
            print(f"step {100 * (k + 1)}: free height {r:.2e} water {rw:.2e} sediment {rs:.2e} | "
                  f"shadow height {sr:.2e} water {sw:.2e} sediment {ss:.2e}", flush=True)
            assert r < 1e-4 and rw < 1e-3 and rs < 1e-2
            assert sr < 1e-5 and sw < 1e-4 and ss < 1e-4


def test_channel_io_roundtrip():
    n = 64
    rng = np.random.default_rng(0)
    with hg.Terrain(n) as t:
        for ch in (hg.CH_HEIGHT, hg.CH_FLUX, hg.CH_VELOCITY, hg.CH_SEDIMENT):
            a = rng.uniform(0, 10, (n, n, 4)).astype(np.float32)
            t.write(ch, a)
            b = t.read(ch)
            if ch == hg.CH_HEIGHT:
                assert np.array_equal(b[..., :3], a[..., :3])
                assert np.allclose(b[..., 3], a[..., 0] + a[..., 1] + a[..., 2], rtol=1e-6)
            elif ch == hg.CH_FLUX:
                assert np.array_equal(b, a)
            elif ch == hg.CH_VELOCITY:
                assert np.array_equal(b[..., :3], a[..., :3]) and not b[..., 3].any()
            else:
                assert np.array_equal(b[..., :2], a[..., :2]) and not b[..., 2:].any()
        # plane view agrees with the channel view
        h = t.read(hg.CH_HEIGHT)
        assert np.array_equal(t.read_plane(hg.P_WATER), h[..., 2])


def test_errors_are_reported():
    with pytest.raises(hg.HGError) as e:
        hg.Terrain(100)
    assert e.value.code == hg.HG_ERR_INVALID and "multiple of 8" in str(e.value)
    with hg.Terrain(64) as t:
        bad = hg.erosion_settings()
        bad.d_t = float("nan")
        with pytest.raises(hg.HGError) as e:
            t.set_erosion(bad)
        assert e.value.code == hg.HG_ERR_INVALID
        bad = hg.erosion_settings(Kc=-1.0)
        with pytest.raises(hg.HGError):
            t.set_erosion(bad)
        with pytest.raises(hg.HGError):
            t.run_pass(99)
        with pytest.raises(hg.HGError):
            t.dispatch_grid(1, 0x1000)
        # the context still works after errors
        t.set_erosion(hg.erosion_settings())
        t.dispatch_grid(2)
        assert t.step == 2


def test_timing_counts_launches():
    with _terrain(128) as t:
        t.init_fbm(hg.fbm_params(base_freq=1 / 64))
        t.set_timing(True)
        t.dispatch_grid(10)
        st = t.kernel_stats()
        assert st["fused_hydro"][0] == 10 and st["fused_thermal"][0] == 10
        assert st["fused_hydro"][1] > 0
