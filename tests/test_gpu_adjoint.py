"""GPU adjoint (hg_adjoint_rollout) against the dual-number FP64 oracle."""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O

pytestmark = pytest.mark.gpu

COEF = np.array([0.05, 0.02, 0.02, 0.01])


def setup_case(n, seed=2, hi=12.0):
    h0 = O.fbm_init(n, O.fbm_params(seed=seed, base_freq=1.0 / 12, hi=hi), np.float32)
    tgt = O.fbm_init(n, O.fbm_params(seed=seed + 1, base_freq=1.0 / 12, hi=hi), np.float32)
    return h0, tgt


def gpu_grad(n, h0, tgt, steps, k_ckpt, flags=hg.STEP_DEFAULT):
    with hg.Terrain(n) as t:
        t.set_erosion(hg.erosion_settings(Kc=COEF[0], Ks=COEF[1], Kd=COEF[2], Ke=COEF[3]))
        t.set_rain(hg.RAIN_CONSTANT, amount=0.01)
        st = np.zeros((hg.NUM_STATE_PLANES, n, n), np.float32)
        st[hg.P_ROCK] = h0
        t.write_state(st)
        t.step = 0
        t.dispatch_grid(steps, flags)
        h = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
        final = t.read_state(hg.NUM_STATE_PLANES)
        g = (2.0 * (h - tgt) / h.size).astype(np.float32)
        t.write_state(st)
        t.step = 0
        gh, gc = t.adjoint_rollout(steps, k_ckpt, g, flags)
        after = t.read_state(hg.NUM_STATE_PLANES)
        assert t.step == steps
    return gh, gc, final, after


@pytest.mark.parametrize("flags", [hg.STEP_DEFAULT, hg.STEP_RAIN])
def test_adjoint_matches_dual_oracle(flags):
    n, steps = 40, 12
    h0, tgt = setup_case(n)
    gh, gc, _, _ = gpu_grad(n, h0, tgt, steps, 5, flags)
    st = O.new_state(n)
    st[O.ROCK] = h0
    rng = np.random.default_rng(0)
    cells = rng.choice(n * n, size=24, replace=False)
    loss, dc, dh = O.rollout_grad_dual(st, O.config_settings(), COEF, steps, tgt, cells,
                                       flags=flags, rain_mode=O.RAIN_CONSTANT, rain_amount=0.01)
    print("coef gpu", gc, "dual", dc)
    for j in range(4):
        if abs(dc[j]) < 1e-12:
            assert abs(gc[j]) < 1e-8, j
        else:
            assert abs(gc[j] - dc[j]) <= 2e-3 * abs(dc[j]), (j, gc[j], dc[j])
    scale = np.abs(gh).max()
    err = np.abs(gh.flat[cells] - dh)
    print("cells max err / max|g|", err.max() / scale)
    assert err.max() <= 2e-3 * scale


def test_checkpoint_interval_does_not_change_the_gradient():
    n, steps = 32, 9
    h0, tgt = setup_case(n, seed=5)
    ref_h, ref_c, _, _ = gpu_grad(n, h0, tgt, steps, 1)
    for k in (2, 4, 5, 9, 20):      # 5: two segments, both taped by the forward
        gh, gc, _, _ = gpu_grad(n, h0, tgt, steps, k)
        assert np.abs(gh - ref_h).max() <= 1e-5 * np.abs(ref_h).max(), k
        assert np.allclose(gc, ref_c, rtol=1e-5, atol=1e-12), k


def test_repeated_calls_reuse_the_workspace():
    """The workspace and recompute stream stay with the context: a second
    call of the same shape, a smaller one (reuse) and a larger one
    (reallocation) on one context give bit for bit what fresh contexts give."""
    n = 32
    h0, tgt = setup_case(n, seed=6)
    runs = [(9, 4), (9, 4), (5, 2), (9, 9)]
    fresh = [gpu_grad(n, h0, tgt, s, k)[:2] for s, k in runs]
    with hg.Terrain(n) as t:
        t.set_erosion(hg.erosion_settings(Kc=COEF[0], Ks=COEF[1], Kd=COEF[2], Ke=COEF[3]))
        t.set_rain(hg.RAIN_CONSTANT, amount=0.01)
        st = np.zeros((hg.NUM_STATE_PLANES, n, n), np.float32)
        st[hg.P_ROCK] = h0
        for (steps, k), (ref_h, ref_c) in zip(runs, fresh):
            t.write_state(st)
            t.step = 0
            t.dispatch_grid(steps)
            h = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
            g = (2.0 * (h - tgt) / h.size).astype(np.float32)
            t.write_state(st)
            t.step = 0
            gh, gc = t.adjoint_rollout(steps, k, g)
            assert np.array_equal(gh, ref_h), (steps, k)
            assert np.array_equal(gc, ref_c), (steps, k)


def test_adjoint_leaves_the_final_state():
    n, steps = 32, 7
    h0, tgt = setup_case(n, seed=7)
    _, _, final, after = gpu_grad(n, h0, tgt, steps, 3)
    for p in range(hg.NUM_STATE_PLANES):
        assert np.array_equal(final[p], after[p]), p


def test_adjoint_arguments_rejected():
    with hg.Terrain(32) as t:
        g = np.zeros((32, 32), np.float32)
        with pytest.raises(hg.HGError):
            t.adjoint_rollout(0, 4, g)
        with pytest.raises(hg.HGError):
            t.adjoint_rollout(4, 0, g)
        with pytest.raises(hg.HGError):
            t.adjoint_rollout(4, 2, g, hg.STEP_PER_PASS)


def test_adjoint_with_back_traces_beyond_one_cell():
    """Thin water under large old fluxes: back-traces of several cells, so the
    transport adjoint's scatter path (taps more than one cell away) runs."""
    n, steps = 40, 2
    h0, tgt = setup_case(n, seed=4)
    rng = np.random.default_rng(8)
    st = np.zeros((hg.NUM_STATE_PLANES, n, n), np.float32)
    st[hg.P_ROCK] = h0
    st[hg.P_WATER] = rng.uniform(0.0, 0.02, (n, n))
    # old outflow to the right only: the velocity (from old fluxes) is about
    # 2 / dt to the right, the inflow and outflow (K-scaled) keep the depth thin
    st[hg.P_FLUX_R] = rng.uniform(20.0, 60.0, (n, n))
    st[hg.P_FLUX_T] = rng.uniform(0.0, 5.0, (n, n))
    st[hg.P_SED_ROCK] = rng.uniform(0, 1, (n, n))
    st[hg.P_SED_DIRT] = rng.uniform(0, 1, (n, n))
    with hg.Terrain(n) as t:
        t.set_erosion(hg.erosion_settings(Kc=COEF[0], Ks=COEF[1], Kd=COEF[2], Ke=COEF[3]))
        t.set_rain(hg.RAIN_CONSTANT, amount=0.01)
        t.write_state(st)
        t.run_pass(hg.PASS_FLUX)
        vdt = np.abs(t.read_plane(hg.P_VEL_U)) * 0.005
        assert (vdt > 1.0).mean() > 0.2        # many taps are more than a cell away
        t.write_state(st)
        t.step = 0
        t.dispatch_grid(steps)
        h = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
        g = (2.0 * (h - tgt) / h.size).astype(np.float32)
        t.write_state(st)
        t.step = 0
        gh, gc = t.adjoint_rollout(steps, 1, g)
    ost = O.new_state(n)
    ost[: hg.NUM_STATE_PLANES] = st
    cells = rng.choice(n * n, size=24, replace=False)
    _, dc, dh = O.rollout_grad_dual(ost, O.config_settings(), COEF, steps, tgt, cells,
                                    rain_mode=O.RAIN_CONSTANT, rain_amount=0.01)
    print("coef gpu", gc, "dual", dc)
    for j in range(4):
        assert abs(gc[j] - dc[j]) <= 2e-3 * max(abs(dc[j]), 1e-9), (j, gc[j], dc[j])
    err = np.abs(gh.flat[cells] - dh)
    assert err.max() <= 2e-3 * np.abs(gh).max()


def _grad_with_fusion(n, st, g, steps, k_ckpt, fuse, flags=hg.STEP_DEFAULT):
    with hg.Terrain(n) as t:
        t.set_erosion(hg.erosion_settings(Kc=COEF[0], Ks=COEF[1], Kd=COEF[2], Ke=COEF[3]))
        t.set_rain(hg.RAIN_CONSTANT, amount=0.01)
        assert t.get_adjoint_fusion() == 15         # the default: every grouping
        t.set_adjoint_fusion(fuse)
        assert t.get_adjoint_fusion() == fuse
        t.write_state(st)
        t.step = 0
        return t.adjoint_rollout(steps, k_ckpt, g, flags)


@pytest.mark.parametrize("n,flags", [(8, hg.STEP_DEFAULT), (16, hg.STEP_DEFAULT), (40, hg.STEP_DEFAULT),
                                     (104, hg.STEP_DEFAULT), (136, hg.STEP_RAIN),
                                     (72, hg.STEP_RAIN | hg.STEP_THERMAL),
                                     (64, hg.STEP_RAIN | hg.STEP_SMOOTH)])
def test_fused_adjoint_kernels_equal_separate_kernels(n, flags):
    """The erosion + flux VJPs in one kernel, the smoothing + thermal layer-1
    VJPs in one kernel, the transport's far taps in the thermal layer-0
    launch and the transport's gather in the smoothing + layer-1 launch give
    the gradients of the separate kernels bit for bit, in every combination (sides that fill the 64 x 8 gather tiles partly, several
    segments, flag subsets)."""
    h0, tgt = setup_case(n, seed=9)
    st = np.zeros((hg.NUM_STATE_PLANES, n, n), np.float32)
    st[hg.P_ROCK] = h0
    rng = np.random.default_rng(3)
    st[hg.P_WATER] = rng.uniform(0.0, 0.5, (n, n))
    st[hg.P_SED_ROCK] = rng.uniform(0.0, 0.05, (n, n))
    g = rng.standard_normal((n, n)).astype(np.float32)
    print("case", n, flags, flush=True)
    gh0, gc0 = _grad_with_fusion(n, st, g, 7, 3, 0, flags)
    assert np.abs(gh0).max() > 0
    for fuse in range(1, 16):
        gh, gc = _grad_with_fusion(n, st, g, 7, 3, fuse, flags)
        assert np.array_equal(gh, gh0), fuse
        assert np.array_equal(gc, gc0), fuse


def test_fused_erosion_flux_adjoint_with_long_back_traces():
    """Back-traces of several cells (the transport's atomic scatter runs
    before the fused unit): both groupings agree."""
    n = 64
    h0, tgt = setup_case(n, seed=4)
    rng = np.random.default_rng(8)
    st = np.zeros((hg.NUM_STATE_PLANES, n, n), np.float32)
    st[hg.P_ROCK] = h0
    st[hg.P_WATER] = rng.uniform(0.0, 0.02, (n, n))
    st[hg.P_FLUX_R] = rng.uniform(20.0, 60.0, (n, n))
    st[hg.P_FLUX_T] = rng.uniform(0.0, 5.0, (n, n))
    st[hg.P_SED_ROCK] = rng.uniform(0, 1, (n, n))
    st[hg.P_SED_DIRT] = rng.uniform(0, 1, (n, n))
    g = rng.standard_normal((n, n)).astype(np.float32)
    gh1, gc1 = _grad_with_fusion(n, st, g, 2, 1, 15)
    gh0, gc0 = _grad_with_fusion(n, st, g, 2, 1, 0)
    # far taps are summed with float atomics, whose order is not fixed
    assert np.abs(gh1 - gh0).max() <= 1e-6 * np.abs(gh0).max()
    assert np.allclose(gc1, gc0, rtol=1e-6, atol=1e-12)


def test_adjoint_fusion_argument_rejected():
    with hg.Terrain(32) as t:
        with pytest.raises(hg.HGError):
            t.set_adjoint_fusion(16)
        with pytest.raises(hg.HGError):
            t.set_adjoint_fusion(-1)
        assert t.get_adjoint_fusion() == 15
