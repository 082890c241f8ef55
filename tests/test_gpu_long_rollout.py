"""GPU: the adjoint on the differentiable-rollout workload at full size.

1024 x 1024 fBm terrain (seed 11, [0, 64], frequency 1/256), 256 pipe +
thermal steps with rain 0.01 and (Kc, Ks, Kd, Ke) = (0.05, 0.02, 0.02,
0.01), checkpoint every 16 steps; loss = mean((rock + dirt - target)^2), the
target made by the forward model from seed 12 with coefficients scaled by
U(0.8, 1.2) (the benchmark's construction).  All checks start from the
GPU's initial FP32 state.

Three gradients of the same loss are compared:
 * the GPU adjoint (FP32, hand-written reverse sweep);
 * the FP64 oracle in dual numbers: the exact derivative of the branch the
   rollout takes at every clamp, max/min, erosion case, talus test and
   smoothing test;
 * FP64 central differences at a step of 1e-3 (relative to each
   coefficient, absolute in height) with every branch frozen at the
   unperturbed run (orc::Shadow), whose limit is that same derivative.

The rollout itself is discontinuous: the erosion pass processes the rock
layer only when the dirt layer deposits or is exhausted
(glsl/hydro_erosion.glsl:58-83), the talus test switches a direction's
outflow from 0 to S*D_k/bk (glsl/thermal_erosion.glsl:63-91), and the
smoothing test switches the 5-point mean on (glsl/smoothing.glsl:62-73).  A
coefficient moves every cell, so plain central differences straddle some of
these switches at any step size; the frozen-branch differences do not, and
agree with the dual numbers to O(step^2).  Plain differences are run by the
opt-in test in test_gpu_long_adjoint.py."""
from __future__ import annotations

import time

import numpy as np
import pytest

import hydrogen as hg
import pyoracle as O
from helpers import C0

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

COEF = np.array([C0["Kc"], C0["Ks"], C0["Kd"], C0["Ke"]])
N, STEPS, K_CKPT = 1024, 256, 16


def say(capsys, msg):
    with capsys.disabled():
        print("\n" + msg, flush=True)


@pytest.fixture(scope="module")
def rollout():
    """The GPU adjoint of the full rollout and of its first 32 steps, with
    the start state and targets."""
    rng = np.random.default_rng(9)
    names = ("Kc", "Ks", "Kd", "Ke")
    pert = {nm: float(c) * float(rng.uniform(0.8, 1.2)) for nm, c in zip(names, COEF)}
    out = {}
    with hg.Terrain(N) as t:
        t.set_rain(hg.RAIN_CONSTANT, amount=0.01)
        for steps in (STEPS, 32):
            t.set_erosion(hg.erosion_settings(**pert))
            t.init_fbm(hg.fbm_params(seed=12, base_freq=1.0 / 256, hi=64.0))
            t.step = 0
            t.dispatch_grid(steps)
            target = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
            t.set_erosion(hg.erosion_settings(**C0))
            t.init_fbm(hg.fbm_params(seed=11, base_freq=1.0 / 256, hi=64.0))
            t.step = 0
            start = t.read_state()
            t.dispatch_grid(steps)
            h = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
            g = (2.0 * (h - target) / h.size).astype(np.float32)
            t.write_state(start[: hg.NUM_STATE_PLANES])
            t.step = 0
            gh, gc = t.adjoint_rollout(steps, K_CKPT, g)
            out[steps] = dict(start=start.astype(np.float64), target=target, gh=gh, gc=gc)
    return out


def coef_err(a, ref):
    """|a - ref| relative to each coefficient's gradient, with a floor of 1e-3
    of the largest (Kd's gradient is ~1e-5 of Kc's here)."""
    return np.abs(a - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())


CELLS = np.random.default_rng(3).choice(N * N, size=12, replace=False)


@pytest.fixture(scope="module")
def dual(rollout):
    """The FP64 dual-number gradient of the full rollout: (Kc, Ks, Kd, Ke)
    and the initial rock height at 12 cells, and the seconds it took."""
    r = rollout[STEPS]
    t0 = time.perf_counter()
    loss, dc, dh = O.rollout_grad_dual(r["start"], O.config_settings(**C0), COEF, STEPS, r["target"], CELLS,
                                       rain_mode=O.RAIN_CONSTANT, rain_amount=0.01)
    return dc, dh, time.perf_counter() - t0


def test_adjoint_vs_dual_numbers_1024_256_steps(rollout, dual, capsys):
    r = rollout[STEPS]
    dc, dh, secs = dual
    gcell = r["gh"].reshape(-1)[CELLS].astype(np.float64)
    ec, eh = coef_err(r["gc"], dc), np.abs(gcell - dh).max() / np.abs(dh).max()
    say(capsys, f"adjoint vs dual, {N}^2 x {STEPS} (dual numbers {secs:.0f} s): Kc {ec[0]:.2e} Ks {ec[1]:.2e} Kd {ec[2]:.2e} "
                f"Ke {ec[3]:.2e}; 12 cells {eh:.2e} of the largest")
    assert (ec[:3] <= 1e-3).all(), ec
    assert eh <= 1e-3, eh
    # dL/dKe sums water adjoint x water over every cell and step, and the
    # FP32 trajectory's water drifts about 1 % from the FP64 one over 256
    # steps (the FP32 oracle drifts as far): Ke is gated at 1e-3 on the
    # 32-step rollout below, here at 2e-2.
    assert ec[3] <= 2e-2, ec


def test_adjoint_vs_dual_numbers_1024_32_steps(rollout, capsys):
    r = rollout[32]
    t0 = time.perf_counter()
    loss, dc, _ = O.rollout_grad_dual(r["start"], O.config_settings(**C0), COEF, 32, r["target"],
                                      np.zeros(0, np.int32), rain_mode=O.RAIN_CONSTANT, rain_amount=0.01)
    ec = coef_err(r["gc"], dc)
    say(capsys, f"adjoint vs dual, {N}^2 x 32 ({time.perf_counter() - t0:.0f} s): "
                f"Kc {ec[0]:.2e} Ks {ec[1]:.2e} Kd {ec[2]:.2e} Ke {ec[3]:.2e}")
    assert (ec <= 1e-3).all(), ec


def test_frozen_branch_differences_vs_dual_numbers_1024_256_steps(rollout, dual, capsys):
    r = rollout[STEPS]
    dc, dh, _ = dual
    cells = CELLS[:4]
    t0 = time.perf_counter()
    fc, fh = O.rollout_fd_frozen(r["start"], O.config_settings(**C0), COEF, STEPS, r["target"], cells,
                                 eps_coef=1e-3, eps_cell=1e-3, rain_mode=O.RAIN_CONSTANT, rain_amount=0.01)
    ec, eh = coef_err(fc, dc), np.abs(fh - dh[:4]).max() / np.abs(dh[:4]).max()
    gcell = r["gh"].reshape(-1)[cells].astype(np.float64)
    ag = coef_err(r["gc"], fc)
    say(capsys, f"frozen-branch differences (step 1e-3) vs dual ({time.perf_counter() - t0:.0f} s): "
                f"Kc {ec[0]:.2e} Ks {ec[1]:.2e} Kd {ec[2]:.2e} Ke {ec[3]:.2e}; 4 cells {eh:.2e}; "
                f"GPU adjoint vs them: Kc {ag[0]:.2e} Ks {ag[1]:.2e} Kd {ag[2]:.2e} Ke {ag[3]:.2e}, "
                f"cells {np.abs(gcell - fh).max() / np.abs(fh).max():.2e}")
    assert (ec <= 1e-3).all(), ec
    assert eh <= 1e-3, eh
