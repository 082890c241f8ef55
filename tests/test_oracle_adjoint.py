"""CPU: the dual-number FP64 oracle's derivatives against central finite
differences of the FP64 rollout loss (two independent differentiations of the
same model), on a small grid where no branch switches inside +-eps."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O


@pytest.fixture(scope="module")
def case():
    n, steps = 24, 10
    st = O.new_state(n)
    st[O.ROCK] = O.fbm_init(n, O.fbm_params(seed=2, base_freq=1 / 12, hi=12.0))
    tgt = O.fbm_init(n, O.fbm_params(seed=3, base_freq=1 / 12, hi=12.0))
    return n, steps, st, tgt, O.config_settings(), np.array([0.05, 0.02, 0.02, 0.01])


def test_dual_matches_central_differences(case):
    n, steps, st, tgt, s, coef = case
    cells = [3 * n + 4, 12 * n + 12, 20 * n + 5, n * n - 2]
    loss, gc, gh = O.rollout_grad_dual(st, s, coef, steps, tgt, cells)
    assert abs(loss - O.rollout_loss(st, s, coef, steps, tgt)) <= 1e-12 * abs(loss)
    for j in range(4):
        e = 1e-7
        cp, cm = coef.copy(), coef.copy()
        cp[j] += e
        cm[j] -= e
        fd = (O.rollout_loss(st, s, cp, steps, tgt) - O.rollout_loss(st, s, cm, steps, tgt)) / (2 * e)
        assert abs(fd - gc[j]) <= 1e-4 * max(abs(gc[j]), 1e-6), (j, fd, gc[j])
    for i, c in enumerate(cells):
        e = 1e-6
        sp, sm = st.copy(), st.copy()
        sp[O.ROCK].flat[c] += e
        sm[O.ROCK].flat[c] -= e
        fd = (O.rollout_loss(sp, s, coef, steps, tgt) - O.rollout_loss(sm, s, coef, steps, tgt)) / (2 * e)
        assert abs(fd - gh[i]) <= 1e-4 * max(abs(gh[i]), 1e-8), (c, fd, gh[i])


def test_dual_is_linear_in_the_seed(case):
    """Mixed check: d/dKc of the loss equals the derivative along the Kc axis
    of a rollout whose Kc is Kc0 + t (dual seeds are exact, not sampled)."""
    n, steps, st, tgt, s, coef = case
    _, gc, _ = O.rollout_grad_dual(st, s, coef, steps, tgt, [])
    # second-order central difference with two step sizes agrees with the dual
    f = lambda t: O.rollout_loss(st, s, coef + np.array([t, 0, 0, 0]), steps, tgt)
    d1 = (f(1e-6) - f(-1e-6)) / 2e-6
    d2 = (f(2e-6) - f(-2e-6)) / 4e-6
    assert abs(d1 - gc[0]) <= 1e-4 * abs(gc[0]) and abs(d2 - gc[0]) <= 1e-4 * abs(gc[0])


def test_frozen_branch_differences(case):
    """orc::Shadow: copies of the rollout with perturbed inputs that follow the
    unperturbed run's branches.  Their central difference is the derivative
    of the branch taken up to the step's truncation and FP64 cancellation, so
    it meets the dual numbers at 1e-3 with a step of 1e-3 (plain differences
    of a rollout with branch switches need not)."""
    n, steps, st, tgt, s, coef = case
    cells = [3 * n + 4, 12 * n + 12, 20 * n + 5, n * n - 2]
    _, gc, gh = O.rollout_grad_dual(st, s, coef, steps, tgt, cells)
    for e in (1e-2, 1e-3):
        fc, fh = O.rollout_fd_frozen(st, s, coef, steps, tgt, cells, eps_coef=e, eps_cell=e)
        floor = 1e-3 * np.abs(gc).max()
        assert (np.abs(fc - gc) <= 1e-3 * np.maximum(np.abs(gc), floor)).all(), (e, fc, gc)
        assert (np.abs(fh - gh) <= 1e-3 * np.abs(gh).max()).all(), (e, fh, gh)
    # a wrong step size in the quotient (or a copy that missed its
    # perturbation) would be off by O(1)
    fc, _ = O.rollout_fd_frozen(st, s, coef, steps, tgt, cells, eps_coef=1e-3, eps_cell=1e-3)
    assert abs(fc[0] / gc[0] - 1.0) < 1e-3 and abs(fc[1] / gc[1] - 1.0) < 1e-3
