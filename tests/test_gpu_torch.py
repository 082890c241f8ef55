"""GPU: the PyTorch autograd harness over the library's adjoint."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autograd_rollout_gradients_match_dual_oracle():
    import torch
    import hydrogen as hg
    from hydrogen.autograd import rollout
    import pyoracle as O

    n, steps = 32, 8
    h0 = O.fbm_init(n, O.fbm_params(seed=4, base_freq=1.0 / 12, hi=10.0), np.float32)
    tgt = O.fbm_init(n, O.fbm_params(seed=5, base_freq=1.0 / 12, hi=10.0), np.float32)
    coef0 = np.array([0.05, 0.02, 0.02, 0.01])
    with hg.Terrain(n) as t:
        t.set_erosion(hg.erosion_settings())
        t.set_rain(hg.RAIN_CONSTANT, amount=0.01)
        H0 = torch.tensor(h0, requires_grad=True)
        C = torch.tensor(coef0, requires_grad=True)
        H = rollout(H0, C, t, steps, k_ckpt=3)
        loss = ((H.double() - torch.tensor(tgt, dtype=torch.float64)) ** 2).mean()
        loss.backward()
    st = O.new_state(n)
    st[O.ROCK] = h0
    cells = [5 * n + 6, 17 * n + 3, 30 * n + 30]
    l64, dc, dh = O.rollout_grad_dual(st, O.config_settings(), coef0, steps, tgt, cells)
    assert abs(loss.item() - l64) <= 1e-4 * l64
    g = C.grad.numpy()
    for j in range(4):
        assert abs(g[j] - dc[j]) <= 2e-3 * max(abs(dc[j]), 1e-9), (j, g[j], dc[j])
    gh = H0.grad.numpy()
    assert np.abs(gh.flat[cells] - dh).max() <= 2e-3 * np.abs(gh).max()
