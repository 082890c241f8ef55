"""PyTorch autograd harness over the hand-written adjoint of the C ABI.

`rollout(h0, coef, terrain, nsteps)` runs a GRID rollout on the library's
HIP kernels from rock = h0 (dirt, water, flux and sediment zero) with
(Kc, Ks, Kd, Ke) = coef, and returns the final terrain height rock + dirt.
Its backward calls hg_adjoint_rollout, the library's checkpointed reverse
sweep.  Tensors cross to the library through host memory; torch's own device
allocator and streams are not shared with the library.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _abi


def _settings(base: _abi.ErosionSettings, coef: np.ndarray) -> _abi.ErosionSettings:
    s = _abi.ErosionSettings()
    ctypes.memmove(ctypes.byref(s), ctypes.byref(base), ctypes.sizeof(s))
    s.Kc = float(coef[0])
    s.Ks[0] = s.Ks[1] = float(coef[1])
    s.Kd[0] = s.Kd[1] = float(coef[2])
    s.Ke = float(coef[3])
    return s


def _load_initial(terrain: _abi.Terrain, h0: np.ndarray, step0: int):
    st = np.zeros((_abi.NUM_STATE_PLANES, terrain.n, terrain.n), dtype=np.float32)
    st[_abi.P_ROCK] = h0
    terrain.write_state(st)
    terrain.step = step0


class ErosionRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h0, coef, terrain, nsteps, k_ckpt, flags, base):
        h0_np = h0.detach().to("cpu", torch.float32).numpy()
        coef_np = coef.detach().to("cpu", torch.float64).numpy()
        terrain.set_erosion(_settings(base, coef_np))
        _load_initial(terrain, h0_np, 0)
        terrain.dispatch_grid(nsteps, flags)
        h = terrain.read_plane(_abi.P_ROCK) + terrain.read_plane(_abi.P_DIRT)
        ctx.terrain, ctx.nsteps, ctx.k_ckpt, ctx.flags, ctx.base = terrain, nsteps, k_ckpt, flags, base
        ctx.h0_np, ctx.coef_np = h0_np, coef_np
        ctx.h0_meta, ctx.coef_meta = (h0.device, h0.dtype), (coef.device, coef.dtype)
        return torch.from_numpy(h).to(h0.device, h0.dtype)

    @staticmethod
    def backward(ctx, grad_h):
        t = ctx.terrain
        t.set_erosion(_settings(ctx.base, ctx.coef_np))
        _load_initial(t, ctx.h0_np, 0)
        g = grad_h.detach().to("cpu", torch.float32).contiguous().numpy()
        g_h0, g_coef = t.adjoint_rollout(ctx.nsteps, ctx.k_ckpt, g, ctx.flags)
        gh = torch.from_numpy(g_h0).to(*ctx.h0_meta)
        gc = torch.from_numpy(g_coef).to(*ctx.coef_meta)
        return gh, gc, None, None, None, None, None


def rollout(h0: torch.Tensor, coef: torch.Tensor, terrain: _abi.Terrain, nsteps: int, k_ckpt: int = 16,
            flags: int = _abi.STEP_DEFAULT, base: _abi.ErosionSettings | None = None) -> torch.Tensor:
    """Differentiable GRID rollout; gradients flow to h0 and coef = (Kc, Ks, Kd, Ke)."""
    if base is None:
        base = terrain.get_erosion()
    return ErosionRollout.apply(h0, coef, terrain, nsteps, k_ckpt, flags, base)
