"""Shared helpers for the parity tests: seeded synthetic states and the
comparison metrics (the oracle is the checker; the GPU library is measured)."""
from __future__ import annotations

import numpy as np

import pyoracle as O

# Config-0 coefficients (512^2 plain pipe model).
C0 = dict(Kc=0.05, Ks=0.02, Kd=0.02, Ke=0.01, talus_deg=35.0, slippage=0.1, G=9.81, dt=0.005)


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """||a - b||_2 / ||b||_2 (b is the reference)."""
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def spun_up_state(n: int, steps: int, seed: int = 1, base_freq: float = 1 / 64, hi: float = 64.0,
                  rain_mode: int = O.RAIN_UNIFORM, rain_amount: float = 0.02) -> np.ndarray:
    """A non-trivial float32 state: fBm rock, then `steps` FP64 oracle steps
    (so flux, velocity, sediment and thermal planes are populated), rounded to
    float32."""
    p = O.fbm_params(seed=seed, base_freq=base_freq, hi=hi)
    st = O.new_state(n)
    st[O.ROCK] = O.fbm_init(n, p, np.float64)
    st[O.DIRT] = 0.5 + 0.25 * np.sin(np.arange(n)[None, :] / 7.0) * np.cos(np.arange(n)[:, None] / 5.0)
    if steps:
        O.run(st, O.config_settings(**C0), steps, O.STEP_DEFAULT, rain_mode, rain_amount, seed=seed)
    return st.astype(np.float32)


def close_fraction(got: np.ndarray, ref: np.ndarray, rtol: float, atol: float) -> float:
    """Fraction of cells outside |got-ref| <= atol + rtol*|ref|."""
    bad = np.abs(got.astype(np.float64) - ref) > atol + rtol * np.abs(ref)
    return float(bad.mean())
