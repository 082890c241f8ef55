"""Writes the regression fixtures of tests/golden/ from the CPU oracle.

    python tests/golden/make_fixtures.py

grid16x12_c0.npz: a 16 x 12 terrain (simplex fBm rock, sinusoidal dirt) run
for 30 GRID steps with the config-0 coefficients and uniform Philox rain
0.02 (seed 5) by the FP64 oracle; `initial` holds all 20 planes, `final`
the 9 persistent planes after the run.  The reference holds no golden
vectors of its own; philox4x32_10_kat.json holds the published Philox
known-answer vectors.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

import pyoracle as O  # noqa: E402

C0 = dict(Kc=0.05, Ks=0.02, Kd=0.02, Ke=0.01, talus_deg=35.0, slippage=0.1, G=9.81, dt=0.005)


def main():
    nx, ny, steps, seed = 16, 12, 30, 5
    st = O.new_state(nx, ny=ny)
    st[O.ROCK] = O.fbm_init(nx, O.fbm_params(seed=seed, base_freq=1 / 8, hi=6.0), np.float64, ny=ny)
    st[O.DIRT] = 0.3 + 0.2 * np.sin(np.arange(nx)[None, :] / 3.0) * np.cos(np.arange(ny)[:, None] / 2.0)
    initial = st.copy()
    O.run(st, O.config_settings(**C0), steps, O.STEP_DEFAULT, O.RAIN_UNIFORM, 0.02, seed=seed)
    np.savez(os.path.join(HERE, "grid16x12_c0.npz"), initial=initial, final=st[:9], steps=steps, seed=seed)


if __name__ == "__main__":
    main()
