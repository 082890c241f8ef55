"""GPU: the largest terrain the library accepts, 32768 x 32768 (2^30 cells,
4 GiB per plane, 160 GiB of device memory for the 20 planes and their
second buffers).  Plane offsets pass 2^32 bytes here, so every index must
be 64-bit.  Two steps of the fused kernels equal two steps of the per-pass
kernels bit for bit in all nine state planes (compared by 128-bit hashes,
one plane at a time, to keep host memory at one plane)."""
from __future__ import annotations

import gc

import numpy as np
import pytest
import xxhash

import hydrogen as hg
from helpers import C0

pytestmark = pytest.mark.gpu

N = 32768


def _device_memory():
    """(free, total) bytes of the current HIP device."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value, total.value


def _hashes(t, buf):
    out = []
    for p in range(hg.NUM_STATE_PLANES):
        t.read_plane(p, out=buf)
        out.append((xxhash.xxh3_128_hexdigest(memoryview(buf)), bool(np.isfinite(buf).all()),
                    float(buf[::4096, ::4096].astype(np.float64).sum())))
    return out


def test_max_size_fused_equals_per_pass():
    """One context: initialise, two fused steps, hash; initialise again, two
    per-pass steps, hash."""
    gc.collect()
    free, total = _device_memory()
    print(f"max size {N}^2: device memory free {free / 2**30:.1f} of {total / 2**30:.1f} GiB", flush=True)
    buf = np.empty((N, N), np.float32)
    res = []
    with hg.Terrain(N) as t:
        t.set_erosion(hg.erosion_settings(**C0))
        t.set_rain(hg.RAIN_UNIFORM, amount=0.02, seed=5)
        for name, flags in (("fused", hg.STEP_DEFAULT), ("per pass", hg.STEP_DEFAULT | hg.STEP_PER_PASS)):
            print(f"max size {N}^2: {name}", flush=True)
            t.init_fbm(hg.fbm_params(seed=5, base_freq=1.0 / 4096, hi=512.0))
            t.step = 0
            t.dispatch_grid(2, flags)
            t.sync()
            res.append(_hashes(t, buf))
    fused, plain = res
    for p, (a, b) in enumerate(zip(fused, plain)):
        print(f"plane {p}: {a[0]} finite {a[1]} sample sum {a[2]:.6e}", flush=True)
        assert a[1] and b[1], p
        assert a[0] == b[0], (p, a, b)
    assert fused[hg.P_WATER][2] > 0 and fused[hg.P_ROCK][2] > 0
