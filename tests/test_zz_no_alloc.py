"""GPU: hg_droplets_noalloc, the droplet iterations that allocate nothing,
against hg_step_particle, the entry point they are a variant of.

The no-allocation call runs on planes of a second context (their device
addresses and that context's stream, as a caller's own buffers would be) and
writes its totals to device memory the caller owns (planes of a third
context); everything goes through the C ABI.  One droplet per
iteration has no concurrency, so both entry points give the same heights,
exported mass and droplet steps bit for bit; that is checked on two calls in
a row, the second on a smaller grid.  Many droplets update shared cells with
float atomics in an order the hardware picks, so a 2^16-droplet call is
judged as the droplet model is: mass balance with the call's own exported
total (a reduction that summed stale or missing partials would break it),
non-negative layers up to FP32 rounding, droplet steps, and the distance of
the heights to hg_step_particle's, against the distance between two runs of
hg_step_particle itself; then a smaller call must balance its own mass too
(the fixed array keeps the previous call's partials)."""
from __future__ import annotations

import numpy as np
import pytest

import hydrogen as hg

pytestmark = pytest.mark.gpu


def fresh(n, seed):
    t = hg.Terrain(n)
    t.init_fbm(hg.fbm_params(seed=seed, base_freq=4.0 / n, hi=128.0))
    return t


def existing(n, seed, d, iters):
    with fresh(n, seed) as t:
        t.set_droplets(d)
        t.dispatch_particle(iters)
        r, dd = t.read_plane(hg.P_ROCK), t.read_plane(hg.P_DIRT)
        ex, steps = t.particle_stats()
    return r, dd, ex, steps


def no_alloc(n, seed, d, iters):
    """The call on another context's planes, as on a caller's own device
    buffers, with its totals written to device memory the caller owns (two
    planes of a third, 8 x 8 context, pre-filled with a sentinel)."""
    with fresh(n, seed) as t, hg.Terrain(8) as out:
        sentinel = np.full((8, 8), -1.0, np.float32)
        out.write_plane(hg.P_VEL_U, sentinel)
        out.write_plane(hg.P_VEL_V, sentinel)
        h0 = t.read_plane(hg.P_ROCK).astype(np.float64).sum()
        hg.droplets_noalloc(t.device_ptr(hg.P_ROCK), t.device_ptr(hg.P_DIRT), n, d, 0, iters, t.stream(),
                            out.device_ptr(hg.P_VEL_U), out.device_ptr(hg.P_VEL_V))
        t.sync()
        r, dd = t.read_plane(hg.P_ROCK), t.read_plane(hg.P_DIRT)
        ex = float(out.read_plane(hg.P_VEL_U).view(np.float64).ravel()[0])
        steps = int(out.read_plane(hg.P_VEL_V).view(np.uint64).ravel()[0])
    return r, dd, ex, steps, h0


@pytest.mark.parametrize("n,iters", [(256, 3), (128, 2)])
def test_single_droplet_equals_step_particle(n, iters):
    d = hg.droplet_settings(count=1, seed=5)
    print(f"no-alloc droplets: n {n}, 1 droplet x {iters}", flush=True)
    r0, d0, ex0, st0 = existing(n, 3, d, iters)
    r1, d1, ex1, st1, _ = no_alloc(n, 3, d, iters)
    assert st1 == st0 and st1 > iters
    assert ex1 == ex0
    assert np.array_equal(r1, r0) and np.array_equal(d1, d0)


def test_many_droplets_two_calls_then_limits():
    for n, count, iters in ((512, 1 << 16, 2), (256, 4096, 1)):
        d = hg.droplet_settings(count=count, seed=9)
        print(f"no-alloc droplets: n {n}, {count} droplets x {iters}", flush=True)
        ra, da, exa, sta = existing(n, 4, d, iters)
        rb, db, _, _ = existing(n, 4, d, iters)
        r1, d1, ex1, st1, m0 = no_alloc(n, 4, d, iters)
        with fresh(n, 4) as t:
            h_init = t.read_plane(hg.P_ROCK).astype(np.float64)
        ha, hb = ra.astype(np.float64) + da, rb.astype(np.float64) + db
        h1 = r1.astype(np.float64) + d1
        # the call's own totals balance its own heights (a sum over stale or
        # missing partials would not)
        assert abs(h1.sum() + ex1 - m0) <= 1e-6 * m0
        assert abs(ha.sum() + exa - m0) <= 1e-6 * m0
        assert count * iters <= st1 <= count * iters * d.max_lifetime
        assert abs(st1 - sta) <= 0.01 * sta
        assert min(r1.min(), d1.min()) >= -1e-5       # FP32 rounding of raced takes only
        # the two entry points differ by no more than one differs from itself
        # (the order of the concurrent droplets' atomics)
        change = np.linalg.norm(ha - h_init)
        d_self = np.linalg.norm(ha - hb) / change
        d_cross = np.linalg.norm(h1 - ha) / change
        print(f"  heights: no-alloc vs existing {d_cross:.3e}, existing vs itself {d_self:.3e} of the change",
              flush=True)
        assert d_cross <= 3 * d_self + 0.05
    with fresh(64, 1) as t:
        big = hg.droplet_settings(count=hg.HG_DROPLETS_NOALLOC_MAX + 1)
        with pytest.raises(hg.HGError) as e:
            hg.droplets_noalloc(t.device_ptr(hg.P_ROCK), t.device_ptr(hg.P_DIRT), 64, big, 0, 1, t.stream())
        assert "HG_DROPLETS_NOALLOC_MAX" in str(e.value)
