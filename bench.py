"""Benchmark: the pipe+thermal GRID step on one MI355X.

Workload (the metric's headline configuration): one 4096 x 4096 FP32 terrain,
seeded simplex fBm (6 octaves, lacunarity 2, gain 0.5, base frequency 1/1024)
rescaled to heights [0, 512]; rain U(0, 0.02) per cell per step from
Philox4x32-10; coefficients g 9.81, dt 0.005, A = L = 1, Kc 0.05,
Ks = Kd = 0.02, Ke 0.01, talus 35 deg, slippage 0.1; thermal erosion and
smoothing every step.  W warm-up steps, then K timed steps (default: steps
100-2000 of the 2000-step run).

Prints one JSON line.  `value` is cell-steps/s over the timed steps with the
state resident in HBM.  `roofline` is the dominant kernel's algorithmic HBM
bytes per launch over its average launch time, measured with HIP events on
the library's stream inside the timed region; its `step_frac` is the whole
step's rate at 72 B per cell-step over the same peak.  `cpu_baseline` is the FP32
oracle (a CPU port of the same passes, OpenMP over rows) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hydro-gen_amd"))

import numpy as np  # noqa: E402

import hydrogen as hg  # noqa: E402

METRIC = "cell-steps/s at 4096² + % HBM roofline; batched and fwd+adjoint vs plain"
HBM_PEAK = 8.0e12            # MI355X HBM3E, bytes/s
# Algorithmic HBM bytes per cell for each fused kernel (each persistent plane
# read once and written once):
#   fused_hydro  : rock, dirt, water, flux L/R/T/B, sediment rock/dirt -> 9 x (4+4) = 72 B
#   fused_thermal: rock, dirt -> 2 x (4+4) = 16 B
#   step_sweep   : the whole step in one kernel, the 9 planes once -> 72 B
ALGO_BYTES = {"fused_hydro": 72, "fused_thermal": 16, "step_sweep": 72}
KERNEL_INDEX = {"fused_hydro": 0, "fused_thermal": 1, "step_sweep": 2}
STEP_ALGO_BYTES = 72         # minimum for a step that materialises the 9 planes once


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "fused_step_4096.md")


def profiled_traffic(kernel: str, n: int, variant: tuple) -> int | None:
    """HBM bytes per launch of `kernel` (tile variant (tx, ty, halo, threads)) on
    an n x n grid, from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary;
    None when that summary has no figure for this kernel, grid and variant."""
    key = f"traffic[{kernel} {n} {'x'.join(str(v) for v in variant)}]:"
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                line = line.strip().lstrip("- ")
                if line.startswith(key):
                    return int(line[len(key):].strip())
    except OSError:
        pass
    return None


def build_terrain(n: int, seed: int = 1) -> hg.Terrain:
    t = hg.Terrain(n, device=0)
    t.set_erosion(hg.erosion_settings(Kc=0.05, Ks=0.02, Kd=0.02, Ke=0.01, talus_deg=35.0,
                                      slippage=0.1, G=9.81, dt=0.005))
    t.set_rain(hg.RAIN_UNIFORM, amount=0.02, seed=seed)
    t.init_fbm(hg.fbm_params(seed=seed, octaves=6, lacunarity=2.0, gain=0.5,
                             base_freq=1.0 / 1024, lo=0.0, hi=512.0))
    return t


def batched_workload(seed: int = 2026, count: int = 64):
    """Batched config: `count` terrains, sides drawn uniformly from
    {256, 384, ..., 2048}, own fBm seed, coefficients Kc~U(0.02,0.1),
    Ks,Kd~U(0.01,0.05), Ke~U(0.005,0.02), talus~U(30,45) deg."""
    rng = np.random.default_rng(seed)
    sides = [int(x) for x in rng.choice(np.arange(256, 2049, 128), size=count)]
    coeffs = [dict(Kc=float(rng.uniform(0.02, 0.1)), Ks=float(rng.uniform(0.01, 0.05)),
                   Kd=float(rng.uniform(0.01, 0.05)), Ke=float(rng.uniform(0.005, 0.02)),
                   talus_deg=float(rng.uniform(30, 45)), slippage=0.1, G=9.81, dt=0.005)
              for _ in range(count)]
    return sides, coeffs


def bench_batched(steps: int, base_steps: int, graph_steps: int = 100) -> dict:
    sides, coeffs = batched_workload()
    cells = float(sum(n * n for n in sides))
    b = hg.Batch(sides)
    b.set_rain(hg.RAIN_UNIFORM, amount=0.02, seed=7)
    for i, n in enumerate(sides):
        b.set_erosion(hg.erosion_settings(**coeffs[i]), terrain=i)
        b.init_fbm(i, hg.fbm_params(seed=1000 + i, base_freq=1.0 / 256, hi=float(n) / 8))
    out = {"terrains": len(sides), "cells": int(cells), "sides_min": min(sides), "sides_max": max(sides)}
    # tile shapes for the batched kernels, timed on the batch itself (state restored)
    hv, tv = b.autotune(4)
    out["tile_variants"] = {"hydro": hv, "thermal": tv}
    log(f"batched tiles: hydro {hv}, thermal {tv}")
    # baselines: per-terrain launches, single-terrain fused kernels / one kernel
    # per pass, eager
    for key, flags, k in (("per_terrain_fused", hg.STEP_DEFAULT | hg.STEP_PER_TERRAIN, base_steps),
                          ("per_terrain_per_pass",
                           hg.STEP_DEFAULT | hg.STEP_PER_TERRAIN | hg.STEP_PER_PASS, base_steps)):
        b.step(2, flags)
        b.sync()
        t0 = time.perf_counter()
        b.step(k, flags)
        b.sync()
        dt = time.perf_counter() - t0
        out[key] = {"steps": k, "cell_steps_per_s": cells * k / dt, "ms_per_step": dt / k * 1e3}
    # baseline: per-terrain launches of the same fused kernels (each over its
    # terrain's tiles), `graph_steps` steps captured in one hipGraph, so launch
    # overhead is off the clock for the baseline as well
    if graph_steps > 0:
        b.set_persistent(False)
        b.capture(graph_steps, hg.STEP_DEFAULT | hg.STEP_PER_TERRAIN)
        b.run()
        b.sync()
        t0 = time.perf_counter()
        b.run()
        b.sync()
        dt = time.perf_counter() - t0
        out["per_terrain_fused_graph"] = {"steps": graph_steps, "cell_steps_per_s": cells * graph_steps / dt,
                                          "ms_per_step": dt / graph_steps * 1e3}
    # batched: one launch per pass over all terrains' tiles, `steps` steps
    # captured in one hipGraph; flat grid (a workgroup per tile) and persistent grid
    best = None
    for persistent in (False, True):
        b.set_persistent(persistent)
        b.capture(steps)
        b.run()          # warm-up replay
        b.sync()
        t0 = time.perf_counter()
        b.run()
        b.sync()
        dt = time.perf_counter() - t0
        v = cells * steps / dt
        key = "batched_graph_persistent" if persistent else "batched_graph"
        out[key] = {"steps": steps, "cell_steps_per_s": v, "ms_per_step": dt / steps * 1e3,
                    "hbm_roofline_frac": v * STEP_ALGO_BYTES / HBM_PEAK}
        best = v if best is None else max(best, v)
    v = best
    out["speedup_vs_per_terrain_per_pass"] = v / out["per_terrain_per_pass"]["cell_steps_per_s"]
    out["speedup_vs_per_terrain_fused"] = v / out["per_terrain_fused"]["cell_steps_per_s"]
    if "per_terrain_fused_graph" in out:
        out["speedup_vs_per_terrain_fused_graph"] = v / out["per_terrain_fused_graph"]["cell_steps_per_s"]
    h = b.read_plane(len(sides) - 1, hg.P_ROCK)
    out["state_finite"] = bool(np.isfinite(h).all())
    b.close()
    return out


def bench_reference_particles(iterations: int, n: int = 2048, count: int = 262144) -> dict:
    """The reference's own particle mode (hg_step_particle_reference) with its
    defaults (src/state.cpp:60-78, particle_count 262144 from the reference's
    config default): n^2 fBm [0, 256], every dispatch = movement, erosion,
    thermal erosion of both layers and smoothing with its particle branch."""
    t = hg.Terrain(n)
    t.init_fbm(hg.fbm_params(seed=3, base_freq=1.0 / 512, hi=256.0))
    t.set_erosion(hg.reference_particle_settings(count))
    t.ref_particles_init(count)
    t.dispatch_particle_reference(5, True, 0.0, 1.0 / 60)
    t.sync()
    t0 = time.perf_counter()
    t.dispatch_particle_reference(iterations, True, 5.0 / 60, 1.0 / 60)
    t.sync()
    dt = time.perf_counter() - t0
    h = t.read_plane(hg.P_ROCK)
    out = {"grid": n, "particles": count, "dispatches": iterations, "ms_per_dispatch": dt / iterations * 1e3,
           "particle_steps_per_s": count * iterations / dt, "cas_giveups": t.ref_particle_stats(),
           "state_finite": bool(np.isfinite(h).all())}
    t.close()
    return out


def bench_particles(iterations: int, oracle_iters: int) -> dict:
    """Particle config: 2048^2 fBm (frequency 1/512) in [0, 256], 2^20 droplets
    per iteration, 64 lifetime steps, inertia 0.05, capacity 4, deposition and
    erosion 0.3, evaporation 0.01, gravity 4, min slope 0.01, brush radius 3.
    Reports droplet-steps/s, the mass-balance residual, and, for the first
    `oracle_iters` iterations, the height-histogram L1 distance to the serial
    FP64 oracle next to that of the serial FP32 oracle (the rounding floor)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    n = 2048
    d = hg.droplet_settings(count=1 << 20, seed=42)
    t = hg.Terrain(n)
    t.init_fbm(hg.fbm_params(seed=3, base_freq=1.0 / 512, hi=256.0))
    h0 = t.read_plane(hg.P_ROCK)
    t.set_droplets(d)
    out = {"grid": n, "droplets_per_iteration": 1 << 20, "iterations": iterations}
    t.dispatch_particle(oracle_iters)
    h_early = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
    t.sync()
    t0 = time.perf_counter()
    t.dispatch_particle(iterations - oracle_iters)
    t.sync()
    dt = time.perf_counter() - t0
    ex, steps = t.particle_stats()
    h = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
    t.close()
    m0 = float(h0.astype(np.float64).sum())
    frac_timed = (iterations - oracle_iters) / iterations
    out["droplet_steps"] = int(steps)
    out["droplet_steps_per_s"] = steps * frac_timed / dt
    out["ms_per_iteration"] = dt / (iterations - oracle_iters) * 1e3
    out["mass_residual_rel"] = (float(h.sum()) + ex - m0) / m0
    out["exported_mass"] = ex
    # the same workload with the droplets sorted by start tile (an option:
    # faster, statistically further from the serial model)
    with hg.Terrain(n) as ts:
        ts.init_fbm(hg.fbm_params(seed=3, base_freq=1.0 / 512, hi=256.0))
        ts.set_droplets(d)
        ts.set_droplet_order(1)
        ts.dispatch_particle(1)
        ts.sync()
        k = max(1, min(5, iterations))
        t0 = time.perf_counter()
        ts.dispatch_particle(k)
        ts.sync()
        dts = time.perf_counter() - t0
        _, steps_s = ts.particle_stats()
        out["sorted_order"] = {"iterations": k, "ms_per_iteration": dts / k * 1e3,
                               "droplet_steps_per_s": steps_s * k / (k + 1) / dts}
    if oracle_iters > 0:
        od = O.droplet_settings(count=1 << 20, seed=42)
        r64, d64 = h0.astype(np.float64), np.zeros((n, n))
        r32, d32 = h0.copy(), np.zeros((n, n), np.float32)
        t0 = time.perf_counter()
        for it in range(oracle_iters):
            O.droplets(r64, d64, od, it)
        cpu_dt = time.perf_counter() - t0
        for it in range(oracle_iters):
            O.droplets(r32, d32, od, it)
        h64, h32 = r64 + d64, r32.astype(np.float64) + d32

        def l1(a, b, bins=256):
            lo, hi = min(a.min(), b.min()), max(a.max(), b.max())
            ha, _ = np.histogram(a, bins=bins, range=(lo, hi))
            hb, _ = np.histogram(b, bins=bins, range=(lo, hi))
            return float(np.abs(ha / ha.sum() - hb / hb.sum()).sum())
        out["oracle_check"] = {"iterations": oracle_iters, "hist_l1_gpu_vs_fp64": l1(h_early, h64),
                               "hist_l1_fp32_vs_fp64": l1(h32, h64), "hist_l1_init_vs_fp64": l1(h0, h64),
                               "fp64_oracle_seconds": cpu_dt, "cores": 1}
    return out


def bench_adjoint(n: int = 1024, steps: int = 256, k_ckpt: int = 16) -> dict:
    """Differentiable-rollout config: n^2 fBm terrain (6 octaves, frequency
    1/256, [0, 64]), `steps` pipe+thermal steps with the plain-config
    coefficients and rain 0.01/step; loss = MSE between the final rock+dirt and
    a target made by the forward model from another seed with coefficients
    scaled by U(0.8, 1.2); gradients w.r.t. the initial height and
    (Kc, Ks, Kd, Ke) through the adjoint with a checkpoint every k_ckpt steps.
    Reports t(forward + adjoint) / t(forward), each the median of 3 warm runs
    (forward and adjoint alternating)."""
    coef = dict(Kc=0.05, Ks=0.02, Kd=0.02, Ke=0.01)
    rng = np.random.default_rng(9)
    pert = {k: v * float(rng.uniform(0.8, 1.2)) for k, v in coef.items()}
    t = hg.Terrain(n)
    t.set_rain(hg.RAIN_CONSTANT, amount=0.01)
    # the forward and the adjoint's recompute use the tile variants tuned for
    # this grid (the state is re-initialised below)
    t.set_erosion(hg.erosion_settings(**coef))
    t.init_fbm(hg.fbm_params(seed=11, base_freq=1.0 / 256, hi=64.0))
    tiles = t.autotune(40)        # 40 steps per variant: 10 (0.4 ms) at 1024^2 picked unstably
    # target terrain
    t.set_erosion(hg.erosion_settings(**pert))
    t.init_fbm(hg.fbm_params(seed=12, base_freq=1.0 / 256, hi=64.0))
    t.dispatch_grid(steps)
    target = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
    t.set_erosion(hg.erosion_settings(**coef))
    t.init_fbm(hg.fbm_params(seed=11, base_freq=1.0 / 256, hi=64.0))
    init = t.read_state(hg.NUM_STATE_PLANES)
    reps = 3

    def time_forward() -> float:
        t.write_state(init)
        t.step = 0
        t.sync()
        t0 = time.perf_counter()
        t.dispatch_grid(steps)
        t.sync()
        return time.perf_counter() - t0

    def time_adjoint(k: int):
        t.write_state(init)
        t.step = 0
        t.sync()
        t0 = time.perf_counter()
        out = t.adjoint_rollout(steps, k, g)
        return time.perf_counter() - t0, out

    # forward alone (one untimed run first), then the loss gradient at its end
    t.dispatch_grid(steps)
    t.sync()
    time_forward()
    h = t.read_plane(hg.P_ROCK).astype(np.float64) + t.read_plane(hg.P_DIRT)
    loss = float(((h - target) ** 2).mean())
    g = (2.0 * (h - target) / h.size).astype(np.float32)
    # forward + adjoint (the adjoint call runs its own checkpointed forward),
    # one untimed call first; then `reps` forwards and adjoints, alternating,
    # and the median of each
    time_adjoint(k_ckpt)
    fwd_s, adj_s = [], []
    for _ in range(reps):
        fwd_s.append(time_forward())
        dt, (gh, gc) = time_adjoint(k_ckpt)
        adj_s.append(dt)
    t_fwd, t_adj = float(np.median(fwd_s)), float(np.median(adj_s))
    # the same rollout taped whole by the forward (one segment: no
    # recompute; 11 + 9 planes per step stay in HBM, 20 GB at 1024^2 x 256)
    time_adjoint(steps)
    full_s = []
    for _ in range(reps):
        dt, (gh_full, gc_full) = time_adjoint(steps)
        full_s.append(dt)
    t_full = float(np.median(full_s))
    # the checkpointed rollout again with one launch per pass adjoint (no
    # fused reverse launches)
    t.set_adjoint_fusion(0)
    time_adjoint(k_ckpt)
    sep_s = []
    for _ in range(reps):
        dt, (gh_sep, gc_sep) = time_adjoint(k_ckpt)
        sep_s.append(dt)
    t_sep = float(np.median(sep_s))
    t.set_adjoint_fusion(15)
    t.close()
    return {"grid": n, "steps": steps, "k_ckpt": k_ckpt, "tile_variants": list(tiles), "loss": loss,
            "t_forward_ms": t_fwd * 1e3, "t_forward_adjoint_ms": t_adj * 1e3,
            "cost_ratio": t_adj / t_fwd, "timing": f"median of {reps} runs each, forward and adjoint alternating",
            "runs_ms": {"forward": [round(x * 1e3, 3) for x in fwd_s],
                        "forward_adjoint": [round(x * 1e3, 3) for x in adj_s]},
            "grad_coef": [float(x) for x in gc],
            "grad_h0_absmax": float(np.abs(gh).max()), "grad_finite": bool(np.isfinite(gh).all()),
            "full_tape": {"k_ckpt": steps, "t_forward_adjoint_ms": t_full * 1e3, "cost_ratio": t_full / t_fwd,
                          "same_gradient": bool(np.array_equal(gh_full, gh)
                                                and np.array_equal(np.asarray(gc_full), np.asarray(gc)))},
            "separate_kernels": {"t_forward_adjoint_ms": t_sep * 1e3, "cost_ratio": t_sep / t_fwd,
                                      "same_gradient": bool(np.array_equal(gh_sep, gh)
                                                            and np.array_equal(np.asarray(gc_sep), np.asarray(gc)))}}


def calibrate_copy(nbytes: int = 256 << 20, reps: int = 3) -> dict:
    """A plain device-to-device copy of known size (for scaling the PMC
    FETCH_SIZE / WRITE_SIZE counters, which a profiler attributes to the copy
    kernel of the HIP runtime)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    dt = (time.perf_counter() - t0) / reps
    hip.hipFree(src)
    hip.hipFree(dst)
    return {"bytes": nbytes, "copy_GBps": 2 * nbytes / dt / 1e9}


def cpu_baseline(t: hg.Terrain, steps: int) -> dict:
    """FP32 oracle (CPU port of the same passes) on `steps` steps of the same
    4096^2 state; OpenMP threads from OMP_NUM_THREADS, bound close.  Also the
    512^2 plain configuration (fBm [0, 64], rain 0.01) for 100 steps with the
    FP32 port and with the FP64 oracle, timed apart."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    state = np.zeros((O.NUM_PLANES, t.n, t.n), dtype=np.float32)
    state[: hg.NUM_STATE_PLANES] = t.read_state(hg.NUM_STATE_PLANES)
    s = O.config_settings(Kc=0.05, Ks=0.02, Kd=0.02, Ke=0.01, talus_deg=35.0, slippage=0.1,
                          G=9.81, dt=0.005)
    O.run(state, s, 1, O.STEP_DEFAULT, O.RAIN_UNIFORM, 0.02, seed=1, step0=t.step)  # warm caches
    t0 = time.perf_counter()
    O.run(state, s, steps, O.STEP_DEFAULT, O.RAIN_UNIFORM, 0.02, seed=1, step0=t.step + 1)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # the 512^2 configuration, FP32 port and FP64 oracle
    c0 = {}
    n0, k0 = 512, 100
    init = O.fbm_init(n0, O.fbm_params(seed=1, base_freq=1.0 / 256, hi=64.0), np.float64)
    for name, dtype in (("fp32_port", np.float32), ("fp64_oracle", np.float64)):
        st = O.new_state(n0).astype(dtype)
        st[O.ROCK] = init
        t0 = time.perf_counter()
        O.run(st, s, k0, O.STEP_DEFAULT, O.RAIN_CONSTANT, 0.01, seed=1)
        c0[name] = n0 * n0 * k0 / (time.perf_counter() - t0)
    return {"value": t.n * t.n * steps / dt, "unit": "cell-steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} FP32 steps of the same {t.n}x{t.n} workload (state after the "
                      f"warm-up), oracle/liboracle.so, {cores} OpenMP threads, {dt:.1f} s",
            "plain_512": {"steps": k0, "unit": "cell-steps/s", **c0}}


def log(msg: str) -> None:
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1900)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--per-pass", action="store_true", help="time the unfused per-pass kernels")
    ap.add_argument("--cpu-steps", type=int, default=96, help="CPU baseline sample (0 = skip)")
    ap.add_argument("--batched-steps", type=int, default=1000,
                    help="steps of the 64-terrain batched graph (0 = skip)")
    ap.add_argument("--per-terrain-graph-steps", type=int, default=100,
                    help="steps of the per-terrain baseline graph (128 launches per step; 0 = skip)")
    ap.add_argument("--particle-iters", type=int, default=100,
                    help="droplet iterations of the particle workload (0 = skip)")
    ap.add_argument("--particle-oracle-iters", type=int, default=1)
    ap.add_argument("--adjoint-steps", type=int, default=256,
                    help="steps of the differentiable-rollout workload (0 = skip)")
    ap.add_argument("--adjoint-ckpt", type=int, default=16,
                    help="checkpoint interval of the differentiable-rollout workload")
    ap.add_argument("--calibrate-copy", action="store_true", default=True,
                    help="run a 256 MiB device-to-device copy: the box's copy bandwidth next to the "
                         "8 TB/s peak, and the PMC byte counters' calibration (default)")
    ap.add_argument("--no-calibrate-copy", dest="calibrate_copy", action="store_false")
    ap.add_argument("--sweep", type=int, default=None,
                    help="force the step kernel: -1 two kernels, v >= 0 one-kernel variant v "
                         "(default: the autotuner's choice)")
    ap.add_argument("--tune", action="store_true", default=True)
    ap.add_argument("--no-tune", dest="tune", action="store_false")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: the erosion path is measured per GPU; "
              "the multi-GPU sharding is designed but not implemented", file=sys.stderr)
        return 2

    n = args.n
    flags = hg.STEP_DEFAULT | (hg.STEP_PER_PASS if args.per_pass else 0)
    t = build_terrain(n)
    tiles = None
    sweep = -1
    if args.tune and not args.per_pass:
        tiles = t.autotune(10, flags)
        sweep = t.get_sweep()
    if args.sweep is not None and not args.per_pass:
        t.set_sweep(args.sweep)
        sweep = args.sweep
    t.dispatch_grid(args.warmup, flags)
    t.sync()

    t.set_timing(True)
    t.reset_timing()
    t.sync()
    t0 = time.perf_counter()
    t.dispatch_grid(args.steps, flags)
    t.sync()
    wall = time.perf_counter() - t0
    stats = t.kernel_stats()
    t.set_timing(False)

    cell_steps = float(n) * n * args.steps
    value = cell_steps / wall
    kern = {k: {"launches": c, "avg_us": ms / c * 1e3} for k, (c, ms) in stats.items()}
    dom = max(stats, key=lambda k: stats[k][1])
    dom_avg_s = stats[dom][1] / stats[dom][0] / 1e3
    per_cell = ALGO_BYTES.get(dom)
    roofline = None
    if per_cell is not None:
        ach = per_cell * n * n / dom_avg_s
        hv, tv = t.get_tiles()
        pick = {"fused_hydro": hv, "fused_thermal": tv, "step_sweep": sweep}[dom]
        variant = hg.tile_variants(KERNEL_INDEX[dom])[pick]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9,
                    "unit": "GB/s", "frac": ach / HBM_PEAK,
                    "traffic": profiled_traffic(dom, n, variant),
                    "algo_bytes_per_launch": per_cell * n * n,
                    # the whole step (every kernel of it) at 72 B per cell-step
                    "step_frac": value * STEP_ALGO_BYTES / HBM_PEAK}

    log(f"{args.steps} GRID steps at {n}^2: {wall / args.steps * 1e3:.3f} ms/step")
    h = t.read_plane(hg.P_ROCK)
    finite = bool(np.isfinite(h).all())
    cpu = cpu_baseline(t, args.cpu_steps) if args.cpu_steps > 0 else None
    if cpu:
        log(f"CPU baseline: {cpu['value']:.3e} cell-steps/s")
    t.close()
    batched = (bench_batched(args.batched_steps, 10, args.per_terrain_graph_steps)
               if args.batched_steps > 0 else None)
    if batched:
        log("batched terrains done")
    particles = (bench_particles(args.particle_iters, args.particle_oracle_iters)
                 if args.particle_iters > 0 else None)
    if particles is not None:
        particles["reference_mode"] = bench_reference_particles(args.particle_iters)
        log("reference particle mode done")
    if particles:
        log("droplets done")
    adjoint = (bench_adjoint(steps=args.adjoint_steps, k_ckpt=args.adjoint_ckpt)
               if args.adjoint_steps > 0 else None)
    if adjoint:
        log("adjoint done")
    copy = calibrate_copy() if args.calibrate_copy else None

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"pipe+thermal GRID step, {n}x{n} FP32, simplex fBm [0,512], "
                        "rain U(0,0.02) Philox per step, thermal+smoothing every step",
            "grid": n,
            "path": "per-pass" if args.per_pass else "fused",
            "step_hbm_roofline_frac": value * STEP_ALGO_BYTES / HBM_PEAK,
            "kernels": kern,
            "tile_variants": {"hydro": tiles[0], "thermal": tiles[1]} if tiles else None,
            "sweep_variant": sweep,      # -1: hydro + thermal kernels; v >= 0: one-kernel variant v
            "state_finite": finite,
            "batched": batched,
            "particles": particles,
            "adjoint": adjoint,
            "copy_calibration": copy,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
