"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product library never imports it.

State is a float64 (or float32) numpy array of shape (NUM_PLANES, ny, nx) in
the HG_P_* plane order of include/hg.h.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NUM_PLANES = 20
ROCK, DIRT, WATER, FL, FR, FT, FB, S0, S1, VX, VY, VZ = range(12)
THERM = list(range(12, 20))
PASS_RAIN, PASS_FLUX, PASS_EROSION, PASS_TRANSPORT, PASS_TFLUX0, PASS_TTRANS0, PASS_TFLUX1, \
    PASS_TTRANS1, PASS_SMOOTH = range(9)
STEP_RAIN, STEP_THERMAL, STEP_SMOOTH = 1, 2, 4
STEP_DEFAULT = 7
REFP_REVERSE_ORDER = 0x100   # reference particle mode: erode in descending id order
RAIN_OFF, RAIN_CONSTANT, RAIN_UNIFORM = 0, 1, 2


class ErosionSettings(ctypes.Structure):
    """Erosion_data (glsl/bindings.glsl:39-60), std140 layout, 96 bytes."""
    _fields_ = [
        ("particle_count", ctypes.c_uint32), ("Kc", ctypes.c_float),
        ("Kalpha", ctypes.c_float * 2), ("Kconv", ctypes.c_float), ("_pad0", ctypes.c_float),
        ("Ks", ctypes.c_float * 2), ("Kd", ctypes.c_float * 2), ("Ke", ctypes.c_float),
        ("energy_kept", ctypes.c_float), ("Kspeed", ctypes.c_float * 2), ("G", ctypes.c_float),
        ("d_t", ctypes.c_float), ("density", ctypes.c_float), ("init_volume", ctypes.c_float),
        ("friction", ctypes.c_float), ("inertia", ctypes.c_float), ("min_volume", ctypes.c_float),
        ("min_velocity", ctypes.c_float), ("ttl", ctypes.c_uint32), ("_pad1", ctypes.c_uint32),
    ]


class RainSettings(ctypes.Structure):
    """Rain_data (glsl/bindings.glsl:62-68), std140, 32 bytes."""
    _fields_ = [("amount", ctypes.c_float), ("mountain_thresh", ctypes.c_float),
                ("mountain_multip", ctypes.c_float), ("period", ctypes.c_int32), ("drops", ctypes.c_float),
                ("_pad", ctypes.c_uint32 * 3)]


def rain_settings(amount=0.01, mountain_thresh=0.55, mountain_multip=0.05, period=512, drops=0.02):
    """Defaults: src/settings.hpp:15-21."""
    r = RainSettings()
    r.amount = amount; r.mountain_thresh = mountain_thresh; r.mountain_multip = mountain_multip
    r.period = period; r.drops = drops
    return r


class MapSettings(ctypes.Structure):
    """Map_settings_data (glsl/bindings.glsl:70-97), std140, 96 bytes."""
    _fields_ = [("max_height", ctypes.c_float), ("max_dirt", ctypes.c_float), ("hmap_dims", ctypes.c_int32 * 2),
                ("height_mult", ctypes.c_float), ("water_lvl", ctypes.c_float), ("seed", ctypes.c_float),
                ("persistance", ctypes.c_float), ("lacunarity", ctypes.c_float), ("scale", ctypes.c_float),
                ("redistribution", ctypes.c_float), ("octaves", ctypes.c_int32), ("fake_erosion", ctypes.c_uint32),
                ("mask_round", ctypes.c_uint32), ("mask_exp", ctypes.c_uint32), ("mask_power", ctypes.c_uint32),
                ("mask_slope", ctypes.c_uint32), ("uplift", ctypes.c_uint32), ("uplift_scale", ctypes.c_float),
                ("domain_warp", ctypes.c_int32), ("domain_warp_scale", ctypes.c_float), ("terrace", ctypes.c_int32),
                ("terrace_scale", ctypes.c_float), ("_pad", ctypes.c_uint32)]


def map_settings(**kw) -> MapSettings:
    """The reference defaults (src/settings.hpp:29-55) with a fixed seed."""
    d = dict(max_height=256.0, max_dirt=2.0, height_mult=1.0, water_lvl=96.0, seed=1234.5, persistance=0.44,
             lacunarity=2.0, scale=0.00075, redistribution=1.0, octaves=8, fake_erosion=0, mask_round=0,
             mask_exp=1, mask_power=1, mask_slope=0, uplift=0, uplift_scale=1.16, domain_warp=1,
             domain_warp_scale=100.0, terrace=0, terrace_scale=0.5)
    d.update(kw)
    m = MapSettings()
    m.hmap_dims[0] = m.hmap_dims[1] = 1024
    for k, v in d.items():
        setattr(m, k, v)
    return m


class FbmParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("octaves", ctypes.c_int32),
                ("lacunarity", ctypes.c_float), ("gain", ctypes.c_float),
                ("base_freq", ctypes.c_float), ("height_lo", ctypes.c_float),
                ("height_hi", ctypes.c_float)]


class DropletSettings(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint32), ("max_lifetime", ctypes.c_uint32),
                ("inertia", ctypes.c_float), ("capacity", ctypes.c_float),
                ("deposition", ctypes.c_float), ("erosion", ctypes.c_float),
                ("evaporation", ctypes.c_float), ("gravity", ctypes.c_float),
                ("min_slope", ctypes.c_float), ("brush_radius", ctypes.c_int32),
                ("init_water", ctypes.c_float), ("init_speed", ctypes.c_float),
                ("seed", ctypes.c_uint64)]


def droplet_settings(count=1 << 20, max_lifetime=64, inertia=0.05, capacity=4.0, deposition=0.3,
                     erosion=0.3, evaporation=0.01, gravity=4.0, min_slope=0.01, brush_radius=3,
                     init_water=1.0, init_speed=1.0, seed=0) -> DropletSettings:
    d = DropletSettings()
    d.count = count; d.max_lifetime = max_lifetime; d.inertia = inertia; d.capacity = capacity
    d.deposition = deposition; d.erosion = erosion; d.evaporation = evaporation
    d.gravity = gravity; d.min_slope = min_slope; d.brush_radius = brush_radius
    d.init_water = init_water; d.init_speed = init_speed; d.seed = seed
    return d


def droplets(rock: np.ndarray, dirt: np.ndarray, s: DropletSettings, iteration: int) -> float:
    """One serial droplet iteration in place (FP64 or FP32 by dtype); returns
    the exported mass."""
    assert rock.dtype == dirt.dtype
    n = rock.shape[0]
    fn = lib().orc_droplets_f64 if rock.dtype == np.float64 else lib().orc_droplets_f32
    return fn(_ptr(rock), _ptr(dirt), n, ctypes.byref(s), iteration)


class RefParticle(ctypes.Structure):
    """Particle (glsl/bindings.glsl:101-111), std430, 48 bytes."""
    _fields_ = [("sc", ctypes.c_float), ("iters", ctypes.c_int32), ("position", ctypes.c_float * 2),
                ("velocity", ctypes.c_float * 2), ("volume", ctypes.c_float), ("_pad0", ctypes.c_float),
                ("sediment", ctypes.c_float * 2), ("to_kill", ctypes.c_uint32), ("_pad1", ctypes.c_uint32)]


def reference_particle_defaults(particle_count: int) -> ErosionSettings:
    """src/state.cpp:60-78 (particle mode)."""
    s = ErosionSettings()
    s.particle_count = particle_count
    s.Kc = 0.2
    s.Kalpha[0] = 1.3; s.Kalpha[1] = 0.6
    s.Kconv = 0.001
    s.Ks[0] = 0.03; s.Ks[1] = 0.09
    s.Kd[0] = 0.01; s.Kd[1] = 0.03
    s.Ke = 0.03
    s.Kspeed[0] = 0.002; s.Kspeed[1] = 0.008
    s.G = 9.81
    s.d_t = 0.25
    s.density = 1.0
    s.init_volume = 1.0
    s.friction = 0.2
    s.inertia = 1.0
    s.min_volume = 0.0
    s.min_velocity = 0.001
    s.ttl = 15000
    return s


def ref_particle_run(planes: np.ndarray, parts, s: ErosionSettings, iterations: int, should_rain: bool,
                     time0: float, time_step: float, flags: int, dims=(0, 0)) -> None:
    """`iterations` reference particle dispatches in place: planes (NUM_PLANES,
    n, n) float64 or float32 (FLUX planes = momentum map x, y, z, w), parts a
    ctypes array of RefParticle (rounded to FP32 after the run).  `dims` is
    Map_settings_data.hmap_dims as the shader reads it for the spawn range and
    border test ((0, 0) = the grid's n).  flags: STEP_THERMAL, STEP_SMOOTH, and
    REFP_REVERSE_ORDER to erode in descending particle id order."""
    n = planes.shape[-1]
    fn = lib().orc_ref_particle_run_f64 if planes.dtype == np.float64 else lib().orc_ref_particle_run_f32
    rc = fn(_ptr(planes), n, parts, len(parts), ctypes.byref(s), iterations, int(should_rain), time0, time_step,
            flags, int(dims[0]), int(dims[1]))
    assert rc == 0


def ref_particle_rand(x: float, y: float, dtype=np.float32) -> float:
    """rand(vec2 p) of glsl/particle.glsl:41-44."""
    return (lib().orc_ref_particle_rand_f32 if dtype == np.float32 else lib().orc_ref_particle_rand_f64)(x, y)


def config_settings(Kc=0.05, Ks=0.02, Kd=0.02, Ke=0.01, talus_deg=35.0, slippage=0.1,
                    G=9.81, dt=0.005, Kconv=0.001, E=1.0) -> ErosionSettings:
    """Scalar coefficients (one per process) broadcast to both layers; the unnamed
    fields take the reference grid defaults (src/state.cpp:81-92)."""
    s = ErosionSettings()
    s.particle_count = 0
    s.Kc = Kc
    a = math.radians(talus_deg)
    s.Kalpha[0] = a; s.Kalpha[1] = a
    s.Kconv = Kconv
    s.Ks[0] = Ks; s.Ks[1] = Ks
    s.Kd[0] = Kd; s.Kd[1] = Kd
    s.Ke = Ke
    s.energy_kept = E
    s.Kspeed[0] = slippage; s.Kspeed[1] = slippage
    s.G = G
    s.d_t = dt
    return s


def reference_grid_defaults() -> ErosionSettings:
    """src/state.cpp:81-92."""
    s = ErosionSettings()
    s.Kc = 0.2
    s.Kalpha[0] = 1.3; s.Kalpha[1] = 0.6
    s.Kconv = 0.001
    s.Ks[0] = 0.03; s.Ks[1] = 0.09
    s.Kd[0] = 0.01; s.Kd[1] = 0.03
    s.Ke = 0.03
    s.energy_kept = 1.0
    s.Kspeed[0] = 0.5; s.Kspeed[1] = 2.0
    s.G = 1.0
    s.d_t = 0.001
    return s


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_rain_u.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int,
                                 ctypes.c_int]
        L.orc_rain_u.restype = ctypes.c_double
        L.orc_fbm_permutation.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint8)]
        L.orc_simplex2_f64.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
        L.orc_simplex2_f64.restype = ctypes.c_double
        L.orc_droplets_f64.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                       ctypes.c_int, ctypes.POINTER(DropletSettings), ctypes.c_uint32]
        L.orc_droplets_f64.restype = ctypes.c_double
        L.orc_droplets_f32.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                       ctypes.c_int, ctypes.POINTER(DropletSettings), ctypes.c_uint32]
        L.orc_droplets_f32.restype = ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        for name, t in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            getattr(L, f"orc_ref_particle_run_{name}").argtypes = [
                ctypes.POINTER(t), ctypes.c_int, ctypes.POINTER(RefParticle), ctypes.c_uint32,
                ctypes.POINTER(ErosionSettings), ctypes.c_uint32, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                ctypes.c_uint, ctypes.c_int, ctypes.c_int]
        for name, ct in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            getattr(L, f"orc_smooth_flags_{name}").argtypes = [ctypes.POINTER(ct), ctypes.c_int, ctypes.c_int,
                                                               ctypes.POINTER(ctypes.c_uint8)]
        L.orc_gen_exp_f32.argtypes = [ctypes.c_float]
        L.orc_gen_exp_f32.restype = ctypes.c_float
        L.orc_hash_sin_f32.argtypes = [ctypes.c_float]
        L.orc_hash_sin_f32.restype = ctypes.c_float
        L.orc_ref_particle_rand_f32.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_ref_particle_rand_f32.restype = ctypes.c_float
        L.orc_ref_particle_rand_f64.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_ref_particle_rand_f64.restype = ctypes.c_double
        L.orc_rollout_loss_f64.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ErosionSettings), dp,
                                           ctypes.c_int, ctypes.c_float, ctypes.c_uint64, ctypes.c_int,
                                           ctypes.c_uint, dp, dp]
        L.orc_rollout_grad_dual.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ErosionSettings), dp,
                                            ctypes.c_int, ctypes.c_float, ctypes.c_uint64, ctypes.c_int,
                                            ctypes.c_uint, dp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, dp, dp]
        L.orc_rollout_fd_frozen.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ErosionSettings), dp,
                                            ctypes.c_int, ctypes.c_float, ctypes.c_uint64, ctypes.c_int,
                                            ctypes.c_uint, dp, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double, dp]
        for name, t in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            getattr(L, f"orc_gen_heightmap_{name}").argtypes = [ctypes.POINTER(t), ctypes.POINTER(t), ctypes.c_int,
                                                                 ctypes.POINTER(MapSettings)]
        L.orc_elias_hash.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.orc_noised_f64.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.orc_ref_simplex_f64.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_ref_simplex_f64.restype = ctypes.c_double
        L.orc_ref_simplex_f32.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_ref_simplex_f32.restype = ctypes.c_float
        L.orc_ref_sfbm_f64.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int]
        L.orc_ref_sfbm_f64.restype = ctypes.c_double
        for name, t in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            getattr(L, f"orc_rain_reference_{name}").argtypes = [
                ctypes.POINTER(t), ctypes.c_int, ctypes.c_int, ctypes.POINTER(RainSettings), ctypes.c_float,
                ctypes.c_float]
        for name, t in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            p = ctypes.POINTER(t)
            getattr(L, f"orc_fbm_init_{name}").argtypes = [p, ctypes.c_int, ctypes.c_int,
                                                           ctypes.POINTER(FbmParams)]
            getattr(L, f"orc_run_{name}").argtypes = [
                p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ErosionSettings), ctypes.c_int,
                ctypes.c_float, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                ctypes.c_uint]
            getattr(L, f"orc_pass_{name}").argtypes = [
                p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ErosionSettings), ctypes.c_int,
                ctypes.c_float, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    t = ctypes.c_double if a.dtype == np.float64 else ctypes.c_float
    return a.ctypes.data_as(ctypes.POINTER(t))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def rain_u(seed, step, terrain, x, y) -> float:
    return lib().orc_rain_u(seed, step, terrain, x, y)


def fbm_permutation(seed: int) -> np.ndarray:
    p = (ctypes.c_uint8 * 256)()
    lib().orc_fbm_permutation(seed, p)
    return np.frombuffer(bytes(p), dtype=np.uint8).copy()


def simplex2(seed: int, x: float, y: float) -> float:
    return lib().orc_simplex2_f64(seed, x, y)


def fbm_params(seed=1, octaves=6, lacunarity=2.0, gain=0.5, base_freq=1.0 / 256, lo=0.0, hi=64.0):
    p = FbmParams()
    p.seed = seed; p.octaves = octaves; p.lacunarity = lacunarity; p.gain = gain
    p.base_freq = base_freq; p.height_lo = lo; p.height_hi = hi
    return p


def fbm_init(n: int, params: FbmParams, dtype=np.float64, ny: int | None = None) -> np.ndarray:
    ny = n if ny is None else ny
    out = np.zeros((ny, n), dtype=dtype)
    fn = lib().orc_fbm_init_f64 if dtype == np.float64 else lib().orc_fbm_init_f32
    if fn(_ptr(out), n, ny, ctypes.byref(params)) != 0:
        raise ValueError("orc_fbm_init failed")
    return out


def new_state(n: int, dtype=np.float64, ny: int | None = None) -> np.ndarray:
    ny = n if ny is None else ny
    return np.zeros((NUM_PLANES, ny, n), dtype=dtype)


def run(state: np.ndarray, settings: ErosionSettings, nsteps: int, flags: int = STEP_DEFAULT,
        rain_mode: int = RAIN_OFF, rain_amount: float = 0.0, seed: int = 0, terrain: int = 0,
        step0: int = 0) -> np.ndarray:
    """nsteps GRID steps in place; returns state."""
    _, ny, nx = state.shape
    fn = lib().orc_run_f64 if state.dtype == np.float64 else lib().orc_run_f32
    rc = fn(_ptr(state), nx, ny, ctypes.byref(settings), rain_mode, rain_amount, seed, terrain,
            step0, nsteps, flags)
    if rc != 0:
        raise ValueError("orc_run failed")
    return state


def run_pass(state: np.ndarray, settings: ErosionSettings, pass_id: int, rain_mode: int = RAIN_OFF,
             rain_amount: float = 0.0, seed: int = 0, terrain: int = 0, step: int = 0) -> np.ndarray:
    _, ny, nx = state.shape
    fn = lib().orc_pass_f64 if state.dtype == np.float64 else lib().orc_pass_f32
    rc = fn(_ptr(state), nx, ny, ctypes.byref(settings), rain_mode, rain_amount, seed, terrain,
            step, pass_id)
    if rc != 0:
        raise ValueError("orc_pass failed")
    return state


def gen_exp_f32(x: float) -> float:
    """FP32 exp of the reference generator's exp mask (orc::gen_exp)."""
    return lib().orc_gen_exp_f32(x)


def hash_sin_f32(x: float) -> float:
    """FP32 sin of the reference particle spawn hash (orc::hash_sin)."""
    return lib().orc_hash_sin_f32(x)


def smooth_flags(state: np.ndarray) -> np.ndarray:
    """Extremum flags of the smoothing pass per cell (bit 0 rock, bit 1
    surface; border cells 0), from the state's rock and dirt planes."""
    _, ny, nx = state.shape
    out = np.zeros((ny, nx), np.uint8)
    fn = lib().orc_smooth_flags_f64 if state.dtype == np.float64 else lib().orc_smooth_flags_f32
    ct = ctypes.c_double if state.dtype == np.float64 else ctypes.c_float
    st = np.ascontiguousarray(state)
    rc = fn(st.ctypes.data_as(ctypes.POINTER(ct)), nx, ny, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    assert rc == 0
    return out


def ref_simplex(x: float, y: float, dtype=np.float64) -> float:
    """gln_simplex restated (glsl/simplex_noise.glsl:377-402)."""
    return lib().orc_ref_simplex_f64(x, y) if dtype == np.float64 else lib().orc_ref_simplex_f32(x, y)


def ref_sfbm(x: float, y: float, seed: float, scale: float, octaves: int = 8) -> float:
    """gln_sfbm restated with persistence 0.5, lacunarity 2, redistribution 1."""
    return lib().orc_ref_sfbm_f64(x, y, seed, scale, octaves)


def rain_reference(state: np.ndarray, rs: RainSettings, time: float, max_height: float) -> np.ndarray:
    """One reference rain pass (glsl/rain.glsl:32-56) in place; returns state."""
    _, ny, nx = state.shape
    fn = lib().orc_rain_reference_f64 if state.dtype == np.float64 else lib().orc_rain_reference_f32
    if fn(_ptr(state), nx, ny, ctypes.byref(rs), time, max_height) != 0:
        raise ValueError("orc_rain_reference failed")
    return state


def gen_heightmap(n: int, m: MapSettings, dtype=np.float64):
    """Reference heightmap generator: (rock, dirt) planes of n x n."""
    rock = np.zeros((n, n), dtype=dtype)
    dirt = np.zeros((n, n), dtype=dtype)
    fn = lib().orc_gen_heightmap_f64 if dtype == np.float64 else lib().orc_gen_heightmap_f32
    if fn(_ptr(rock), _ptr(dirt), n, ctypes.byref(m)) != 0:
        raise ValueError("orc_gen_heightmap failed")
    return rock, dirt


def elias_hash(px: int, py: int):
    out = (ctypes.c_double * 2)()
    lib().orc_elias_hash(px, py, out)
    return out[0], out[1]


def noised(x: float, y: float):
    """(value, d/dx, d/dy) of the reference gradient noise."""
    out = (ctypes.c_double * 3)()
    lib().orc_noised_f64(x, y, out)
    return out[0], out[1], out[2]


def rollout_loss(state: np.ndarray, settings: ErosionSettings, coef, nsteps: int, target: np.ndarray,
                 flags: int = STEP_DEFAULT, rain_mode: int = RAIN_CONSTANT, rain_amount: float = 0.01,
                 seed: int = 0) -> float:
    """FP64 loss mean((rock + dirt - target)^2) after nsteps from `state`, with
    (Kc, Ks, Kd, Ke) = coef (double precision) and the rest from settings."""
    _, ny, nx = state.shape
    st = np.ascontiguousarray(state, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    tg = np.ascontiguousarray(target, dtype=np.float64)
    out = ctypes.c_double()
    if lib().orc_rollout_loss_f64(_ptr(st), nx, ny, ctypes.byref(settings), _ptr(c), rain_mode, rain_amount,
                                  seed, nsteps, flags, _ptr(tg), ctypes.byref(out)) != 0:
        raise ValueError("orc_rollout_loss_f64 failed")
    return out.value


def rollout_fd_frozen(state: np.ndarray, settings: ErosionSettings, coef, nsteps: int, target: np.ndarray,
                      cells, eps_coef: float = 1e-3, eps_cell: float = 1e-3, flags: int = STEP_DEFAULT,
                      rain_mode: int = RAIN_CONSTANT, rain_amount: float = 0.01, seed: int = 0):
    """FP64 central differences of rollout_loss with every branch frozen at the
    unperturbed run: (dL/d(Kc,Ks,Kd,Ke) at steps eps_coef * coef, dL/drock0
    at cells with step eps_cell).  At most 12 cells."""
    _, ny, nx = state.shape
    st = np.ascontiguousarray(state, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    tg = np.ascontiguousarray(target, dtype=np.float64)
    cl = np.ascontiguousarray(cells, dtype=np.int32)
    g = np.zeros(4 + len(cl))
    if lib().orc_rollout_fd_frozen(_ptr(st), nx, ny, ctypes.byref(settings), _ptr(c), rain_mode, rain_amount,
                                   seed, nsteps, flags, _ptr(tg), cl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   len(cl), eps_coef, eps_cell, _ptr(g)) != 0:
        raise ValueError("orc_rollout_fd_frozen failed")
    return g[:4], g[4:]


def rollout_grad_dual(state: np.ndarray, settings: ErosionSettings, coef, nsteps: int, target: np.ndarray,
                      cells, flags: int = STEP_DEFAULT, rain_mode: int = RAIN_CONSTANT,
                      rain_amount: float = 0.01, seed: int = 0):
    """Dual-number FP64 rollout: (loss, dL/d(Kc,Ks,Kd,Ke), dL/drock0[cells])."""
    _, ny, nx = state.shape
    st = np.ascontiguousarray(state, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    tg = np.ascontiguousarray(target, dtype=np.float64)
    cl = np.ascontiguousarray(cells, dtype=np.int32)
    g = np.zeros(4 + len(cl))
    out = ctypes.c_double()
    if lib().orc_rollout_grad_dual(_ptr(st), nx, ny, ctypes.byref(settings), _ptr(c), rain_mode, rain_amount,
                                   seed, nsteps, flags, _ptr(tg), cl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   len(cl), ctypes.byref(out), _ptr(g)) != 0:
        raise ValueError("orc_rollout_grad_dual failed")
    return out.value, g[:4], g[4:]
