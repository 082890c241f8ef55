"""ctypes binding of libhydrogen.so (the C ABI declared in include/hg.h).

The library is the product; this module only marshals arguments.  Loading
fails loudly when the built library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))   # hydro-gen_amd/
LIB_PATH = os.path.join(PKG_ROOT, "libhydrogen.so")

# ---- enums mirrored from include/hg.h -------------------------------------
HG_OK, HG_ERR_INVALID, HG_ERR_HIP, HG_ERR_OOM, HG_ERR_STATE, HG_ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5
HG_DROPLETS_NOALLOC_MAX = 4194304   # droplets per iteration of hg_droplets_noalloc
RAIN_OFF, RAIN_CONSTANT, RAIN_UNIFORM, RAIN_REFERENCE = 0, 1, 2, 3
CH_HEIGHT, CH_FLUX, CH_VELOCITY, CH_SEDIMENT = 0, 1, 2, 3
(P_ROCK, P_DIRT, P_WATER, P_FLUX_L, P_FLUX_R, P_FLUX_T, P_FLUX_B, P_SED_ROCK, P_SED_DIRT,
 P_VEL_U, P_VEL_V, P_VEL_DEPTH, P_THERM_L, P_THERM_R, P_THERM_T, P_THERM_B, P_THERM_LT,
 P_THERM_RT, P_THERM_LB, P_THERM_RB) = range(20)
NUM_PLANES = 20
NUM_STATE_PLANES = 9
STEP_RAIN, STEP_THERMAL, STEP_SMOOTH, STEP_PER_PASS, STEP_KEEP_VELOCITY = 0x1, 0x2, 0x4, 0x8, 0x10
STEP_PER_TERRAIN = 0x20
STEP_DEFAULT = STEP_RAIN | STEP_THERMAL | STEP_SMOOTH
(PASS_RAIN, PASS_FLUX, PASS_EROSION, PASS_TRANSPORT, PASS_THERMAL_FLUX_0, PASS_THERMAL_TRANSPORT_0,
 PASS_THERMAL_FLUX_1, PASS_THERMAL_TRANSPORT_1, PASS_SMOOTH) = range(9)
NUM_PASSES = 9

# Every exported symbol of include/hg.h (checked by the ABI test).
EXPORTS = [
    "hg_create", "hg_destroy", "hg_last_error", "hg_size", "hg_set_erosion", "hg_get_erosion",
    "hg_set_rain", "hg_set_rain_time", "hg_init_fbm", "hg_gen_heightmap", "hg_rain", "hg_step_grid", "hg_run_pass", "hg_get_step",
    "hg_set_step", "hg_read", "hg_write", "hg_save_height", "hg_load_height", "hg_read_plane", "hg_write_plane", "hg_plane_device_ptr",
    "hg_sync", "hg_get_stream", "hg_set_timing", "hg_kernel_stats", "hg_reset_timing",
    "hg_set_droplets", "hg_step_particle", "hg_particle_stats", "hg_droplets_noalloc", "hg_adjoint_rollout",
    "hg_set_droplet_order", "hg_get_droplet_order", "hg_set_adjoint_fusion", "hg_get_adjoint_fusion",
    "hg_ref_particles_init", "hg_ref_particles_read", "hg_ref_particles_write", "hg_step_particle_reference",
    "hg_ref_particle_stats", "hg_set_ref_particle_dims",
    "hg_tile_variant_count", "hg_tile_variant_info", "hg_set_tiles", "hg_get_tiles", "hg_set_xcd_order", "hg_autotune",
    "hg_set_sweep", "hg_get_sweep",
    "hg_batch_create", "hg_batch_destroy", "hg_batch_last_error", "hg_batch_count",
    "hg_batch_set_erosion", "hg_batch_set_rain", "hg_batch_init_fbm", "hg_batch_set_tiles", "hg_batch_autotune",
    "hg_batch_set_persistent", "hg_batch_set_xcd_order",
    "hg_batch_step", "hg_batch_capture", "hg_batch_run", "hg_batch_sync", "hg_batch_get_step",
    "hg_batch_read_plane", "hg_batch_write_plane",
]
KERNEL_HYDRO, KERNEL_THERMAL, KERNEL_SWEEP = 0, 1, 2


class ErosionSettings(ctypes.Structure):
    """hg_erosion_settings = Erosion_data (glsl/bindings.glsl:39-60), 96 bytes."""
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
    """hg_rain_settings = Rain_data (glsl/bindings.glsl:62-68), 32 bytes."""
    _fields_ = [("amount", ctypes.c_float), ("mountain_thresh", ctypes.c_float),
                ("mountain_multip", ctypes.c_float), ("period", ctypes.c_int32),
                ("drops", ctypes.c_float), ("_pad", ctypes.c_uint32 * 3)]


class MapSettings(ctypes.Structure):
    """hg_map_settings = Map_settings_data (glsl/bindings.glsl:70-97), 96 bytes."""
    _fields_ = [("max_height", ctypes.c_float), ("max_dirt", ctypes.c_float), ("hmap_dims", ctypes.c_int32 * 2),
                ("height_mult", ctypes.c_float), ("water_lvl", ctypes.c_float), ("seed", ctypes.c_float),
                ("persistance", ctypes.c_float), ("lacunarity", ctypes.c_float), ("scale", ctypes.c_float),
                ("redistribution", ctypes.c_float), ("octaves", ctypes.c_int32), ("fake_erosion", ctypes.c_uint32),
                ("mask_round", ctypes.c_uint32), ("mask_exp", ctypes.c_uint32), ("mask_power", ctypes.c_uint32),
                ("mask_slope", ctypes.c_uint32), ("uplift", ctypes.c_uint32), ("uplift_scale", ctypes.c_float),
                ("domain_warp", ctypes.c_int32), ("domain_warp_scale", ctypes.c_float), ("terrace", ctypes.c_int32),
                ("terrace_scale", ctypes.c_float), ("_pad", ctypes.c_uint32)]


def map_settings(**kw) -> MapSettings:
    """Map settings with the reference defaults (src/settings.hpp:29-55); the
    reference draws its seed from rand(), here it is a parameter (1234.5)."""
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


def erosion_settings(Kc=0.05, Ks=0.02, Kd=0.02, Ke=0.01, talus_deg=35.0, slippage=0.1, G=9.81,
                     dt=0.005, Kconv=0.001, E=1.0) -> ErosionSettings:
    """Scalar coefficients broadcast to both layers (rock, dirt).  Unnamed
    fields take the reference grid defaults (src/state.cpp:81-92); the talus
    angle is given in degrees and stored in radians like the reference UI."""
    s = ErosionSettings()
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


def rain_settings(amount=0.01, mountain_thresh=0.55, mountain_multip=0.05, period=512,
                  drops=0.02) -> RainSettings:
    """Rain_data; the defaults are the reference's (src/settings.hpp:15-21)."""
    r = RainSettings()
    r.amount = amount
    r.mountain_thresh = mountain_thresh
    r.mountain_multip = mountain_multip
    r.period = period
    r.drops = drops
    return r


def fbm_params(seed=1, octaves=6, lacunarity=2.0, gain=0.5, base_freq=1.0 / 256, lo=0.0,
               hi=64.0) -> FbmParams:
    p = FbmParams()
    p.seed = seed; p.octaves = octaves; p.lacunarity = lacunarity; p.gain = gain
    p.base_freq = base_freq; p.height_lo = lo; p.height_hi = hi
    return p


class RefParticle(ctypes.Structure):
    """hg_ref_particle: Particle of glsl/bindings.glsl:101-111, std430, 48 bytes."""
    _fields_ = [("sc", ctypes.c_float), ("iters", ctypes.c_int32), ("position", ctypes.c_float * 2),
                ("velocity", ctypes.c_float * 2), ("volume", ctypes.c_float), ("_pad0", ctypes.c_float),
                ("sediment", ctypes.c_float * 2), ("to_kill", ctypes.c_uint32), ("_pad1", ctypes.c_uint32)]


def reference_particle_settings(particle_count: int) -> "ErosionSettings":
    """The reference's particle-mode Erosion_data defaults (src/state.cpp:60-78)."""
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


class DropletSettings(ctypes.Structure):
    """hg_droplet_settings (include/hg.h)."""
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
    """Defaults: the particle configuration (2^20 droplets, 64 steps, ...)."""
    d = DropletSettings()
    d.count = count; d.max_lifetime = max_lifetime; d.inertia = inertia; d.capacity = capacity
    d.deposition = deposition; d.erosion = erosion; d.evaporation = evaporation
    d.gravity = gravity; d.min_slope = min_slope; d.brush_radius = brush_radius
    d.init_water = init_water; d.init_speed = init_speed; d.seed = seed
    return d


class HGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hg error {code}: {msg}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    """Load libhydrogen.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    L.hg_create.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(vp)]
    L.hg_destroy.argtypes = [vp]
    L.hg_last_error.argtypes = [vp]
    L.hg_last_error.restype = ctypes.c_char_p
    L.hg_size.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32)]
    L.hg_set_erosion.argtypes = [vp, ctypes.POINTER(ErosionSettings)]
    L.hg_get_erosion.argtypes = [vp, ctypes.POINTER(ErosionSettings)]
    L.hg_set_rain.argtypes = [vp, ctypes.POINTER(RainSettings), ctypes.c_int, ctypes.c_uint64]
    L.hg_set_rain_time.argtypes = [vp, ctypes.c_float, ctypes.c_float]
    L.hg_init_fbm.argtypes = [vp, ctypes.POINTER(FbmParams)]
    L.hg_save_height.argtypes = [vp, ctypes.c_char_p]
    L.hg_load_height.argtypes = [vp, ctypes.c_char_p]
    L.hg_gen_heightmap.argtypes = [vp, ctypes.POINTER(MapSettings)]
    L.hg_rain.argtypes = [vp]
    L.hg_step_grid.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32]
    L.hg_run_pass.argtypes = [vp, ctypes.c_int]
    L.hg_get_step.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.hg_set_step.argtypes = [vp, ctypes.c_uint64]
    L.hg_read.argtypes = [vp, ctypes.c_int, fp, ctypes.c_size_t]
    L.hg_write.argtypes = [vp, ctypes.c_int, fp, ctypes.c_size_t]
    L.hg_read_plane.argtypes = [vp, ctypes.c_int, fp, ctypes.c_size_t]
    L.hg_write_plane.argtypes = [vp, ctypes.c_int, fp, ctypes.c_size_t]
    L.hg_plane_device_ptr.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.hg_sync.argtypes = [vp]
    L.hg_get_stream.argtypes = [vp, ctypes.POINTER(vp)]
    L.hg_set_timing.argtypes = [vp, ctypes.c_int]
    L.hg_kernel_stats.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                  ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]
    L.hg_reset_timing.argtypes = [vp]
    L.hg_set_droplets.argtypes = [vp, ctypes.POINTER(DropletSettings)]
    L.hg_adjoint_rollout.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, fp, fp,
                                     ctypes.POINTER(ctypes.c_double)]
    L.hg_step_particle.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32]
    L.hg_particle_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)]
    L.hg_droplets_noalloc.argtypes = [vp, vp, ctypes.c_uint32, ctypes.POINTER(DropletSettings), ctypes.c_uint32,
                                      ctypes.c_uint32, vp, vp, vp]
    L.hg_ref_particles_init.argtypes = [vp, ctypes.c_uint32]
    L.hg_ref_particles_read.argtypes = [vp, ctypes.POINTER(RefParticle), ctypes.c_uint32]
    L.hg_ref_particles_write.argtypes = [vp, ctypes.POINTER(RefParticle), ctypes.c_uint32]
    L.hg_step_particle_reference.argtypes = [vp, ctypes.c_uint32, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                             ctypes.c_uint32]
    L.hg_ref_particle_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.hg_set_ref_particle_dims.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    ip = ctypes.POINTER(ctypes.c_int)
    L.hg_tile_variant_count.argtypes = [ctypes.c_int]
    L.hg_tile_variant_info.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip, ip, ip]
    L.hg_set_tiles.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.hg_set_droplet_order.argtypes = [vp, ctypes.c_int]
    L.hg_get_droplet_order.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.hg_set_adjoint_fusion.argtypes = [vp, ctypes.c_int]
    L.hg_get_adjoint_fusion.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.hg_set_sweep.argtypes = [vp, ctypes.c_int]
    L.hg_get_sweep.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.hg_get_tiles.argtypes = [vp, ip, ip]
    L.hg_set_xcd_order.argtypes = [vp, ctypes.c_int]
    L.hg_autotune.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ip, ip]
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.hg_batch_create.argtypes = [ctypes.c_int, ctypes.c_int, u32p, ctypes.POINTER(vp)]
    L.hg_batch_destroy.argtypes = [vp]
    L.hg_batch_last_error.argtypes = [vp]
    L.hg_batch_last_error.restype = ctypes.c_char_p
    L.hg_batch_count.argtypes = [vp, ip]
    L.hg_batch_set_erosion.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ErosionSettings)]
    L.hg_batch_set_rain.argtypes = [vp, ctypes.POINTER(RainSettings), ctypes.c_int, ctypes.c_uint64]
    L.hg_batch_init_fbm.argtypes = [vp, ctypes.c_int, ctypes.POINTER(FbmParams)]
    L.hg_batch_set_tiles.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.hg_batch_autotune.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ip, ip]
    L.hg_batch_step.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32]
    L.hg_batch_set_persistent.argtypes = [vp, ctypes.c_int]
    L.hg_batch_set_xcd_order.argtypes = [vp, ctypes.c_int]
    L.hg_batch_capture.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32]
    L.hg_batch_run.argtypes = [vp]
    L.hg_batch_sync.argtypes = [vp]
    L.hg_batch_get_step.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.hg_batch_read_plane.argtypes = [vp, ctypes.c_int, ctypes.c_int, fp, ctypes.c_size_t]
    L.hg_batch_write_plane.argtypes = [vp, ctypes.c_int, ctypes.c_int, fp, ctypes.c_size_t]
    _lib = L
    return L


def _fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Terrain:
    """One n x n terrain on one HIP device: the owner of an hg_ctx.

    Method names follow the reference erosion layer (src/erosion.hpp:38-47):
    dispatch_grid = one or more GRID steps, dispatch_grid_rain = one rain pass.
    """

    def __init__(self, n: int, device: int = 0, particle_count: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.hg_create(device, n, particle_count, ctypes.byref(h))
        if rc != HG_OK:
            raise HGError(rc, self.lib.hg_last_error(None).decode())
        self.h = h
        self.n = n

    # -- lifetime --
    def close(self):
        if getattr(self, "h", None):
            self.lib.hg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int):
        if rc != HG_OK:
            raise HGError(rc, self.lib.hg_last_error(self.h).decode())

    # -- settings --
    def set_erosion(self, s: ErosionSettings):
        self._check(self.lib.hg_set_erosion(self.h, ctypes.byref(s)))

    def get_erosion(self) -> ErosionSettings:
        s = ErosionSettings()
        self._check(self.lib.hg_get_erosion(self.h, ctypes.byref(s)))
        return s

    def set_rain(self, mode: int, amount: float = 0.01, seed: int = 0, settings: RainSettings | None = None):
        r = settings if settings is not None else rain_settings(amount)
        self._check(self.lib.hg_set_rain(self.h, ctypes.byref(r), mode, seed))

    def set_rain_time(self, time: float, max_height: float = 256.0):
        """Time seed and map max height of RAIN_REFERENCE (hg_set_rain_time)."""
        self._check(self.lib.hg_set_rain_time(self.h, time, max_height))

    def save_height(self, path: str):
        """HEIGHT channel to a .npy file of shape (n, n, 4) (hg_save_height)."""
        self._check(self.lib.hg_save_height(self.h, os.fsencode(path)))

    def load_height(self, path: str):
        """HEIGHT channel from a .npy file written by save_height (hg_load_height)."""
        self._check(self.lib.hg_load_height(self.h, os.fsencode(path)))

    def gen_heightmap(self, m: MapSettings):
        """Reference heightmap generator (hg_gen_heightmap)."""
        self._check(self.lib.hg_gen_heightmap(self.h, ctypes.byref(m)))

    def init_fbm(self, p: FbmParams):
        self._check(self.lib.hg_init_fbm(self.h, ctypes.byref(p)))

    # -- stepping --
    def dispatch_grid(self, nsteps: int = 1, flags: int = STEP_DEFAULT):
        self._check(self.lib.hg_step_grid(self.h, nsteps, flags))

    def adjoint_rollout(self, nsteps: int, k_ckpt: int, dL_dH_final: np.ndarray,
                        flags: int = STEP_DEFAULT):
        """Run nsteps from the current state and return (dL/dH0, dL/d(Kc, Ks, Kd, Ke))."""
        g = np.ascontiguousarray(dL_dH_final, dtype=np.float32)
        assert g.shape == (self.n, self.n)
        out = np.empty((self.n, self.n), dtype=np.float32)
        coef = (ctypes.c_double * 4)()
        self._check(self.lib.hg_adjoint_rollout(self.h, nsteps, k_ckpt, flags, _fptr(g), _fptr(out), coef))
        return out, np.array(list(coef))

    def set_droplets(self, d: DropletSettings):
        self._check(self.lib.hg_set_droplets(self.h, ctypes.byref(d)))

    def set_droplet_order(self, mode: int):
        """hg_set_droplet_order: 0 thread j runs droplet j, 1 droplets sorted
        by the tile of their start cell."""
        self._check(self.lib.hg_set_droplet_order(self.h, mode))

    def get_droplet_order(self) -> int:
        v = ctypes.c_int()
        self._check(self.lib.hg_get_droplet_order(self.h, ctypes.byref(v)))
        return v.value

    def set_adjoint_fusion(self, on: int):
        """hg_set_adjoint_fusion: bit 0 the erosion and flux adjoints of a step
        in one kernel, bit 1 the smoothing and thermal layer-1 adjoints, bit 2
        the transport's far-tap scatter in the thermal layer-0 launch, bit 3
        the transport's gather in bit 1's launch (all set by default); a clear
        bit runs them one launch each.  Every combination gives the same
        gradients."""
        self._check(self.lib.hg_set_adjoint_fusion(self.h, on))

    def get_adjoint_fusion(self) -> int:
        v = ctypes.c_int()
        self._check(self.lib.hg_get_adjoint_fusion(self.h, ctypes.byref(v)))
        return v.value

    def dispatch_particle(self, iterations: int = 1, flags: int = 0):
        """Droplet iterations (Erosion::dispatch_particle counterpart)."""
        self._check(self.lib.hg_step_particle(self.h, iterations, flags))

    # -- the reference's particle mode --
    def ref_particles_init(self, count: int):
        """`count` zeroed particles and a zeroed momentum map (hg_ref_particles_init)."""
        self._check(self.lib.hg_ref_particles_init(self.h, count))
        self._refp_count = count

    def set_ref_particle_dims(self, dims_x: int = 0, dims_y: int = 0):
        """Map_settings_data.hmap_dims of the particle spawn range and border
        test (hg_set_ref_particle_dims); the reference keeps (1024, 1024) on
        every map size, (0, 0) uses the grid's n."""
        self._check(self.lib.hg_set_ref_particle_dims(self.h, dims_x, dims_y))

    def read_ref_particles(self):
        arr = (RefParticle * self._refp_count)()
        self._check(self.lib.hg_ref_particles_read(self.h, arr, self._refp_count))
        return arr

    def write_ref_particles(self, arr):
        self._check(self.lib.hg_ref_particles_write(self.h, arr, len(arr)))

    def dispatch_particle_reference(self, iterations: int = 1, should_rain: bool = True, time0: float = 0.0,
                                    time_step: float = 0.0, flags: int = STEP_THERMAL | STEP_SMOOTH):
        """Erosion::dispatch_particle of the reference (hg_step_particle_reference)."""
        self._check(self.lib.hg_step_particle_reference(self.h, iterations, int(should_rain), time0, time_step,
                                                         flags))

    def ref_particle_stats(self) -> int:
        g = ctypes.c_uint64()
        self._check(self.lib.hg_ref_particle_stats(self.h, ctypes.byref(g)))
        return g.value

    def particle_stats(self):
        ex, st = ctypes.c_double(), ctypes.c_uint64()
        self._check(self.lib.hg_particle_stats(self.h, ctypes.byref(ex), ctypes.byref(st)))
        return ex.value, st.value

    def dispatch_grid_rain(self):
        self._check(self.lib.hg_rain(self.h))

    def run_pass(self, p: int):
        self._check(self.lib.hg_run_pass(self.h, p))

    @property
    def step(self) -> int:
        v = ctypes.c_uint64()
        self._check(self.lib.hg_get_step(self.h, ctypes.byref(v)))
        return v.value

    @step.setter
    def step(self, v: int):
        self._check(self.lib.hg_set_step(self.h, v))

    def sync(self):
        self._check(self.lib.hg_sync(self.h))

    # -- state I/O --
    def read_plane(self, p: int, out: np.ndarray | None = None) -> np.ndarray:
        """Plane p (hg_read_plane), into `out` (contiguous float32 n x n) if given."""
        a = np.empty((self.n, self.n), dtype=np.float32) if out is None else out
        assert a.dtype == np.float32 and a.shape == (self.n, self.n) and a.flags.c_contiguous
        self._check(self.lib.hg_read_plane(self.h, p, _fptr(a), a.size))
        return a

    def write_plane(self, p: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.shape == (self.n, self.n)
        self._check(self.lib.hg_write_plane(self.h, p, _fptr(a), a.size))

    def read(self, ch: int) -> np.ndarray:
        a = np.empty((self.n, self.n, 4), dtype=np.float32)
        self._check(self.lib.hg_read(self.h, ch, _fptr(a), a.size))
        return a

    def write(self, ch: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.shape == (self.n, self.n, 4)
        self._check(self.lib.hg_write(self.h, ch, _fptr(a), a.size))

    def read_state(self, nplanes: int = NUM_PLANES) -> np.ndarray:
        st = np.empty((nplanes, self.n, self.n), dtype=np.float32)
        for p in range(nplanes):
            self.read_plane(p, st[p])
        return st

    def write_state(self, st: np.ndarray):
        for p in range(st.shape[0]):
            self.write_plane(p, st[p])

    def device_ptr(self, p: int) -> int:
        v = ctypes.c_void_p()
        self._check(self.lib.hg_plane_device_ptr(self.h, p, ctypes.byref(v)))
        return v.value

    def stream(self) -> int:
        v = ctypes.c_void_p()
        self._check(self.lib.hg_get_stream(self.h, ctypes.byref(v)))
        return v.value

    # -- timing --
    def set_timing(self, on: bool):
        self._check(self.lib.hg_set_timing(self.h, 1 if on else 0))

    def kernel_stats(self) -> dict:
        out = {}
        i = 0
        while True:
            name = ctypes.c_char_p()
            n = ctypes.c_uint64()
            ms = ctypes.c_double()
            rc = self.lib.hg_kernel_stats(self.h, i, ctypes.byref(name), ctypes.byref(n), ctypes.byref(ms))
            if rc != HG_OK:
                break
            out[name.value.decode()] = (n.value, ms.value)
            i += 1
        return out

    def reset_timing(self):
        self._check(self.lib.hg_reset_timing(self.h))

    # -- tiles / autotuning --
    def set_tiles(self, hydro: int = -1, thermal: int = -1):
        self._check(self.lib.hg_set_tiles(self.h, hydro, thermal))

    def get_tiles(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.hg_get_tiles(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_sweep(self, variant: int):
        """One-kernel steps (hg_set_sweep): a KERNEL_SWEEP variant, or -1 for
        the hydro and thermal kernels."""
        self._check(self.lib.hg_set_sweep(self.h, variant))

    def get_sweep(self) -> int:
        v = ctypes.c_int()
        self._check(self.lib.hg_get_sweep(self.h, ctypes.byref(v)))
        return v.value

    def set_xcd_order(self, mode: int):
        """Workgroup-to-tile order of the fused kernels (hg_set_xcd_order):
        0 row-major, 1 contiguous runs per XCD, 2 column strips per XCD."""
        self._check(self.lib.hg_set_xcd_order(self.h, int(mode)))

    def autotune(self, steps: int = 10, flags: int = STEP_DEFAULT):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.hg_autotune(self.h, steps, flags, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def tile_variants(kernel: int) -> list:
    """[(tile_x, tile_y, halo, threads), ...] of the compiled variants of a fused kernel."""
    L = load()
    out = []
    for v in range(L.hg_tile_variant_count(kernel)):
        a = [ctypes.c_int() for _ in range(4)]
        if L.hg_tile_variant_info(kernel, v, *[ctypes.byref(x) for x in a]) == HG_OK:
            out.append(tuple(x.value for x in a))
    return out


def droplets_noalloc(rock_ptr: int, dirt_ptr: int, n: int, s: "DropletSettings", first_iteration: int,
                     iterations: int, stream: int = 0, exported_ptr: int = 0, steps_ptr: int = 0) -> None:
    """hg_droplets_noalloc: droplet iterations on caller-owned device planes
    (device addresses as ints), allocating nothing; exported_ptr / steps_ptr
    are device addresses of a float64 / uint64 receiving the totals (0 = none)."""
    L = load()
    rc = L.hg_droplets_noalloc(ctypes.c_void_p(rock_ptr), ctypes.c_void_p(dirt_ptr), n, ctypes.byref(s),
                               first_iteration, iterations, ctypes.c_void_p(stream or None),
                               ctypes.c_void_p(exported_ptr or None), ctypes.c_void_p(steps_ptr or None))
    if rc != HG_OK:
        raise HGError(rc, L.hg_last_error(None).decode())


class Batch:
    """Many square terrains of uneven sides advanced in lockstep (hg_batch_*)."""

    def __init__(self, sizes, device: int = 0):
        self.lib = load()
        self.sizes = [int(x) for x in sizes]
        arr = (ctypes.c_uint32 * len(self.sizes))(*self.sizes)
        h = ctypes.c_void_p()
        rc = self.lib.hg_batch_create(device, len(self.sizes), arr, ctypes.byref(h))
        if rc != HG_OK:
            raise HGError(rc, self.lib.hg_batch_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.hg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != HG_OK:
            raise HGError(rc, self.lib.hg_batch_last_error(self.h).decode())

    def set_erosion(self, s: ErosionSettings, terrain: int = -1):
        self._check(self.lib.hg_batch_set_erosion(self.h, terrain, ctypes.byref(s)))

    def set_rain(self, mode: int, amount: float = 0.01, seed: int = 0):
        r = rain_settings(amount)
        self._check(self.lib.hg_batch_set_rain(self.h, ctypes.byref(r), mode, seed))

    def init_fbm(self, terrain: int, p: FbmParams):
        self._check(self.lib.hg_batch_init_fbm(self.h, terrain, ctypes.byref(p)))

    def set_tiles(self, hydro: int = -1, thermal: int = -1):
        self._check(self.lib.hg_batch_set_tiles(self.h, hydro, thermal))

    def autotune(self, steps: int = 4, flags: int = STEP_DEFAULT):
        """Select the fastest tile variants for this batch (hg_batch_autotune);
        the state and step counter are restored and a captured graph is dropped."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.hg_batch_autotune(self.h, steps, flags, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def step(self, nsteps: int = 1, flags: int = STEP_DEFAULT):
        self._check(self.lib.hg_batch_step(self.h, nsteps, flags))

    def set_xcd_order(self, on: bool):
        self._check(self.lib.hg_batch_set_xcd_order(self.h, 1 if on else 0))

    def set_persistent(self, on: bool):
        self._check(self.lib.hg_batch_set_persistent(self.h, 1 if on else 0))

    def capture(self, nsteps: int, flags: int = STEP_DEFAULT):
        self._check(self.lib.hg_batch_capture(self.h, nsteps, flags))

    def run(self):
        self._check(self.lib.hg_batch_run(self.h))

    def sync(self):
        self._check(self.lib.hg_batch_sync(self.h))

    @property
    def step_count(self) -> int:
        v = ctypes.c_uint64()
        self._check(self.lib.hg_batch_get_step(self.h, ctypes.byref(v)))
        return v.value

    def read_plane(self, terrain: int, p: int) -> np.ndarray:
        n = self.sizes[terrain]
        a = np.empty((n, n), dtype=np.float32)
        self._check(self.lib.hg_batch_read_plane(self.h, terrain, p, _fptr(a), a.size))
        return a

    def write_plane(self, terrain: int, p: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32)
        self._check(self.lib.hg_batch_write_plane(self.h, terrain, p, _fptr(a), a.size))

    def read_state(self, terrain: int) -> np.ndarray:
        return np.stack([self.read_plane(terrain, p) for p in range(NUM_STATE_PLANES)])

    def write_state(self, terrain: int, st: np.ndarray):
        for p in range(NUM_STATE_PLANES):
            self.write_plane(terrain, p, st[p])
