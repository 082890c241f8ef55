/*
 * hg.h — C ABI of the MI355X (gfx950) erosion library.
 *
 * This is the drop-in boundary for hydro-gen's erosion dispatch layer
 * (src/erosion.hpp:38-47, src/state.hpp:78-85).  Every entry point takes
 * plain pointers and sizes; no C++ or torch types cross it.  Status codes are
 * int: 0 = OK, < 0 = error; hg_last_error() returns the message of the last
 * failing call on a context.
 *
 * Threading contract: one host thread per context.  All device work is
 * enqueued on the context's HIP stream and is asynchronous until hg_sync()
 * or a read.
 */
#ifndef HG_H
#define HG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HG_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum {
    HG_OK = 0,
    HG_ERR_INVALID = -1,   /* bad argument (null pointer, size, enum, non-finite setting) */
    HG_ERR_HIP = -2,       /* a HIP runtime call failed */
    HG_ERR_OOM = -3,       /* device allocation failed */
    HG_ERR_STATE = -4,     /* call not valid in the context's current state */
    HG_ERR_UNSUPPORTED = -5
};

typedef struct hg_ctx hg_ctx;

/* ---- settings blocks --------------------------------------------------- */

/* Erosion_data, glsl/bindings.glsl:39-60: same field order and std140 offsets
 * (96 bytes), so a reference Erosion_settings::data can be copied in as is.
 * Index [0] of the per-layer vec2 fields is rock, [1] is dirt. */
typedef struct hg_erosion_settings {
    uint32_t particle_count; /*  0 */
    float Kc;                /*  4  sediment capacity */
    float Kalpha[2];         /*  8  talus angle per layer, radians */
    float Kconv;             /* 16  rock->dirt sediment conversion */
    float _pad0;             /* 20 */
    float Ks[2];             /* 24  dissolving constant per layer */
    float Kd[2];             /* 32  deposition constant per layer */
    float Ke;                /* 40  evaporation */
    float energy_kept;       /* 44  pipe flux damping E */
    float Kspeed[2];         /* 48  thermal slippage speed per layer */
    float G;                 /* 56  gravity */
    float d_t;               /* 60  time step */
    float density;           /* 64  (particle mode) */
    float init_volume;       /* 68 */
    float friction;          /* 72 */
    float inertia;           /* 76 */
    float min_volume;        /* 80 */
    float min_velocity;      /* 84 */
    uint32_t ttl;            /* 88 */
    uint32_t _pad1;          /* 92 */
} hg_erosion_settings;

/* Rain_data, glsl/bindings.glsl:62-68 (std140, 32 bytes). */
typedef struct hg_rain_settings {
    float amount;            /*  0 */
    float mountain_thresh;   /*  4 */
    float mountain_multip;   /*  8 */
    int32_t period;          /* 12 */
    float drops;             /* 16 */
    uint32_t _pad[3];        /* 20 */
} hg_rain_settings;

/* Map_settings_data, glsl/bindings.glsl:70-97 (std140, 96 bytes): settings of
 * the reference heightmap generator (defaults src/settings.hpp:29-55). */
typedef struct hg_map_settings {
    float max_height;        /*  0 */
    float max_dirt;          /*  4 */
    int32_t hmap_dims[2];    /*  8  not used: the grid is the context's n x n */
    float height_mult;       /* 16 */
    float water_lvl;         /* 20  not used by the generator */
    float seed;              /* 24 */
    float persistance;       /* 28 */
    float lacunarity;        /* 32 */
    float scale;             /* 36 */
    float redistribution;    /* 40  not used: the shader passes 1.0 */
    int32_t octaves;         /* 44 */
    uint32_t fake_erosion;   /* 48 */
    uint32_t mask_round;     /* 52 */
    uint32_t mask_exp;       /* 56 */
    uint32_t mask_power;     /* 60 */
    uint32_t mask_slope;     /* 64 */
    uint32_t uplift;         /* 68 */
    float uplift_scale;      /* 72 */
    int32_t domain_warp;     /* 76 */
    float domain_warp_scale; /* 80 */
    int32_t terrace;         /* 84 */
    float terrace_scale;     /* 88 */
    uint32_t _pad;           /* 92 */
} hg_map_settings;

/* Rain modes.
 *  CONSTANT:  water += amount on every cell, every step.
 *  UNIFORM:   water += amount * u, u = (Philox4x32-10(ctr, key) >> 8) * 2^-24 with
 *             ctr = {x >> 2, y, step, terrain}, key = {seed lo, seed hi}, output
 *             word (x & 3).  Resampled every step; bit-identical host/device.
 *  REFERENCE: the reference's own rain (glsl/rain.glsl:32-56): amount times
 *             max(0, simplex fBm of the cell) plus the mountain bonus, from all
 *             Rain_data fields; its wall-clock seed is replaced by the explicit
 *             time of hg_set_rain_time.  hg_step_grid with HG_STEP_RAIN applies
 *             it before step s when (s + 1) % period == 0, the reference's
 *             cadence (src/main.cpp:313-319); hg_rain applies it once. */
enum { HG_RAIN_OFF = 0, HG_RAIN_CONSTANT = 1, HG_RAIN_UNIFORM = 2, HG_RAIN_REFERENCE = 3 };

/* Simplex fBm heightmap initialisation (this project's seeded-permutation
 * simplex; see hg_init_fbm).  Heights are min-max rescaled to
 * [height_lo, height_hi] and stored in the rock layer; dirt, water, flux,
 * sediment and velocity are zeroed. */
typedef struct hg_fbm_params {
    uint64_t seed;
    int32_t octaves;
    float lacunarity;
    float gain;
    float base_freq;         /* cycles per cell of octave 0 */
    float height_lo;
    float height_hi;
} hg_fbm_params;

/* ---- state channels ----------------------------------------------------- */

/* RGBA32F channel views in the reference's texture layout
 * (src/state.hpp:59-76):
 *   HEIGHT   = (rock, dirt, water, rock+dirt+water)
 *   FLUX     = (L, R, T, B)
 *   VELOCITY = (u, v, d1+d2, 0)
 *   SEDIMENT = (s_rock, s_dirt, 0, 0)
 * Row-major, x fastest, y = 0 is the "bottom" row (direction B = -y). */
enum { HG_CH_HEIGHT = 0, HG_CH_FLUX = 1, HG_CH_VELOCITY = 2, HG_CH_SEDIMENT = 3 };

/* Single FP32 planes (structure-of-arrays storage on the device). */
enum {
    HG_P_ROCK = 0, HG_P_DIRT, HG_P_WATER,
    HG_P_FLUX_L, HG_P_FLUX_R, HG_P_FLUX_T, HG_P_FLUX_B,
    HG_P_SED_ROCK, HG_P_SED_DIRT,
    HG_P_VEL_U, HG_P_VEL_V, HG_P_VEL_DEPTH,
    /* thermal outflows of the last thermal flux pass (per-pass path only):
     * cross (L, R, T, B) then diagonal (LT, RT, LB, RB) */
    HG_P_THERM_L, HG_P_THERM_R, HG_P_THERM_T, HG_P_THERM_B,
    HG_P_THERM_LT, HG_P_THERM_RT, HG_P_THERM_LB, HG_P_THERM_RB,
    HG_NUM_PLANES
};
#define HG_NUM_STATE_PLANES 9   /* rock..sed_dirt persist across steps */

/* ---- step flags (hg_step_grid) ------------------------------------------ */
#define HG_STEP_RAIN          0x1u  /* apply the rain mode before each step */
#define HG_STEP_THERMAL       0x2u  /* thermal flux+transport for both layers */
#define HG_STEP_SMOOTH        0x4u  /* extremum smoothing pass */
#define HG_STEP_PER_PASS      0x8u  /* one kernel per reference pass (unfused) */
#define HG_STEP_KEEP_VELOCITY 0x10u /* fused path also stores the velocity planes */
#define HG_STEP_PER_TERRAIN   0x20u /* batches: launch each terrain separately (baseline) */
#define HG_STEP_DEFAULT (HG_STEP_RAIN | HG_STEP_THERMAL | HG_STEP_SMOOTH)

/* Single reference passes, for pass-by-pass parity (hg_run_pass). */
enum {
    HG_PASS_RAIN = 0,
    HG_PASS_FLUX = 1,               /* glsl/hydro_flux.glsl */
    HG_PASS_EROSION = 2,            /* glsl/hydro_erosion.glsl */
    HG_PASS_TRANSPORT = 3,          /* glsl/sediment_transport.glsl */
    HG_PASS_THERMAL_FLUX_0 = 4,     /* glsl/thermal_erosion.glsl, t_layer 0 */
    HG_PASS_THERMAL_TRANSPORT_0 = 5,/* glsl/thermal_transport.glsl, t_layer 0 */
    HG_PASS_THERMAL_FLUX_1 = 6,
    HG_PASS_THERMAL_TRANSPORT_1 = 7,
    HG_PASS_SMOOTH = 8,             /* glsl/smoothing.glsl, grid branch */
    HG_NUM_PASSES = 9
};

/* ---- context ------------------------------------------------------------ */

/* Replaces World::gen_textures + Erosion::setup_shaders (src/state.cpp:3-44,
 * src/erosion.cpp:21-74): allocates the n x n state on HIP device `device`
 * and a stream.  particle_count only sets the default droplet count of
 * hg_step_particle (the droplets need no buffers of their own); the
 * reference particle mode's particle buffer comes from hg_ref_particles_init.
 * The settings start at the reference grid defaults (src/state.cpp:81-92). */
int hg_create(int device, uint32_t n, uint32_t particle_count, hg_ctx** out);

/* Replaces World::delete_textures (src/state.cpp:46-55).  NULL is a no-op. */
int hg_destroy(hg_ctx* ctx);

/* Message of the last failing call on ctx (or, when ctx is NULL, of the last
 * failing hg_create or hg_droplets_noalloc on this thread).  Never NULL. */
const char* hg_last_error(const hg_ctx* ctx);

/* Grid side length. */
int hg_size(const hg_ctx* ctx, uint32_t* n);

/* Replaces Erosion_settings::push_data (src/settings.hpp:61-67): the new
 * values apply from the next enqueued pass.  Rejects non-finite values,
 * d_t <= 0 and negative coefficients. */
int hg_set_erosion(hg_ctx* ctx, const hg_erosion_settings* s);
int hg_get_erosion(const hg_ctx* ctx, hg_erosion_settings* s);

/* Replaces Rain_settings::push_data; `mode` is one of HG_RAIN_*. */
int hg_set_rain(hg_ctx* ctx, const hg_rain_settings* s, int mode, uint64_t seed);

/* Time seed of HG_RAIN_REFERENCE (the `time` uniform the reference sets from
 * glfwGetTime, src/erosion.cpp:76-89) and the map's max_height
 * (Map_settings_data, glsl/bindings.glsl:70-71) that scales the mountain
 * bonus.  max_height must be > 0. */
int hg_set_rain_time(hg_ctx* ctx, float time, float max_height);

/* Config heightmap init (in place of World::gen_heightmap,
 * src/state.cpp:116-147).  Resets the step counter to 0. */
int hg_init_fbm(hg_ctx* ctx, const hg_fbm_params* p);

/* The reference heightmap (replaces World::gen_heightmap, src/state.cpp:116-147,
 * and glsl/heightmap.glsl:52-158): domain-warped, derivative-damped gradient
 * noise fBm (erosion_perlfbm), optional ridged simplex uplift, round / exp /
 * power / slope masks and terraces into rock, perlfbm into dirt; water, flux,
 * velocity and sediment are zeroed.  Two choices where GLSL is undefined or
 * inexact: pow(x, 2) of a negative x in the round mask is x * x, and
 * pow(x, 3) of the power mask is x * x * x.  Resets the step counter to 0. */
int hg_gen_heightmap(hg_ctx* ctx, const hg_map_settings* m);

/* ---- stepping ----------------------------------------------------------- */

/* One rain application at the current step counter, as a pass of its own
 * (replaces Erosion::dispatch_grid_rain, src/erosion.cpp:76-89). */
int hg_rain(hg_ctx* ctx);

/* nsteps GRID steps (replaces Erosion::dispatch_grid, src/erosion.cpp:158-200)
 * in the reference pass order: [rain] -> flux -> erosion -> sediment transport
 * -> thermal layer 0 -> thermal layer 1 -> smoothing.  The step counter
 * advances by nsteps. */
int hg_step_grid(hg_ctx* ctx, uint32_t nsteps, uint32_t flags);

/* One reference pass (HG_PASS_*) through the per-pass kernels. */
int hg_run_pass(hg_ctx* ctx, int pass);

int hg_get_step(const hg_ctx* ctx, uint64_t* step);
int hg_set_step(hg_ctx* ctx, uint64_t step);

/* ---- state I/O ---------------------------------------------------------- */

/* Channel I/O in the reference RGBA32F layout; n_floats must be 4*n*n.
 * Writing HEIGHT ignores the .w (total) component; writing VELOCITY ignores .w.
 * Reads synchronise the stream. */
int hg_read(hg_ctx* ctx, int channel, float* rgba, size_t n_floats);
int hg_write(hg_ctx* ctx, int channel, const float* rgba, size_t n_floats);

/* Heightmap files (no counterpart in the reference, which has none): the
 * HEIGHT channel as a NumPy .npy v1.0 file, little-endian float32, shape
 * (n, n, 4) = (rock, dirt, water, total) row-major.  Load requires the same n
 * and ignores the stored total. */
int hg_save_height(hg_ctx* ctx, const char* path);
int hg_load_height(hg_ctx* ctx, const char* path);

/* Plane I/O (HG_P_*), n_floats must be n*n. */
int hg_read_plane(hg_ctx* ctx, int plane, float* dst, size_t n_floats);
int hg_write_plane(hg_ctx* ctx, int plane, const float* src, size_t n_floats);

/* Device address of the current buffer of a plane (valid until the next
 * pass swaps buffers). */
int hg_plane_device_ptr(hg_ctx* ctx, int plane, void** ptr);

/* Block until all work enqueued on the context has finished. */
int hg_sync(hg_ctx* ctx);

/* The context's hipStream_t, as void*. */
int hg_get_stream(hg_ctx* ctx, void** stream);

/* ---- per-kernel timing -------------------------------------------------- */

/* When enabled, every kernel launch is bracketed by HIP events on the
 * context's stream and its duration is accumulated per kernel name.
 * hg_kernel_stats() synchronises and reports kernel i (0-based, in first-
 * launch order); returns HG_ERR_INVALID past the last one. */
int hg_set_timing(hg_ctx* ctx, int enable);
int hg_kernel_stats(hg_ctx* ctx, int i, const char** name, uint64_t* launches,
                    double* total_ms);
int hg_reset_timing(hg_ctx* ctx);

/* ---- particle (droplet) erosion ------------------------------------------ */

/* Droplet model parameters (a mass-conserving Beyer-style model; the
 * reference's own particle mode, glsl/particle.glsl, is a different model
 * that does not conserve mass, and is hg_step_particle_reference below).
 * Terrain height is rock + dirt; erosion takes dirt first, then rock;
 * deposits go to dirt. */
typedef struct hg_droplet_settings {
    uint32_t count;          /* droplets per iteration */
    uint32_t max_lifetime;   /* steps per droplet */
    float inertia;
    float capacity;
    float deposition;
    float erosion;
    float evaporation;
    float gravity;
    float min_slope;
    int32_t brush_radius;    /* 1..6 cells */
    float init_water;
    float init_speed;
    uint64_t seed;
} hg_droplet_settings;

int hg_set_droplets(hg_ctx* ctx, const hg_droplet_settings* s);
/* `iterations` droplet iterations (in place of Erosion::dispatch_particle,
 * src/erosion.cpp:132-156).  Every droplet of an iteration runs in parallel
 * and writes heights with float atomic adds, so results depend on the order
 * the hardware applies them in.  flags: HG_STEP_THERMAL / HG_STEP_SMOOTH run
 * the thermal and smoothing passes after each iteration, like the reference's
 * particle dispatch. */
int hg_step_particle(hg_ctx* ctx, uint32_t iterations, uint32_t flags);
/* Which thread runs which droplet of an iteration (the droplets and their
 * start positions are the same in both modes; only the timing of their
 * atomic updates relative to each other changes):
 *   0  (default) thread j runs droplet j, so the 64 droplets of a wave start
 *      at unrelated places;
 *   1  droplets sorted by the tile of their start cell (16 x 16 cells, or
 *      n/128 for n > 2048), so a wave's droplets start close together and
 *      share cache lines: measured 1.19x faster at 2048^2 with 2^20
 *      droplets, but neighbouring droplets then run in lockstep and do not
 *      see each other's changes within a step, which moves the statistics
 *      further from the serial model (the spread of heights after 4
 *      iterations of 8192 droplets on 256^2: 5.7 % off the serial oracle,
 *      against 0.6-2.6 % in mode 0).
 * Mode 1 keeps count + (number of tiles, at most 16384) words of device
 * memory in the context.  No counterpart in the reference. */
int hg_set_droplet_order(hg_ctx* ctx, int mode);
int hg_get_droplet_order(const hg_ctx* ctx, int* mode);
/* Cumulative since hg_init_fbm: mass exported over the map edge and the
 * number of droplet steps simulated. */
int hg_particle_stats(hg_ctx* ctx, double* exported_mass, uint64_t* droplet_steps);

/* Droplet iterations that allocate nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * droplet iterations of hg_step_particle (without the thermal and smoothing
 * flags) on planes the caller owns, enqueued on `stream` (a hipStream_t; NULL
 * is the default stream) on the current HIP device:
 *   rock, dirt      device pointers to n x n FP32 planes, eroded in place;
 *   s               the droplet model (hg_droplet_settings); at most
 *                   HG_DROPLETS_NOALLOC_MAX droplets per iteration;
 *   first_iteration the first iteration's index (the start positions'
 *                   counter), then first_iteration + 1, ...;
 *   exported_mass, droplet_steps   device pointers (either may be NULL) that
 *                   receive this call's totals, summed in a fixed order.
 * Each workgroup's sums go to a fixed 256 KiB device array of the library (one
 * per GPU, held until the process ends); nothing is allocated, freed or
 * cleared through the runtime.  Calls that overlap on one GPU share that
 * array and race.  Errors: hg_last_error(NULL). */
#define HG_DROPLETS_NOALLOC_MAX 4194304u
int hg_droplets_noalloc(float* rock, float* dirt, uint32_t n, const hg_droplet_settings* s,
                        uint32_t first_iteration, uint32_t iterations, void* stream,
                        double* exported_mass, uint64_t* droplet_steps);

/* ---- the reference's particle mode ---------------------------------------- */

/* Particle of glsl/bindings.glsl:101-111 in its std430 layout (48 bytes), the
 * element of the reference's particle SSBO. */
typedef struct hg_ref_particle {
    float sc;                /*  0  sediment capacity */
    int32_t iters;           /*  4  steps lived; 0 = not spawned */
    float position[2];       /*  8 */
    float velocity[2];       /* 16 */
    float volume;            /* 24 */
    float _pad0;             /* 28 */
    float sediment[2];       /* 32  rock, dirt */
    uint32_t to_kill;        /* 40  GLSL bool */
    uint32_t _pad1;          /* 44 */
} hg_ref_particle;

/* Creates (or re-creates) `count` zeroed particles and zeroes the momentum
 * map, as World::gen_textures does for particle mode (src/state.cpp:3-44).
 * The momentum map (the reference's velocity texture in particle mode,
 * rgba) is held in the four FLUX planes: L = x, R = y, T = z, B = w. */
int hg_ref_particles_init(hg_ctx* ctx, uint32_t count);
/* Map_settings_data.hmap_dims as glsl/particle.glsl:78-79,111-112 reads it
 * for the spawn range (2 + rand * (dims - 4)) and the border test
 * (position >= dims - 2).  The reference leaves it at (1024, 1024) whatever
 * the map size (src/settings.hpp:34); set (1024, 1024) to keep that quirk on
 * another grid size.  Particles outside the grid then read heights of 0 and
 * their corner updates are dropped, like imageLoad / imageStore outside an
 * image.  (0, 0), the default, uses the grid's n.  Both 0 or both in
 * [8, 65536]. */
int hg_set_ref_particle_dims(hg_ctx* ctx, int32_t dims_x, int32_t dims_y);
int hg_ref_particles_read(hg_ctx* ctx, hg_ref_particle* out, uint32_t count);
int hg_ref_particles_write(hg_ctx* ctx, const hg_ref_particle* in, uint32_t count);
/* `iterations` Erosion::dispatch_particle calls (src/erosion.cpp:132-156):
 * particle movement (glsl/particle.glsl:64-136), particle erosion at the 4
 * corners of each particle's quad (glsl/particle_erosion.glsl:22-128),
 * thermal erosion of both layers (flags HG_STEP_THERMAL) and smoothing with
 * its particle branch (HG_STEP_SMOOTH; glsl/smoothing.glsl:22-103).  Iteration
 * j uses time = time0 + j * time_step for the shader's `time` uniform;
 * should_rain as the dispatch's argument.  The erosion settings'
 * particle_count, Kc, Ks, Kd, Kconv, Ke, G, d_t, init_volume, friction,
 * inertia, min_volume, min_velocity and ttl are used as in the shaders.
 * Corner updates of different particles at one pixel are atomic per layer
 * (no per-pixel spinlock). */
int hg_step_particle_reference(hg_ctx* ctx, uint32_t iterations, int should_rain, float time0,
                               float time_step, uint32_t flags);
/* Atomic height updates that gave up retrying and fell back to an unclamped
 * add since hg_ref_particles_init (0 unless the contention is extreme). */
int hg_ref_particle_stats(hg_ctx* ctx, uint64_t* cas_giveups);

/* ---- adjoint -------------------------------------------------------------- */

/* Gradient of a GRID rollout (no counterpart in the reference).  Runs nsteps
 * steps (flags: rain / thermal / smoothing) from the context's current state,
 * saving a checkpoint every k_ckpt steps, then the hand-written reverse sweep.
 * dL_dH_final (n*n host floats) is the gradient of the caller's loss with
 * respect to the final terrain height rock + dirt.  Outputs: dL_dH0 (n*n host
 * floats), the gradient with respect to the initial rock height, and
 * dL_dcoef[4], with respect to (Kc, Ks, Kd, Ke) (Ks and Kd shared by both
 * layers).  Clamps, max/min, the talus test and the erosion cases contribute
 * the derivative of the branch the forward run took.  The context is left at
 * the final state, its step counter advanced by nsteps.  The device workspace
 * ((#checkpoints + s (k + 1)) x 9 + s k x 11 + 66 planes of n*n floats, with
 * k = min(k_ckpt, nsteps) and s = 2, or 1 for a single segment) and a second
 * stream stay with the context for the next call, which reuses them when they
 * are large enough, until hg_destroy. */
/* Reverse-launch grouping of hg_adjoint_rollout (no counterpart in the
 * reference): a combination of the bits below, all set by default.  With a
 * bit set, the adjoints it names share one launch per step, otherwise each has
 * its own; every combination gives the same gradients bit for bit (the float
 * atomics that sum back-traces longer than one cell run in no fixed order in
 * any grouping). */
enum {
    HG_ADJ_FUSE_EROSION_FLUX = 1,      /* hydro_erosion + hydro_flux */
    HG_ADJ_FUSE_SMOOTH_THERMAL1 = 2,   /* smoothing + thermal layer 1 (with both flags) */
    HG_ADJ_FUSE_FAR_THERMAL0 = 4,      /* sediment_transport's scatter of back-trace taps beyond one
                                          cell + thermal layer 0 (with the thermal flag) */
    HG_ADJ_FUSE_TRANSPORT_BESIDE = 8   /* sediment_transport's gather in the launch of bit 2's kernel
                                          (with bits 2 and 4 and both flags) */
};
int hg_set_adjoint_fusion(hg_ctx* ctx, int on);
int hg_get_adjoint_fusion(const hg_ctx* ctx, int* on);
int hg_adjoint_rollout(hg_ctx* ctx, uint32_t nsteps, uint32_t k_ckpt, uint32_t flags,
                       const float* dL_dH_final, float* dL_dH0, double* dL_dcoef);

/* ---- tile variants and autotuning -------------------------------------- */

/* The two fused kernels of a GRID step are compiled for several tile shapes
 * (tile width x height, halo, threads per workgroup).  The autotuner times
 * each on the context's own grid and keeps the fastest; the choice is cached
 * per (grid size, device architecture) for contexts created later in the
 * process.  These have no counterpart in the reference (its workgroup is a
 * fixed 8x8, glsl/bindings.glsl:9-10). */
enum { HG_KERNEL_HYDRO = 0, HG_KERNEL_THERMAL = 1, HG_KERNEL_SWEEP = 2 };
int hg_tile_variant_count(int kernel);
/* For HG_KERNEL_SWEEP (the one-kernel step, hg_set_sweep): tile_x = strip
 * width in columns, tile_y = rows of the segment a workgroup sweeps, halo =
 * rows per iteration of the sweep. */
int hg_tile_variant_info(int kernel, int variant, int* tile_x, int* tile_y, int* halo, int* threads);
/* -1 keeps the current choice. */
int hg_set_tiles(hg_ctx* ctx, int hydro_variant, int thermal_variant);
int hg_get_tiles(const hg_ctx* ctx, int* hydro_variant, int* thermal_variant);
/* Tile order of the fused kernels' workgroups, which the hardware deals to
 * the 8 XCDs round robin.  0: row-major tile order.  1: each XCD gets a
 * contiguous run of tiles.  2: each XCD gets a column strip of tiles (when
 * the tile columns divide by 8) and walks it row by row, so the halos a tile
 * shares with its horizontal and vertical neighbours are read through one
 * L2 (the default).  Returns HG_ERR_INVALID for other modes. */
int hg_set_xcd_order(hg_ctx* ctx, int mode);
/* One-kernel steps.  With a HG_KERNEL_SWEEP variant, hg_step_grid runs
 * each fused step as one kernel: each wave (variants 0-4) or workgroup (the
 * rest) sweeps a strip of columns row by row and runs every pass of the step
 * a few rows behind the one it depends on, so the step reads and writes its
 * 9 state planes once.  -1 (the default, kept by hg_autotune unless a
 * one-kernel variant times faster) runs the hydro kernel and then the
 * thermal kernel.  Both give the per-pass result bit for bit; steps with
 * HG_STEP_KEEP_VELOCITY always use the two kernels. */
int hg_set_sweep(hg_ctx* ctx, int variant);
int hg_get_sweep(const hg_ctx* ctx, int* variant);
/* Times `steps` fused steps per variant (state and step counter are restored
 * afterwards) and selects the fastest of each kernel, then the one-kernel
 * variant if one is faster than the two kernels together (else -1, see
 * hg_set_sweep). */
int hg_autotune(hg_ctx* ctx, uint32_t steps, uint32_t flags, int* hydro_best, int* thermal_best);

/* ---- batches of terrains -------------------------------------------------- */

/* A batch holds `count` square terrains of uneven sides (multiples of 8), each
 * with its own erosion coefficients, all advanced in lockstep.  A batched step
 * is one persistent launch of each fused kernel over the tiles of every
 * terrain.  The rain counter carries the terrain index (ctr word 3), so each
 * terrain of a batch rains independently.  No counterpart in the reference,
 * which erodes one terrain per context. */
typedef struct hg_batch hg_batch;
int hg_batch_create(int device, int count, const uint32_t* sizes, hg_batch** out);
int hg_batch_destroy(hg_batch* b);
const char* hg_batch_last_error(const hg_batch* b);
int hg_batch_count(const hg_batch* b, int* count);
/* terrain = -1 sets every terrain. */
int hg_batch_set_erosion(hg_batch* b, int terrain, const hg_erosion_settings* s);
int hg_batch_set_rain(hg_batch* b, const hg_rain_settings* s, int mode, uint64_t seed);
int hg_batch_init_fbm(hg_batch* b, int terrain, const hg_fbm_params* p);
int hg_batch_set_tiles(hg_batch* b, int hydro_variant, int thermal_variant);
/* Times `steps` batched steps per tile variant (state, buffers and step
 * counter are restored afterwards) and selects the fastest of each kernel;
 * drops a captured graph. */
int hg_batch_autotune(hg_batch* b, uint32_t steps, uint32_t flags, int* hydro_best, int* thermal_best);
/* 0 (default): each batched launch has one workgroup per tile of the work-list;
 * 1: a persistent grid of resident workgroups strides over the work-list. */
int hg_batch_set_persistent(hg_batch* b, int persistent);
/* 1 (default): the work-list is cut into 8 contiguous chunks interleaved so
 * that each XCD works through one chunk (its L2 then holds the halo rows its
 * tiles share); 0: natural (terrain, row, column) order. */
int hg_batch_set_xcd_order(hg_batch* b, int on);
/* nsteps steps with persistent launches.  HG_STEP_PER_TERRAIN instead launches
 * the single-terrain fused kernels terrain by terrain; with HG_STEP_PER_PASS as
 * well, one kernel per reference pass per terrain (the reference's dispatch
 * pattern).  All three give identical results. */
int hg_batch_step(hg_batch* b, uint32_t nsteps, uint32_t flags);
/* Records nsteps (even) persistent steps into a hipGraph; hg_batch_run
 * replays it, continuing the step counter (kept in device memory).  With
 * HG_STEP_PER_TERRAIN in flags the graph instead holds one launch of each
 * fused kernel per terrain and step, over that terrain's tiles only (the
 * pattern of eroding the terrains one by one, for comparison; same results).
 * The
 * graph's launches name the plane buffers current at capture time: after an
 * odd number of eager hg_batch_step steps (which swap them) hg_batch_run
 * returns HG_ERR_STATE until another step restores them or a new capture
 * replaces the graph. */
int hg_batch_capture(hg_batch* b, uint32_t nsteps, uint32_t flags);
int hg_batch_run(hg_batch* b);
int hg_batch_sync(hg_batch* b);
int hg_batch_get_step(const hg_batch* b, uint64_t* step);
/* State planes (HG_P_ROCK .. HG_P_SED_DIRT) of one terrain, n*n floats. */
int hg_batch_read_plane(hg_batch* b, int terrain, int plane, float* dst, size_t n_floats);
int hg_batch_write_plane(hg_batch* b, int terrain, int plane, const float* src, size_t n_floats);

#ifdef __cplusplus
}
#endif
#endif /* HG_H */
