// hg_kernels.hpp — kernel launch interface shared by the runtime and the
// kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "hg_cell.hpp"
#include "../../include/hg.h"

namespace hgk {

// Byte distance between consecutive FP32 planes of one allocation: the plane
// rounded up to 256 B plus a 16,640-B stagger.  Planes of power-of-two size
// placed back to back all start on the same HBM channel and bank, and a kernel
// streaming 9 planes in and 9 out then reaches 5.0 TB/s at 4096^2; with the
// stagger it reaches 5.5 TB/s.
constexpr size_t kPlaneStagger = 16640;
inline size_t plane_stride_bytes(size_t cells) { return (cells * sizeof(float) + 255) / 256 * 256 + kPlaneStagger; }

// Whether the fused kernels write with non-temporal stores, given the bytes
// of state a step reads and writes (both buffers of the 9 state planes).
// Measured per step: 1024^2 (72 MB, within MI355X's 256 MB Infinity Cache)
// 0.0350 ms with ordinary stores, 0.0386 ms non-temporal; 2048^2 (288 MB)
// 0.1193 vs 0.1158 ms; 4096^2 0.435 vs 0.425 ms.
constexpr size_t kLastLevelCacheBytes = size_t(256) << 20;
inline int stream_stores(size_t step_state_bytes) { return step_state_bytes > kLastLevelCacheBytes ? 1 : 0; }

// Device plane pointers, indexed by HG_P_*.
struct Planes {
    float* p[HG_NUM_PLANES];
};

// Rain increment of cell (x, y) at `step` (config rain modes).
__device__ __forceinline__ float rain_incr(int x, int y, uint32_t step, uint32_t terrain,
                                           const KParams& k) {
    if (k.rain_mode == HG_RAIN_CONSTANT) return k.rain_amount;
    if (k.rain_mode == HG_RAIN_UNIFORM) {
        uint32_t o[4];
        philox4x32_10(uint32_t(x) >> 2, uint32_t(y), step, terrain, k.seed_lo, k.seed_hi, o);
        const uint32_t w = (x & 3) == 0 ? o[0] : (x & 3) == 1 ? o[1] : (x & 3) == 2 ? o[2] : o[3];
        return k.rain_amount * u24(w);
    }
    return 0.0f;
}

// ---- launchers (all enqueue on `s`; return hipError_t of the launch) -------

// per-pass kernels (hg_pass.hip): one per reference dispatch
hipError_t launch_pass(int pass, const Planes& in, const Planes& out, int nx, int ny, uint32_t step,
                       uint32_t terrain, const KParams& k, hipStream_t s);

// fused kernels (hg_fused.hip)
struct HydroIO {
    const float *r, *d, *b, *fl, *fr, *ft, *fb, *s0, *s1;   // step input
    float *ro, *dout, *bo, *flo, *fro, *fto, *fbo, *s0o, *s1o; // step output
    float *vx, *vy, *vz;                                     // velocity out (null = not kept)
    // optional tape of intermediates for the adjoint (all set or all null):
    // water after rain, water before evaporation, eroded sediment before transport
    float *tb1 = nullptr, *tb2 = nullptr, *ts0 = nullptr, *ts1 = nullptr;
};
struct ThermIO {
    const float *r, *d;
    float *ro, *dout;
    // optional tape: rock after thermal layer 0, dirt after layer 1 (before smoothing)
    float *tr5 = nullptr, *td6 = nullptr;
};
// ---------------------------------------------------------------------------
// Eroded sediment of one cell, recomputed from the step's input planes
// (rain -> flux -> erosion at q).  Used for back-trace taps outside the halo.
// ---------------------------------------------------------------------------
__device__ inline void eroded_sediment_global(const HydroIO& io, int qx, int qy, int nx, int ny,
                                       uint32_t step, uint32_t terrain, bool rain, const KParams& k,
                                       float& s0o, float& s1o) {
    const float* R = io.r;
    const float* D = io.d;
    const float* B = io.b;
    auto water = [&](int x, int y) -> float {
        float b = B[size_t(y) * size_t(nx) + size_t(x)];
        if (rain) b = b + rain_incr(x, y, step, terrain, k);
        return b;
    };
    auto total = [&](int x, int y) -> float {
        if (x < 0 || y < 0 || x >= nx || y >= ny) return HG_OOB_HEIGHT;
        const size_t j = size_t(y) * size_t(nx) + size_t(x);
        return (R[j] + D[j]) + water(x, y);
    };
    auto rd = [&](int x, int y) -> float {
        if (x < 0 || y < 0 || x >= nx || y >= ny) return 0.0f;
        const size_t j = size_t(y) * size_t(nx) + size_t(x);
        return R[j] + D[j];
    };
    const size_t i = size_t(qy) * size_t(nx) + size_t(qx);
    float f[4] = {io.fl[i], io.fr[i], io.ft[i], io.fb[i]};
    const float b = water(qx, qy);
    const float w = (R[i] + D[i]) + b;
    const float wn[4] = {total(qx - 1, qy), total(qx + 1, qy), total(qx, qy + 1), total(qx, qy - 1)};
    const float fin[4] = {qx > 0 ? io.fr[i - 1] : 0.0f, qx < nx - 1 ? io.fl[i + 1] : 0.0f,
                          qy < ny - 1 ? io.fb[i + nx] : 0.0f, qy > 0 ? io.ft[i - nx] : 0.0f};
    const unsigned wall = (qx == 0 ? 1u : 0u) | (qx == nx - 1 ? 2u : 0u) | (qy == 0 ? 4u : 0u) |
                          (qy == ny - 1 ? 8u : 0u);
    float vx, vy, vz;
    (void)flux_cell(f, b, w, wn, fin, wall, k, vx, vy, vz);
    float r = R[i], d = D[i];
    s0o = io.s0[i];
    s1o = io.s1[i];
    erosion_cell(r, d, s0o, s1o, vx, vy, vz, rd(qx - 1, qy), rd(qx + 1, qy), rd(qx, qy + 1),
                 rd(qx, qy - 1), k);
}

int hydro_variant_count();
int thermal_variant_count();
bool hydro_variant_info(int v, int* tx, int* ty, int* halo, int* nt);
bool thermal_variant_info(int v, int* tx, int* ty, int* halo, int* nt);
hipError_t launch_fused_hydro(int variant, const HydroIO& io, int nx, int ny, uint32_t step,
                              uint32_t terrain, bool rain, int xcd, const KParams& k, hipStream_t s);
hipError_t launch_fused_thermal(int variant, const ThermIO& io, int nx, int ny, bool thermal,
                                bool smooth, int xcd, const KParams& k, hipStream_t s);

// the whole step in one kernel (hg_sweep.hip): strips of columns swept row
// by row; rain / thermal / smoothing as the step's flags say
int sweep_variant_count();
bool sweep_variant_info(int v, int* strip_w, int* rows, int* threads, int* seg_rows);
hipError_t launch_step_sweep(int variant, const HydroIO& io, int n, uint32_t step, uint32_t terrain, bool rain,
                             bool thermal, bool smooth, const KParams& k, hipStream_t s);

// batched persistent kernels (hg_fused.hip)
struct BatchTerrain {
    float* plane[HG_NUM_STATE_PLANES][2];   // both buffers of each state plane
    int n;
    KParams k;
};
hipError_t launch_batch_hydro(int variant, const BatchTerrain* terr, const uint32_t* tiles, int ntiles,
                              int grid, uint32_t cur, const uint32_t* step_ptr, uint32_t step_off,
                              bool rain, bool nts, hipStream_t s);
// *step_ptr += n (a kernel, so it can sit in a captured graph)
hipError_t launch_step_advance(uint32_t* step_ptr, uint32_t n, hipStream_t s);
hipError_t launch_batch_thermal(int variant, const BatchTerrain* terr, const uint32_t* tiles, int ntiles,
                                int grid, uint32_t cur, bool thermal, bool smooth, bool nts, hipStream_t s);
int batch_hydro_blocks_per_cu(int variant);
int batch_thermal_blocks_per_cu(int variant);

// droplet erosion (hg_particle.hip)
#define HG_BRUSH_MAX 128
struct DropletParams {
    uint32_t count, max_lifetime;
    float inertia, capacity, deposition, erosion, evaporation, gravity, min_slope;
    float init_water, init_speed;
    uint32_t seed_lo, seed_hi;
};
struct Brush {
    int r, count;
    int8_t dx[HG_BRUSH_MAX], dy[HG_BRUSH_MAX];
    float w[HG_BRUSH_MAX];   // normalised to sum 1 over the full brush
};
Brush make_brush(int r);
// order: null, or the droplet id each thread slot runs (a permutation of
// [0, count), from launch_droplet_order).
hipError_t launch_droplets(float* rock, float* dirt, int n, const DropletParams& p, const Brush& brush,
                           uint32_t iteration, const uint32_t* order, double* exported,
                           unsigned long long* steps, hipStream_t s);
// Droplet ordering: a counting sort of an iteration's droplets by the tile of
// their start cell (tiles of droplet_tile_side(n) cells, row-major), so the
// 64 droplets of a wave start close together and share cache lines.
// tiles: droplet_tile_count(n) words of scratch; order: count words.
int droplet_tile_side(int n);
int droplet_tile_count(int n);
hipError_t launch_droplet_order(int n, const DropletParams& p, uint32_t iteration, uint32_t* tiles,
                                uint32_t* order, hipStream_t s);
// Per-workgroup sums of one droplet launch (256 droplets per workgroup).
struct DropletPartial {
    double exported;
    unsigned long long steps;
};
// The droplet kernel writing one DropletPartial per workgroup instead of
// adding to counters, and the fixed-order sum of those partials.
hipError_t launch_droplets_partials(float* rock, float* dirt, int n, const DropletParams& p, const Brush& brush,
                                    uint32_t iteration, DropletPartial* partials, hipStream_t s);
hipError_t launch_droplet_totals(const DropletPartial* partials, uint32_t nblocks, double* exported,
                                 uint64_t* steps, bool accumulate, hipStream_t s);

// the reference's particle mode (hg_refparticle.hip)
struct RefParticleParams {
    float Kc, Kconv, Ks[2], Kd[2], Ke, G, dt, init_volume, friction, inertia, min_volume, min_velocity;
    uint32_t ttl, particle_count;
    int dims_x, dims_y;       // hmap_dims of the spawn range and border test
};
hipError_t launch_ref_particle_move(hg_ref_particle* parts, uint32_t count, const float* rock, const float* dirt,
                                    const float* water, const float* mx, const float* my, int n,
                                    const RefParticleParams& k, float time, int should_rain, hipStream_t s);
hipError_t launch_ref_particle_erode(hg_ref_particle* parts, uint32_t count, float* rock, float* dirt, float* water,
                                     float* mz, float* mw, int n, const RefParticleParams& k,
                                     unsigned long long* giveups, hipStream_t s);
hipError_t launch_ref_particle_smooth_extra(float* water, float* mx, float* my, float* mz, float* mw, int n,
                                            uint32_t particle_count, hipStream_t s);

// heightmap init (hg_init.hip)
void fbm_permutation(uint64_t seed, uint8_t perm[256]);
hipError_t launch_fbm(float* out, int nx, int ny, const uint8_t* perm_host, int octaves,
                      float lacunarity, float gain, float base_freq, hipStream_t s);
hipError_t launch_minmax(const float* v, size_t n, unsigned* mm_dev, hipStream_t s);
hipError_t launch_rescale(float* v, size_t n, const unsigned* mm_dev, float lo, float hi,
                          hipStream_t s);
hipError_t launch_fill(float* p, size_t n, float value, hipStream_t s);
// out (4n floats, 16-byte aligned) = RGBA texels of the 4 planes (null: 0);
// `total`: the fourth component is (p0 + p1) + p2
hipError_t launch_interleave_rgba(float* out, const float* const planes[4], size_t n, bool total, hipStream_t s);
// planes[k] (when not null) = component k of the texels in `in` (4n floats)
hipError_t launch_deinterleave_rgba(const float* in, float* const planes[4], size_t n, hipStream_t s);
hipError_t launch_gen_heightmap(float* rock, float* dirt, int n, const hg_map_settings& m, hipStream_t s);

}  // namespace hgk
