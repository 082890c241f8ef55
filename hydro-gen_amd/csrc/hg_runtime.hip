// hg_runtime.hip — host runtime behind the C ABI of include/hg.h.
//
// Replaces the reference's GL resource layer and erosion dispatch
// (src/state.cpp:3-55, src/erosion.cpp:21-200): a context owns the SoA state
// planes of one n x n terrain on one HIP device, a stream, the settings and
// the step counter.  Every pass reads the current buffer of each plane and
// writes the alternate buffer; the planes a pass produced are then swapped,
// which is the Tex_pair::swap of the reference (src/shaderprogram.cpp:51-72)
// per plane.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "hg_context.hpp"

using namespace hgctx;

namespace {

thread_local std::string g_create_error;

// Autotuner results per (grid size, device architecture): {hydro, thermal}.
std::mutex g_tune_mu;
struct TunedTiles {
    int hydro, thermal, sweep;
};
std::map<std::pair<int, std::string>, TunedTiles> g_tune_cache;
constexpr int DEFAULT_HYDRO_VARIANT = 0;
constexpr int DEFAULT_THERMAL_VARIANT = 1;
constexpr int DEFAULT_SWEEP_VARIANT = -1;

}  // namespace



namespace hgctx {

int fail(hg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg; else g_create_error = msg;
    return code;
}

int hip_fail(hg_ctx* c, hipError_t e, const char* what) {
    return fail(c, HG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}


size_t plane_n(const hg_ctx* c) { return size_t(c->n) * size_t(c->n); }

Planes planes_of(float* const* p) {
    Planes P;
    for (int i = 0; i < HG_NUM_PLANES; ++i) P.p[i] = p[i];
    return P;
}

hg_erosion_settings reference_grid_defaults() {
    // src/state.cpp:81-92
    hg_erosion_settings s{};
    s.particle_count = 0;
    s.Kc = 0.2f;
    s.Kalpha[0] = 1.3f; s.Kalpha[1] = 0.6f;
    s.Kconv = 0.001f;
    s.Ks[0] = 0.03f; s.Ks[1] = 0.09f;
    s.Kd[0] = 0.01f; s.Kd[1] = 0.03f;
    s.Ke = 0.03f;
    s.energy_kept = 1.0f;
    s.Kspeed[0] = 0.5f; s.Kspeed[1] = 2.0f;
    s.G = 1.0f;
    s.d_t = 0.001f;
    return s;
}

hg_rain_settings reference_rain_defaults() {
    // src/settings.hpp:13-25
    hg_rain_settings r{};
    r.amount = 0.01f;
    r.mountain_thresh = 0.55f;
    r.mountain_multip = 0.05f;
    r.period = 512;
    r.drops = 0.02f;
    return r;
}

void update_kparams(hg_ctx* c) {
    const hg_erosion_settings& s = c->es;
    KParams& k = c->kp;
    k.Kc = s.Kc; k.Kconv = s.Kconv; k.Ke = s.Ke; k.E = s.energy_kept; k.G = s.G; k.dt = s.d_t;
    for (int l = 0; l < 2; ++l) {
        k.Ks[l] = s.Ks[l]; k.Kd[l] = s.Kd[l]; k.Kalpha[l] = s.Kalpha[l]; k.Kspeed[l] = s.Kspeed[l];
        k.tan_alpha[l] = std::tan(s.Kalpha[l]);
        k.thr_diag[l] = std::sqrt(2.0f) * k.tan_alpha[l];
    }
    const float m = s.Kspeed[1] * s.d_t;                      // glsl/smoothing.glsl:75
    k.smooth_m = m < 0.0f ? 0.0f : (m > 1.0f ? 1.0f : m);
    k.evap = 1.0f - s.Ke * s.d_t;                             // sediment_transport.glsl:75
    k.rain_amount = c->rs.amount;
    k.rain_mode = c->rain_mode;
    k.rr_thresh = c->rs.mountain_thresh;
    k.rr_multip = c->rs.mountain_multip;
    k.rr_drops = c->rs.drops;
    k.rr_time = c->rain_time;
    k.rr_max_h = c->map_max_height;
    k.seed_lo = uint32_t(c->seed);
    k.seed_hi = uint32_t(c->seed >> 32);
    k.store_nt = hgk::stream_stores(2 * HG_NUM_STATE_PLANES * hgk::plane_stride_bytes(size_t(c->n) * size_t(c->n)));
}

int stat_index(hg_ctx* c, const char* name) {
    for (size_t i = 0; i < c->stats.size(); ++i)
        if (c->stats[i].name == name) return int(i);
    c->stats.push_back(KernelStat{name, 0, 0.0});
    return int(c->stats.size()) - 1;
}

hipEvent_t take_event(hg_ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Resolve all pending kernel timings (synchronises on their end events).
int resolve_timing(hg_ctx* c) {
    for (auto& p : c->pending) {
        HG_HIP(c, hipEventSynchronize(p.end));
        float ms = 0.0f;
        HG_HIP(c, hipEventElapsedTime(&ms, p.start, p.end));
        c->stats[p.stat].launches += 1;
        c->stats[p.stat].total_ms += double(ms);
        c->event_pool.push_back(p.start);
        c->event_pool.push_back(p.end);
    }
    c->pending.clear();
    return HG_OK;
}

void swap_planes(hg_ctx* c, std::initializer_list<int> ps) {
    for (int p : ps) std::swap(c->cur[p], c->alt[p]);
}

const char* pass_name(int pass) {
    static const char* names[HG_NUM_PASSES] = {
        "pass_rain", "pass_hydro_flux", "pass_hydro_erosion", "pass_sediment_transport",
        "pass_thermal_flux_0", "pass_thermal_transport_0", "pass_thermal_flux_1",
        "pass_thermal_transport_1", "pass_smoothing"};
    return names[pass];
}

int run_pass(hg_ctx* c, int pass) {
    const Planes in = planes_of(c->cur), out = planes_of(c->alt);
    const uint32_t step = uint32_t(c->step);
    int rc = timed_launch(c, pass_name(pass), [&] {
        return hgk::launch_pass(pass, in, out, c->n, c->n, step, 0u, c->kp, c->stream);
    });
    if (rc != HG_OK) return rc;
    switch (pass) {
        case HG_PASS_RAIN: swap_planes(c, {HG_P_WATER}); break;
        case HG_PASS_FLUX:
            swap_planes(c, {HG_P_WATER, HG_P_FLUX_L, HG_P_FLUX_R, HG_P_FLUX_T, HG_P_FLUX_B, HG_P_VEL_U,
                            HG_P_VEL_V, HG_P_VEL_DEPTH});
            break;
        case HG_PASS_EROSION: swap_planes(c, {HG_P_ROCK, HG_P_DIRT, HG_P_SED_ROCK, HG_P_SED_DIRT}); break;
        case HG_PASS_TRANSPORT: swap_planes(c, {HG_P_SED_ROCK, HG_P_SED_DIRT, HG_P_WATER}); break;
        case HG_PASS_THERMAL_FLUX_0:
        case HG_PASS_THERMAL_FLUX_1:
            swap_planes(c, {HG_P_THERM_L, HG_P_THERM_R, HG_P_THERM_T, HG_P_THERM_B, HG_P_THERM_LT,
                            HG_P_THERM_RT, HG_P_THERM_LB, HG_P_THERM_RB});
            break;
        case HG_PASS_THERMAL_TRANSPORT_0: swap_planes(c, {HG_P_ROCK}); break;
        case HG_PASS_THERMAL_TRANSPORT_1: swap_planes(c, {HG_P_DIRT}); break;
        case HG_PASS_SMOOTH: swap_planes(c, {HG_P_ROCK, HG_P_DIRT}); break;
        default: break;
    }
    return HG_OK;
}

int fused_step(hg_ctx* c, bool rain, bool thermal, bool smooth, bool keep_velocity) {
    hgk::HydroIO io;
    io.r = c->cur[HG_P_ROCK]; io.d = c->cur[HG_P_DIRT]; io.b = c->cur[HG_P_WATER];
    io.fl = c->cur[HG_P_FLUX_L]; io.fr = c->cur[HG_P_FLUX_R]; io.ft = c->cur[HG_P_FLUX_T];
    io.fb = c->cur[HG_P_FLUX_B]; io.s0 = c->cur[HG_P_SED_ROCK]; io.s1 = c->cur[HG_P_SED_DIRT];
    io.ro = c->alt[HG_P_ROCK]; io.dout = c->alt[HG_P_DIRT]; io.bo = c->alt[HG_P_WATER];
    io.flo = c->alt[HG_P_FLUX_L]; io.fro = c->alt[HG_P_FLUX_R]; io.fto = c->alt[HG_P_FLUX_T];
    io.fbo = c->alt[HG_P_FLUX_B]; io.s0o = c->alt[HG_P_SED_ROCK]; io.s1o = c->alt[HG_P_SED_DIRT];
    io.vx = keep_velocity ? c->alt[HG_P_VEL_U] : nullptr;
    io.vy = keep_velocity ? c->alt[HG_P_VEL_V] : nullptr;
    io.vz = keep_velocity ? c->alt[HG_P_VEL_DEPTH] : nullptr;
    const uint32_t step = uint32_t(c->step);
    if (c->sweep_variant >= 0 && !keep_velocity) {
        // the whole step in one kernel; rock and dirt come out of its last pass
        const int rc = timed_launch(c, "step_sweep", [&] {
            return hgk::launch_step_sweep(c->sweep_variant, io, c->n, step, 0u, rain, thermal, smooth, c->kp,
                                          c->stream);
        });
        if (rc != HG_OK) return rc;
        swap_planes(c, {HG_P_ROCK, HG_P_DIRT, HG_P_WATER, HG_P_FLUX_L, HG_P_FLUX_R, HG_P_FLUX_T, HG_P_FLUX_B,
                        HG_P_SED_ROCK, HG_P_SED_DIRT});
        return HG_OK;
    }
    int rc = timed_launch(c, "fused_hydro", [&] {
        return hgk::launch_fused_hydro(c->hydro_variant, io, c->n, c->n, step, 0u, rain, c->xcd_order, c->kp,
                                       c->stream);
    });
    if (rc != HG_OK) return rc;
    swap_planes(c, {HG_P_ROCK, HG_P_DIRT, HG_P_WATER, HG_P_FLUX_L, HG_P_FLUX_R, HG_P_FLUX_T, HG_P_FLUX_B,
                    HG_P_SED_ROCK, HG_P_SED_DIRT});
    if (keep_velocity) swap_planes(c, {HG_P_VEL_U, HG_P_VEL_V, HG_P_VEL_DEPTH});
    if (!thermal && !smooth) return HG_OK;
    hgk::ThermIO tio;
    tio.r = c->cur[HG_P_ROCK]; tio.d = c->cur[HG_P_DIRT];
    tio.ro = c->alt[HG_P_ROCK]; tio.dout = c->alt[HG_P_DIRT];
    rc = timed_launch(c, "fused_thermal", [&] {
        return hgk::launch_fused_thermal(c->thermal_variant, tio, c->n, c->n, thermal, smooth, c->xcd_order, c->kp,
                                         c->stream);
    });
    if (rc != HG_OK) return rc;
    swap_planes(c, {HG_P_ROCK, HG_P_DIRT});
    return HG_OK;
}

bool finite_settings(const hg_erosion_settings* s) {
    const float v[] = {s->Kc, s->Kalpha[0], s->Kalpha[1], s->Kconv, s->Ks[0], s->Ks[1], s->Kd[0],
                       s->Kd[1], s->Ke, s->energy_kept, s->Kspeed[0], s->Kspeed[1], s->G, s->d_t,
                       s->density, s->init_volume, s->friction, s->inertia, s->min_volume,
                       s->min_velocity};
    for (float x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

// Device-to-host copy through the context's two pinned chunks: the copy of
// chunk i+1 runs while chunk i is copied out to `dst`.
constexpr size_t kPinChunk = size_t(16) << 20;
int read_pinned(hg_ctx* c, const float* src, float* dst, size_t n_floats) {
    for (int b = 0; b < 2; ++b) {            // each piece stored only once it exists
        if (!c->pin[b]) {
            void* p = nullptr;
            HG_HIP(c, hipHostMalloc(&p, kPinChunk, hipHostMallocDefault));
            c->pin[b] = p;
        }
        if (!c->pin_ev[b]) {
            hipEvent_t e = nullptr;
            HG_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
            c->pin_ev[b] = e;
        }
    }
    const size_t bytes = n_floats * sizeof(float);
    const char* s = reinterpret_cast<const char*>(src);
    char* d = reinterpret_cast<char*>(dst);
    auto issue = [&](size_t off, int b) -> int {
        const size_t len = std::min(kPinChunk, bytes - off);
        HG_HIP(c, hipMemcpyAsync(c->pin[b], s + off, len, hipMemcpyDeviceToHost, c->stream));
        HG_HIP(c, hipEventRecord(c->pin_ev[b], c->stream));
        return HG_OK;
    };
    if (bytes == 0) return HG_OK;
    int rc = issue(0, 0);
    if (rc != HG_OK) return rc;
    for (size_t off = 0, k = 0; off < bytes; off += kPinChunk, ++k) {
        const int b = int(k & 1);
        if (off + kPinChunk < bytes) {           // the other chunk was copied out last round
            rc = issue(off + kPinChunk, 1 - b);
            if (rc != HG_OK) return rc;
        }
        HG_HIP(c, hipEventSynchronize(c->pin_ev[b]));
        std::memcpy(d + off, c->pin[b], std::min(kPinChunk, bytes - off));
    }
    return HG_OK;
}

}  // namespace

extern "C" {

int hg_create(int device, uint32_t n, uint32_t particle_count, hg_ctx** out) {
    if (!out) return fail(nullptr, HG_ERR_INVALID, "hg_create: out is NULL");
    *out = nullptr;
    if (n < 8 || n > 32768 || (n % 8) != 0)
        return fail(nullptr, HG_ERR_INVALID,
                    "hg_create: n must be a multiple of 8 in [8, 32768] (the reference dispatches n/8 "
                    "workgroups of 8x8)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, HG_ERR_HIP, "hg_create: no HIP device");
    if (device < 0 || device >= ndev) return fail(nullptr, HG_ERR_INVALID, "hg_create: bad device");
    (void)particle_count;
    DeviceGuard guard(device);
    hg_ctx* c = new hg_ctx();
    c->device = device;
    c->n = int(n);
    const size_t pn = size_t(n) * n;
    const size_t pitch = hgk::plane_stride_bytes(pn);
    const size_t bytes = pitch * 2 * HG_NUM_PLANES + 256;   // + init reduction, droplet counters
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(nullptr, HG_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    e = hipMalloc(&c->block, bytes);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(nullptr, HG_ERR_OOM, std::string("hipMalloc state planes: ") + hipGetErrorString(e));
    }
    char* base = static_cast<char*>(c->block);
    for (int p = 0; p < HG_NUM_PLANES; ++p) {
        c->cur[p] = reinterpret_cast<float*>(base + pitch * (2 * p));
        c->alt[p] = reinterpret_cast<float*>(base + pitch * (2 * p + 1));
    }
    c->minmax = reinterpret_cast<unsigned*>(base + pitch * 2 * HG_NUM_PLANES);
    c->dev_exported = reinterpret_cast<double*>(base + pitch * 2 * HG_NUM_PLANES + 64);
    c->dev_dsteps = reinterpret_cast<unsigned long long*>(base + pitch * 2 * HG_NUM_PLANES + 128);
    c->dev_giveups = reinterpret_cast<unsigned long long*>(base + pitch * 2 * HG_NUM_PLANES + 192);
    e = hipMemsetAsync(c->block, 0, bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipFree(c->block);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(nullptr, HG_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    }
    c->es = reference_grid_defaults();
    c->rs = reference_rain_defaults();
    update_kparams(c);
    // droplet defaults: the particle configuration's model parameters
    c->ds.count = particle_count;
    c->ds.max_lifetime = 64;
    c->ds.inertia = 0.05f; c->ds.capacity = 4.0f; c->ds.deposition = 0.3f; c->ds.erosion = 0.3f;
    c->ds.evaporation = 0.01f; c->ds.gravity = 4.0f; c->ds.min_slope = 0.01f; c->ds.brush_radius = 3;
    c->ds.init_water = 1.0f; c->ds.init_speed = 1.0f; c->ds.seed = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->arch = prop.gcnArchName;
    c->hydro_variant = DEFAULT_HYDRO_VARIANT;
    c->thermal_variant = DEFAULT_THERMAL_VARIANT;
    c->sweep_variant = DEFAULT_SWEEP_VARIANT;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find({c->n, c->arch});
        if (it != g_tune_cache.end()) {
            c->hydro_variant = it->second.hydro;
            c->thermal_variant = it->second.thermal;
            c->sweep_variant = it->second.sweep;
        }
    }
    *out = c;
    return HG_OK;
}

int hg_destroy(hg_ctx* c) {
    if (!c) return HG_OK;
    DeviceGuard guard(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->pending) { (void)hipEventDestroy(p.start); (void)hipEventDestroy(p.end); }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    (void)hipFree(c->block);
    if (c->refp) (void)hipFree(c->refp);
    if (c->adj_stream) {
        (void)hipStreamSynchronize(c->adj_stream);
        (void)hipStreamDestroy(c->adj_stream);
    }
    for (hipEvent_t e : c->adj_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->adj_block) (void)hipFree(c->adj_block);
    if (c->rgba_dev) (void)hipFree(c->rgba_dev);
    if (c->drop_order) (void)hipFree(c->drop_order);
    for (int b = 0; b < 2; ++b) {
        if (c->pin[b]) (void)hipHostFree(c->pin[b]);
        if (c->pin_ev[b]) (void)hipEventDestroy(c->pin_ev[b]);
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
    return HG_OK;
}

const char* hg_last_error(const hg_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int hg_size(const hg_ctx* c, uint32_t* n) {
    if (!c || !n) return HG_ERR_INVALID;
    *n = uint32_t(c->n);
    return HG_OK;
}

int hg_set_erosion(hg_ctx* c, const hg_erosion_settings* s) {
    if (!c) return HG_ERR_INVALID;
    if (!s) return fail(c, HG_ERR_INVALID, "hg_set_erosion: settings is NULL");
    if (!finite_settings(s)) return fail(c, HG_ERR_INVALID, "hg_set_erosion: non-finite setting");
    if (!(s->d_t > 0.0f)) return fail(c, HG_ERR_INVALID, "hg_set_erosion: d_t must be > 0");
    const float nonneg[] = {s->Kc, s->Kconv, s->Ks[0], s->Ks[1], s->Kd[0], s->Kd[1], s->Ke,
                            s->energy_kept, s->Kspeed[0], s->Kspeed[1], s->G};
    for (float x : nonneg)
        if (x < 0.0f) return fail(c, HG_ERR_INVALID, "hg_set_erosion: negative coefficient");
    for (int l = 0; l < 2; ++l)
        if (!(s->Kalpha[l] >= 0.0f && s->Kalpha[l] < 1.5707963f))
            return fail(c, HG_ERR_INVALID, "hg_set_erosion: Kalpha must be in [0, pi/2) radians");
    c->es = *s;
    update_kparams(c);
    return HG_OK;
}

int hg_get_erosion(const hg_ctx* c, hg_erosion_settings* s) {
    if (!c || !s) return HG_ERR_INVALID;
    *s = c->es;
    return HG_OK;
}

int hg_set_rain(hg_ctx* c, const hg_rain_settings* s, int mode, uint64_t seed) {
    if (!c) return HG_ERR_INVALID;
    if (!s) return fail(c, HG_ERR_INVALID, "hg_set_rain: settings is NULL");
    if (mode < HG_RAIN_OFF || mode > HG_RAIN_REFERENCE) return fail(c, HG_ERR_INVALID, "hg_set_rain: bad mode");
    if (!std::isfinite(s->amount) || s->amount < 0.0f)
        return fail(c, HG_ERR_INVALID, "hg_set_rain: amount must be finite and >= 0");
    if (mode == HG_RAIN_REFERENCE &&
        (!std::isfinite(s->drops) || !std::isfinite(s->mountain_multip) || !(s->mountain_thresh >= 0.0f) ||
         !(s->mountain_thresh < 1.0f)))
        return fail(c, HG_ERR_INVALID, "hg_set_rain: reference rain needs finite drops/multip, 0 <= thresh < 1");
    c->rs = *s;
    c->rain_mode = mode;
    c->seed = seed;
    update_kparams(c);
    return HG_OK;
}

int hg_set_rain_time(hg_ctx* c, float time, float max_height) {
    if (!c) return HG_ERR_INVALID;
    if (!std::isfinite(time) || !std::isfinite(max_height) || !(max_height > 0.0f))
        return fail(c, HG_ERR_INVALID, "hg_set_rain_time: need a finite time and max_height > 0");
    c->rain_time = time;
    c->map_max_height = max_height;
    update_kparams(c);
    return HG_OK;
}

static_assert(sizeof(hg_map_settings) == 96, "Map_settings_data is 96 bytes in std140");

int hg_gen_heightmap(hg_ctx* c, const hg_map_settings* m) {
    if (!c) return HG_ERR_INVALID;
    if (!m) return fail(c, HG_ERR_INVALID, "hg_gen_heightmap: settings is NULL");
    const float f[] = {m->max_height, m->max_dirt, m->height_mult, m->seed, m->persistance, m->lacunarity,
                       m->scale, m->uplift_scale, m->domain_warp_scale, m->terrace_scale};
    for (float v : f)
        if (!std::isfinite(v)) return fail(c, HG_ERR_INVALID, "hg_gen_heightmap: non-finite setting");
    if (m->octaves < 0 || m->octaves > 30 || m->domain_warp < 0 || m->domain_warp > 2 ||
        (m->uplift != 0 && m->uplift_scale == 0.0f))
        return fail(c, HG_ERR_INVALID, "hg_gen_heightmap: octaves 0..30, domain_warp 0..2, uplift_scale != 0");
    DeviceGuard guard(c->device);
    const size_t pn = plane_n(c);
    for (int pl = 0; pl < HG_NUM_PLANES; ++pl)
        if (pl != HG_P_ROCK && pl != HG_P_DIRT) HG_HIP(c, hgk::launch_fill(c->cur[pl], pn, 0.0f, c->stream));
    int rc = timed_launch(c, "gen_heightmap", [&] {
        return hgk::launch_gen_heightmap(c->cur[HG_P_ROCK], c->cur[HG_P_DIRT], c->n, *m, c->stream);
    });
    if (rc != HG_OK) return rc;
    c->step = 0;
    return HG_OK;
}

int hg_init_fbm(hg_ctx* c, const hg_fbm_params* p) {
    if (!c) return HG_ERR_INVALID;
    if (!p) return fail(c, HG_ERR_INVALID, "hg_init_fbm: params is NULL");
    if (p->octaves < 1 || p->octaves > 32 || !std::isfinite(p->base_freq) || !std::isfinite(p->gain) ||
        !std::isfinite(p->lacunarity) || !std::isfinite(p->height_lo) || !std::isfinite(p->height_hi) ||
        !(p->height_hi >= p->height_lo))
        return fail(c, HG_ERR_INVALID, "hg_init_fbm: bad parameters");
    DeviceGuard guard(c->device);
    uint8_t perm[256];
    hgk::fbm_permutation(p->seed, perm);
    const size_t pn = plane_n(c);
    for (int pl = 0; pl < HG_NUM_PLANES; ++pl)
        if (pl != HG_P_ROCK) HG_HIP(c, hgk::launch_fill(c->cur[pl], pn, 0.0f, c->stream));
    int rc = timed_launch(c, "init_fbm", [&] {
        return hgk::launch_fbm(c->cur[HG_P_ROCK], c->n, c->n, perm, p->octaves, p->lacunarity, p->gain,
                               p->base_freq, c->stream);
    });
    if (rc != HG_OK) return rc;
    HG_HIP(c, hgk::launch_minmax(c->cur[HG_P_ROCK], pn, c->minmax, c->stream));
    HG_HIP(c, hgk::launch_rescale(c->cur[HG_P_ROCK], pn, c->minmax, p->height_lo, p->height_hi, c->stream));
    HG_HIP(c, hipMemsetAsync(c->dev_exported, 0, 128, c->stream));
    c->step = 0;
    c->particle_iter = 0;
    return HG_OK;
}

int hg_rain(hg_ctx* c) {
    if (!c) return HG_ERR_INVALID;
    DeviceGuard guard(c->device);
    return run_pass(c, HG_PASS_RAIN);
}

int hg_run_pass(hg_ctx* c, int pass) {
    if (!c) return HG_ERR_INVALID;
    if (pass < 0 || pass >= HG_NUM_PASSES) return fail(c, HG_ERR_INVALID, "hg_run_pass: bad pass");
    DeviceGuard guard(c->device);
    return run_pass(c, pass);
}

int hg_step_grid(hg_ctx* c, uint32_t nsteps, uint32_t flags) {
    if (!c) return HG_ERR_INVALID;
    const uint32_t known = HG_STEP_RAIN | HG_STEP_THERMAL | HG_STEP_SMOOTH | HG_STEP_PER_PASS |
                           HG_STEP_KEEP_VELOCITY;
    if (flags & ~known) return fail(c, HG_ERR_INVALID, "hg_step_grid: unknown flag bits");
    DeviceGuard guard(c->device);
    // config rain: every step, inside the step; reference rain: its own pass
    // before step s when (s + 1) % period == 0 (src/main.cpp:313-319)
    const bool rain = (flags & HG_STEP_RAIN) && (c->rain_mode == HG_RAIN_CONSTANT ||
                                                  c->rain_mode == HG_RAIN_UNIFORM);
    const bool ref_rain = (flags & HG_STEP_RAIN) && c->rain_mode == HG_RAIN_REFERENCE && c->rs.period > 0;
    const bool thermal = (flags & HG_STEP_THERMAL) != 0;
    const bool smooth = (flags & HG_STEP_SMOOTH) != 0;
    for (uint32_t s = 0; s < nsteps; ++s) {
        int rc = HG_OK;
        if (ref_rain && (c->step + 1) % uint64_t(c->rs.period) == 0) {
            rc = run_pass(c, HG_PASS_RAIN);
            if (rc != HG_OK) return rc;
        }
        if (flags & HG_STEP_PER_PASS) {
            if (rain) rc = run_pass(c, HG_PASS_RAIN);
            for (int p = HG_PASS_FLUX; rc == HG_OK && p <= HG_PASS_TRANSPORT; ++p) rc = run_pass(c, p);
            if (thermal)
                for (int p = HG_PASS_THERMAL_FLUX_0; rc == HG_OK && p <= HG_PASS_THERMAL_TRANSPORT_1; ++p)
                    rc = run_pass(c, p);
            if (rc == HG_OK && smooth) rc = run_pass(c, HG_PASS_SMOOTH);
        } else {
            rc = fused_step(c, rain, thermal, smooth, (flags & HG_STEP_KEEP_VELOCITY) != 0);
        }
        if (rc != HG_OK) return rc;
        c->step += 1;
    }
    return HG_OK;
}

int hg_set_droplets(hg_ctx* c, const hg_droplet_settings* s) {
    if (!c) return HG_ERR_INVALID;
    if (!s) return fail(c, HG_ERR_INVALID, "hg_set_droplets: settings is NULL");
    const float v[] = {s->inertia, s->capacity, s->deposition, s->erosion, s->evaporation, s->gravity,
                       s->min_slope, s->init_water, s->init_speed};
    for (float x : v)
        if (!std::isfinite(x) || x < 0.0f) return fail(c, HG_ERR_INVALID, "hg_set_droplets: parameters must be finite and >= 0");
    if (s->inertia > 1.0f || s->evaporation > 1.0f)
        return fail(c, HG_ERR_INVALID, "hg_set_droplets: inertia and evaporation must be <= 1");
    if (s->brush_radius < 1 || s->brush_radius > 6)
        return fail(c, HG_ERR_INVALID, "hg_set_droplets: brush_radius must be in [1, 6]");
    if (s->max_lifetime > 100000) return fail(c, HG_ERR_INVALID, "hg_set_droplets: max_lifetime too large");
    c->ds = *s;
    return HG_OK;
}

int hg_set_droplet_order(hg_ctx* c, int mode) {
    if (!c) return HG_ERR_INVALID;
    if (mode < 0 || mode > 1) return fail(c, HG_ERR_INVALID, "hg_set_droplet_order: mode must be 0 or 1");
    c->droplet_order = mode;
    return HG_OK;
}

int hg_get_droplet_order(const hg_ctx* c, int* mode) {
    if (!c || !mode) return HG_ERR_INVALID;
    *mode = c->droplet_order;
    return HG_OK;
}

int hg_set_adjoint_fusion(hg_ctx* c, int on) {
    if (!c) return HG_ERR_INVALID;
    if (on < 0 || on > (HG_ADJ_FUSE_EROSION_FLUX | HG_ADJ_FUSE_SMOOTH_THERMAL1 | HG_ADJ_FUSE_FAR_THERMAL0 |
                        HG_ADJ_FUSE_TRANSPORT_BESIDE))
        return fail(c, HG_ERR_INVALID, "hg_set_adjoint_fusion: a combination of the HG_ADJ_FUSE_* bits");
    c->adj_fuse = on;
    return HG_OK;
}

int hg_get_adjoint_fusion(const hg_ctx* c, int* on) {
    if (!c || !on) return HG_ERR_INVALID;
    *on = c->adj_fuse;
    return HG_OK;
}

int hg_step_particle(hg_ctx* c, uint32_t iterations, uint32_t flags) {
    if (!c) return HG_ERR_INVALID;
    if (flags & ~(HG_STEP_THERMAL | HG_STEP_SMOOTH)) return fail(c, HG_ERR_INVALID, "hg_step_particle: only thermal/smooth flags");
    DeviceGuard guard(c->device);
    hgk::DropletParams p;
    p.count = c->ds.count; p.max_lifetime = c->ds.max_lifetime;
    p.inertia = c->ds.inertia; p.capacity = c->ds.capacity; p.deposition = c->ds.deposition;
    p.erosion = c->ds.erosion; p.evaporation = c->ds.evaporation; p.gravity = c->ds.gravity;
    p.min_slope = c->ds.min_slope; p.init_water = c->ds.init_water; p.init_speed = c->ds.init_speed;
    p.seed_lo = uint32_t(c->ds.seed); p.seed_hi = uint32_t(c->ds.seed >> 32);
    const hgk::Brush brush = hgk::make_brush(c->ds.brush_radius);
    const bool thermal = (flags & HG_STEP_THERMAL) != 0, smooth = (flags & HG_STEP_SMOOTH) != 0;
    const bool ordered = c->droplet_order > 0 && p.count > 0;
    uint32_t* tiles = nullptr;
    if (ordered) {
        // slot -> id for `count` droplets, then the per-tile counts / offsets
        const size_t need = size_t(p.count) + size_t(hgk::droplet_tile_count(c->n));
        if (need > c->drop_cap) {
            HG_HIP(c, hipStreamSynchronize(c->stream));
            if (c->drop_order) (void)hipFree(c->drop_order);
            c->drop_order = nullptr;
            c->drop_cap = 0;
            void* q = nullptr;
            HG_HIP(c, hipMalloc(&q, size_t(need) * sizeof(uint32_t)));
            c->drop_order = static_cast<uint32_t*>(q);
            c->drop_cap = need;
        }
        tiles = c->drop_order + p.count;
    }
    for (uint32_t it = 0; it < iterations; ++it) {
        const uint32_t iter = c->particle_iter;
        int rc = HG_OK;
        if (ordered) {
            rc = timed_launch(c, "droplet_order", [&] {
                return hgk::launch_droplet_order(c->n, p, iter, tiles, c->drop_order, c->stream);
            });
            if (rc != HG_OK) return rc;
        }
        rc = timed_launch(c, "droplets", [&] {
            return hgk::launch_droplets(c->cur[HG_P_ROCK], c->cur[HG_P_DIRT], c->n, p, brush, iter,
                                        ordered ? c->drop_order : nullptr, c->dev_exported, c->dev_dsteps,
                                        c->stream);
        });
        if (rc != HG_OK) return rc;
        if (thermal || smooth) {
            hgk::ThermIO tio;
            tio.r = c->cur[HG_P_ROCK]; tio.d = c->cur[HG_P_DIRT];
            tio.ro = c->alt[HG_P_ROCK]; tio.dout = c->alt[HG_P_DIRT];
            rc = timed_launch(c, "fused_thermal", [&] {
                return hgk::launch_fused_thermal(c->thermal_variant, tio, c->n, c->n, thermal, smooth, c->xcd_order, c->kp,
                                                 c->stream);
            });
            if (rc != HG_OK) return rc;
            swap_planes(c, {HG_P_ROCK, HG_P_DIRT});
        }
        c->particle_iter += 1;
    }
    return HG_OK;
}

// ---- the reference's particle mode ---------------------------------------

int hg_ref_particles_init(hg_ctx* c, uint32_t count) {
    if (!c) return HG_ERR_INVALID;
    if (count == 0 || count > (1u << 26)) return fail(c, HG_ERR_INVALID, "hg_ref_particles_init: count must be in [1, 2^26]");
    DeviceGuard guard(c->device);
    HG_HIP(c, hipStreamSynchronize(c->stream));
    if (c->refp) {
        (void)hipFree(c->refp);
        c->refp = nullptr;
        c->refp_count = 0;
    }
    {
        void* p = nullptr;
        HG_HIP(c, hipMalloc(&p, size_t(count) * sizeof(hg_ref_particle)));
        c->refp = static_cast<hg_ref_particle*>(p);
    }
    c->refp_count = count;
    HG_HIP(c, hipMemsetAsync(c->refp, 0, size_t(count) * sizeof(hg_ref_particle), c->stream));
    const size_t pn = plane_n(c);
    for (int p = HG_P_FLUX_L; p <= HG_P_FLUX_B; ++p)
        HG_HIP(c, hipMemsetAsync(c->cur[p], 0, pn * sizeof(float), c->stream));
    HG_HIP(c, hipMemsetAsync(c->dev_giveups, 0, sizeof(unsigned long long), c->stream));
    return HG_OK;
}

int hg_set_ref_particle_dims(hg_ctx* c, int32_t dims_x, int32_t dims_y) {
    if (!c) return HG_ERR_INVALID;
    const bool both_zero = dims_x == 0 && dims_y == 0;
    const bool in_range = dims_x >= 8 && dims_y >= 8 && dims_x <= 65536 && dims_y <= 65536;
    if (!both_zero && !in_range)
        return fail(c, HG_ERR_INVALID, "hg_set_ref_particle_dims: dims must both be 0 or both in [8, 65536]");
    c->refp_dims[0] = dims_x;
    c->refp_dims[1] = dims_y;
    return HG_OK;
}

int hg_ref_particles_read(hg_ctx* c, hg_ref_particle* out, uint32_t count) {
    if (!c || !out) return HG_ERR_INVALID;
    if (count != c->refp_count) return fail(c, HG_ERR_INVALID, "hg_ref_particles_read: count differs from hg_ref_particles_init");
    DeviceGuard guard(c->device);
    HG_HIP(c, hipMemcpyAsync(out, c->refp, size_t(count) * sizeof(hg_ref_particle), hipMemcpyDeviceToHost, c->stream));
    HG_HIP(c, hipStreamSynchronize(c->stream));
    return HG_OK;
}

int hg_ref_particles_write(hg_ctx* c, const hg_ref_particle* in, uint32_t count) {
    if (!c || !in) return HG_ERR_INVALID;
    if (count != c->refp_count) return fail(c, HG_ERR_INVALID, "hg_ref_particles_write: count differs from hg_ref_particles_init");
    DeviceGuard guard(c->device);
    HG_HIP(c, hipMemcpyAsync(c->refp, in, size_t(count) * sizeof(hg_ref_particle), hipMemcpyHostToDevice, c->stream));
    HG_HIP(c, hipStreamSynchronize(c->stream));
    return HG_OK;
}

int hg_step_particle_reference(hg_ctx* c, uint32_t iterations, int should_rain, float time0, float time_step,
                               uint32_t flags) {
    if (!c) return HG_ERR_INVALID;
    if (!c->refp) return fail(c, HG_ERR_STATE, "hg_step_particle_reference: call hg_ref_particles_init first");
    if (flags & ~(HG_STEP_THERMAL | HG_STEP_SMOOTH))
        return fail(c, HG_ERR_INVALID, "hg_step_particle_reference: only thermal/smooth flags");
    if (!std::isfinite(time0) || !std::isfinite(time_step))
        return fail(c, HG_ERR_INVALID, "hg_step_particle_reference: time must be finite");
    if (c->n < 8) return fail(c, HG_ERR_INVALID, "hg_step_particle_reference: grid too small");
    DeviceGuard guard(c->device);
    const hg_erosion_settings& e = c->es;
    hgk::RefParticleParams k{};
    k.Kc = e.Kc; k.Kconv = e.Kconv; k.Ke = e.Ke; k.G = e.G; k.dt = e.d_t;
    for (int l = 0; l < 2; ++l) { k.Ks[l] = e.Ks[l]; k.Kd[l] = e.Kd[l]; }
    k.init_volume = e.init_volume; k.friction = e.friction; k.inertia = e.inertia;
    k.min_volume = e.min_volume; k.min_velocity = e.min_velocity; k.ttl = e.ttl;
    k.particle_count = e.particle_count;
    const bool thermal = (flags & HG_STEP_THERMAL) != 0, smooth = (flags & HG_STEP_SMOOTH) != 0;
    const int n = c->n;
    k.dims_x = c->refp_dims[0] > 0 ? c->refp_dims[0] : n;
    k.dims_y = c->refp_dims[1] > 0 ? c->refp_dims[1] : n;
    for (uint32_t it = 0; it < iterations; ++it) {
        const float time = time0 + float(it) * time_step;
        int rc = timed_launch(c, "ref_particle_move", [&] {
            return hgk::launch_ref_particle_move(c->refp, c->refp_count, c->cur[HG_P_ROCK], c->cur[HG_P_DIRT],
                                                 c->cur[HG_P_WATER], c->cur[HG_P_FLUX_L], c->cur[HG_P_FLUX_R], n, k,
                                                 time, should_rain, c->stream);
        });
        if (rc != HG_OK) return rc;
        rc = timed_launch(c, "ref_particle_erode", [&] {
            return hgk::launch_ref_particle_erode(c->refp, c->refp_count, c->cur[HG_P_ROCK], c->cur[HG_P_DIRT],
                                                  c->cur[HG_P_WATER], c->cur[HG_P_FLUX_T], c->cur[HG_P_FLUX_B], n, k,
                                                  c->dev_giveups, c->stream);
        });
        if (rc != HG_OK) return rc;
        if (thermal || smooth) {
            hgk::ThermIO tio;
            tio.r = c->cur[HG_P_ROCK]; tio.d = c->cur[HG_P_DIRT];
            tio.ro = c->alt[HG_P_ROCK]; tio.dout = c->alt[HG_P_DIRT];
            rc = timed_launch(c, "fused_thermal", [&] {
                return hgk::launch_fused_thermal(c->thermal_variant, tio, n, n, thermal, smooth, c->xcd_order, c->kp,
                                                 c->stream);
            });
            if (rc != HG_OK) return rc;
            swap_planes(c, {HG_P_ROCK, HG_P_DIRT});
        }
        if (smooth) {
            rc = timed_launch(c, "ref_particle_smooth_extra", [&] {
                return hgk::launch_ref_particle_smooth_extra(c->cur[HG_P_WATER], c->cur[HG_P_FLUX_L],
                                                             c->cur[HG_P_FLUX_R], c->cur[HG_P_FLUX_T],
                                                             c->cur[HG_P_FLUX_B], n, k.particle_count, c->stream);
            });
            if (rc != HG_OK) return rc;
        }
    }
    return HG_OK;
}

int hg_ref_particle_stats(hg_ctx* c, uint64_t* cas_giveups) {
    if (!c || !cas_giveups) return HG_ERR_INVALID;
    DeviceGuard guard(c->device);
    unsigned long long g = 0;
    HG_HIP(c, hipMemcpyAsync(&g, c->dev_giveups, sizeof(g), hipMemcpyDeviceToHost, c->stream));
    HG_HIP(c, hipStreamSynchronize(c->stream));
    *cas_giveups = uint64_t(g);
    return HG_OK;
}

int hg_particle_stats(hg_ctx* c, double* exported_mass, uint64_t* droplet_steps) {
    if (!c) return HG_ERR_INVALID;
    DeviceGuard guard(c->device);
    double ex = 0;
    unsigned long long st = 0;
    HG_HIP(c, hipMemcpyAsync(&ex, c->dev_exported, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HG_HIP(c, hipMemcpyAsync(&st, c->dev_dsteps, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HG_HIP(c, hipStreamSynchronize(c->stream));
    if (exported_mass) *exported_mass = ex;
    if (droplet_steps) *droplet_steps = uint64_t(st);
    return HG_OK;
}

int hg_get_step(const hg_ctx* c, uint64_t* step) {
    if (!c || !step) return HG_ERR_INVALID;
    *step = c->step;
    return HG_OK;
}

int hg_set_step(hg_ctx* c, uint64_t step) {
    if (!c) return HG_ERR_INVALID;
    c->step = step;
    return HG_OK;
}

int hg_read_plane(hg_ctx* c, int plane, float* dst, size_t n_floats) {
    if (!c) return HG_ERR_INVALID;
    if (plane < 0 || plane >= HG_NUM_PLANES) return fail(c, HG_ERR_INVALID, "hg_read_plane: bad plane");
    if (!dst || n_floats != plane_n(c)) return fail(c, HG_ERR_INVALID, "hg_read_plane: bad buffer size");
    DeviceGuard guard(c->device);
    return read_pinned(c, c->cur[plane], dst, n_floats);
}

int hg_write_plane(hg_ctx* c, int plane, const float* src, size_t n_floats) {
    if (!c) return HG_ERR_INVALID;
    if (plane < 0 || plane >= HG_NUM_PLANES) return fail(c, HG_ERR_INVALID, "hg_write_plane: bad plane");
    if (!src || n_floats != plane_n(c)) return fail(c, HG_ERR_INVALID, "hg_write_plane: bad buffer size");
    DeviceGuard guard(c->device);
    HG_HIP(c, hipMemcpyAsync(c->cur[plane], src, n_floats * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HG_HIP(c, hipStreamSynchronize(c->stream));
    return HG_OK;
}

int hg_read(hg_ctx* c, int channel, float* rgba, size_t n_floats) {
    if (!c) return HG_ERR_INVALID;
    const size_t pn = plane_n(c);
    if (!rgba || n_floats != 4 * pn) return fail(c, HG_ERR_INVALID, "hg_read: bad buffer size");
    int planes[4] = {-1, -1, -1, -1};
    switch (channel) {
        case HG_CH_HEIGHT: planes[0] = HG_P_ROCK; planes[1] = HG_P_DIRT; planes[2] = HG_P_WATER; break;
        case HG_CH_FLUX: planes[0] = HG_P_FLUX_L; planes[1] = HG_P_FLUX_R; planes[2] = HG_P_FLUX_T; planes[3] = HG_P_FLUX_B; break;
        case HG_CH_VELOCITY: planes[0] = HG_P_VEL_U; planes[1] = HG_P_VEL_V; planes[2] = HG_P_VEL_DEPTH; break;
        case HG_CH_SEDIMENT: planes[0] = HG_P_SED_ROCK; planes[1] = HG_P_SED_DIRT; break;
        default: return fail(c, HG_ERR_INVALID, "hg_read: bad channel");
    }
    // texels assembled on the device, then one copy through the pinned chunks
    DeviceGuard guard(c->device);
    if (!c->rgba_dev) {
        void* p = nullptr;
        HG_HIP(c, hipMalloc(&p, 4 * pn * sizeof(float)));
        c->rgba_dev = static_cast<float*>(p);
    }
    const float* src[4];
    for (int k = 0; k < 4; ++k) src[k] = planes[k] < 0 ? nullptr : c->cur[planes[k]];
    HG_HIP(c, hgk::launch_interleave_rgba(c->rgba_dev, src, pn, channel == HG_CH_HEIGHT, c->stream));
    return read_pinned(c, c->rgba_dev, rgba, 4 * pn);
}

// ---- heightmap files: NumPy .npy v1.0, little-endian float32, shape (n, n, 4)
// holding the HEIGHT channel (rock, dirt, water, total) in row-major order.
}  // extern "C"
namespace {
std::string npy_header(size_t n) {
    std::string dict = "{'descr': '<f4', 'fortran_order': False, 'shape': (" + std::to_string(n) + ", " +
                       std::to_string(n) + ", 4), }";
    const size_t base = 10 + dict.size() + 1;              // magic+version+len, dict, '\n'
    dict.append((64 - base % 64) % 64, ' ');
    dict.push_back('\n');
    std::string h("\x93NUMPY\x01\x00", 8);
    h.push_back(char(dict.size() & 0xff));
    h.push_back(char(dict.size() >> 8));
    return h + dict;
}
}  // namespace
extern "C" {

int hg_save_height(hg_ctx* c, const char* path) {
    if (!c) return HG_ERR_INVALID;
    if (!path) return fail(c, HG_ERR_INVALID, "hg_save_height: path is NULL");
    const size_t pn = plane_n(c);
    std::vector<float> rgba(4 * pn);
    int rc = hg_read(c, HG_CH_HEIGHT, rgba.data(), rgba.size());
    if (rc != HG_OK) return rc;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return fail(c, HG_ERR_INVALID, std::string("hg_save_height: cannot open ") + path);
    const std::string h = npy_header(size_t(c->n));
    const bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size() &&
                    std::fwrite(rgba.data(), sizeof(float), rgba.size(), f) == rgba.size();
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return fail(c, HG_ERR_INVALID, std::string("hg_save_height: write failed: ") + path);
    return HG_OK;
}

int hg_load_height(hg_ctx* c, const char* path) {
    if (!c) return HG_ERR_INVALID;
    if (!path) return fail(c, HG_ERR_INVALID, "hg_load_height: path is NULL");
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(c, HG_ERR_INVALID, std::string("hg_load_height: cannot open ") + path);
    const std::string want = npy_header(size_t(c->n));
    std::string got(want.size(), '\0');
    const size_t pn = plane_n(c);
    std::vector<float> rgba(4 * pn);
    const bool ok = std::fread(&got[0], 1, got.size(), f) == got.size() && got == want &&
                    std::fread(rgba.data(), sizeof(float), rgba.size(), f) == rgba.size() &&
                    std::fgetc(f) == EOF;
    std::fclose(f);
    if (!ok)
        return fail(c, HG_ERR_INVALID,
                    std::string("hg_load_height: not a little-endian float32 .npy of shape (n, n, 4) for this n: ") +
                        path);
    return hg_write(c, HG_CH_HEIGHT, rgba.data(), rgba.size());
}

int hg_write(hg_ctx* c, int channel, const float* rgba, size_t n_floats) {
    if (!c) return HG_ERR_INVALID;
    const size_t pn = plane_n(c);
    if (!rgba || n_floats != 4 * pn) return fail(c, HG_ERR_INVALID, "hg_write: bad buffer size");
    int planes[4] = {-1, -1, -1, -1};
    switch (channel) {
        case HG_CH_HEIGHT: planes[0] = HG_P_ROCK; planes[1] = HG_P_DIRT; planes[2] = HG_P_WATER; break;
        case HG_CH_FLUX: planes[0] = HG_P_FLUX_L; planes[1] = HG_P_FLUX_R; planes[2] = HG_P_FLUX_T; planes[3] = HG_P_FLUX_B; break;
        case HG_CH_VELOCITY: planes[0] = HG_P_VEL_U; planes[1] = HG_P_VEL_V; planes[2] = HG_P_VEL_DEPTH; break;
        case HG_CH_SEDIMENT: planes[0] = HG_P_SED_ROCK; planes[1] = HG_P_SED_DIRT; break;
        default: return fail(c, HG_ERR_INVALID, "hg_write: bad channel");
    }
    // texels to the device, split into planes there
    DeviceGuard guard(c->device);
    if (!c->rgba_dev) {
        void* p = nullptr;
        HG_HIP(c, hipMalloc(&p, 4 * pn * sizeof(float)));
        c->rgba_dev = static_cast<float*>(p);
    }
    HG_HIP(c, hipMemcpyAsync(c->rgba_dev, rgba, 4 * pn * sizeof(float), hipMemcpyHostToDevice, c->stream));
    float* dst[4];
    for (int k = 0; k < 4; ++k) dst[k] = planes[k] < 0 ? nullptr : c->cur[planes[k]];
    HG_HIP(c, hgk::launch_deinterleave_rgba(c->rgba_dev, dst, pn, c->stream));
    HG_HIP(c, hipStreamSynchronize(c->stream));
    return HG_OK;
}

int hg_plane_device_ptr(hg_ctx* c, int plane, void** ptr) {
    if (!c) return HG_ERR_INVALID;
    if (plane < 0 || plane >= HG_NUM_PLANES || !ptr) return fail(c, HG_ERR_INVALID, "hg_plane_device_ptr: bad argument");
    *ptr = c->cur[plane];
    return HG_OK;
}

int hg_sync(hg_ctx* c) {
    if (!c) return HG_ERR_INVALID;
    DeviceGuard guard(c->device);
    HG_HIP(c, hipStreamSynchronize(c->stream));
    return HG_OK;
}

int hg_get_stream(hg_ctx* c, void** stream) {
    if (!c || !stream) return HG_ERR_INVALID;
    *stream = static_cast<void*>(c->stream);
    return HG_OK;
}

int hg_set_timing(hg_ctx* c, int enable) {
    if (!c) return HG_ERR_INVALID;
    c->timing = enable != 0;
    return HG_OK;
}

int hg_kernel_stats(hg_ctx* c, int i, const char** name, uint64_t* launches, double* total_ms) {
    if (!c) return HG_ERR_INVALID;
    DeviceGuard guard(c->device);
    const int rc = resolve_timing(c);
    if (rc != HG_OK) return rc;
    if (i < 0 || size_t(i) >= c->stats.size()) return HG_ERR_INVALID;
    if (name) *name = c->stats[i].name.c_str();
    if (launches) *launches = c->stats[i].launches;
    if (total_ms) *total_ms = c->stats[i].total_ms;
    return HG_OK;
}

int hg_tile_variant_count(int kernel) {
    if (kernel == HG_KERNEL_HYDRO) return hgk::hydro_variant_count();
    if (kernel == HG_KERNEL_THERMAL) return hgk::thermal_variant_count();
    if (kernel == HG_KERNEL_SWEEP) return hgk::sweep_variant_count();
    return HG_ERR_INVALID;
}

int hg_tile_variant_info(int kernel, int variant, int* tile_x, int* tile_y, int* halo, int* threads) {
    int tx = 0, ty = 0, h = 0, nt = 0;
    bool ok = false;
    if (kernel == HG_KERNEL_HYDRO) ok = hgk::hydro_variant_info(variant, &tx, &ty, &h, &nt);
    else if (kernel == HG_KERNEL_THERMAL) ok = hgk::thermal_variant_info(variant, &tx, &ty, &h, &nt);
    else if (kernel == HG_KERNEL_SWEEP) ok = hgk::sweep_variant_info(variant, &tx, &h, &nt, &ty);
    if (!ok) return HG_ERR_INVALID;
    if (tile_x) *tile_x = tx;
    if (tile_y) *tile_y = ty;
    if (halo) *halo = h;
    if (threads) *threads = nt;
    return HG_OK;
}

int hg_set_tiles(hg_ctx* c, int hydro_variant, int thermal_variant) {
    if (!c) return HG_ERR_INVALID;
    if (hydro_variant >= hgk::hydro_variant_count() || thermal_variant >= hgk::thermal_variant_count())
        return fail(c, HG_ERR_INVALID, "hg_set_tiles: no such variant");
    if (hydro_variant >= 0) c->hydro_variant = hydro_variant;
    if (thermal_variant >= 0) c->thermal_variant = thermal_variant;
    return HG_OK;
}

int hg_get_tiles(const hg_ctx* c, int* hydro_variant, int* thermal_variant) {
    if (!c) return HG_ERR_INVALID;
    if (hydro_variant) *hydro_variant = c->hydro_variant;
    if (thermal_variant) *thermal_variant = c->thermal_variant;
    return HG_OK;
}

int hg_set_sweep(hg_ctx* c, int variant) {
    if (!c) return HG_ERR_INVALID;
    if (variant < -1 || variant >= hgk::sweep_variant_count())
        return fail(c, HG_ERR_INVALID, "hg_set_sweep: variant is -1 or a HG_KERNEL_SWEEP variant");
    c->sweep_variant = variant;
    return HG_OK;
}

int hg_get_sweep(const hg_ctx* c, int* variant) {
    if (!c || !variant) return HG_ERR_INVALID;
    *variant = c->sweep_variant;
    return HG_OK;
}

int hg_set_xcd_order(hg_ctx* c, int mode) {
    if (!c) return HG_ERR_INVALID;
    if (mode < 0 || mode > 2) return fail(c, HG_ERR_INVALID, "hg_set_xcd_order: mode is 0, 1 or 2");
    c->xcd_order = mode;
    return HG_OK;
}

int hg_autotune(hg_ctx* c, uint32_t steps, uint32_t flags, int* hydro_best, int* thermal_best) {
    if (!c) return HG_ERR_INVALID;
    if (steps == 0) steps = 10;
    DeviceGuard guard(c->device);
    const size_t pn = plane_n(c);
    // back up the persistent planes and the step counter; the trials advance them
    void* backup = nullptr;
    HG_HIP(c, hipMalloc(&backup, pn * sizeof(float) * HG_NUM_STATE_PLANES));
    float* bk = static_cast<float*>(backup);
    for (int p = 0; p < HG_NUM_STATE_PLANES; ++p) {
        const hipError_t e = hipMemcpyAsync(bk + p * pn, c->cur[p], pn * sizeof(float), hipMemcpyDeviceToDevice,
                                            c->stream);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(backup);
            return hgctx::hip_fail(c, e, "hg_autotune: back up the state");
        }
    }
    const uint64_t step0 = c->step;
    const bool saved_timing = c->timing;
    const bool rain = (flags & HG_STEP_RAIN) && (c->rain_mode == HG_RAIN_CONSTANT ||
                                                  c->rain_mode == HG_RAIN_UNIFORM);
    const bool thermal = (flags & HG_STEP_THERMAL) != 0, smooth = (flags & HG_STEP_SMOOTH) != 0;
    int rc = resolve_timing(c);
    std::vector<KernelStat> saved_stats = c->stats;
    c->timing = true;
    const int h0 = c->hydro_variant, t0 = c->thermal_variant, s0 = c->sweep_variant;
    double best[2] = {1e300, 1e300};
    int pick[2] = {h0, t0};
    c->sweep_variant = -1;
    for (int kind = 0; kind < 2 && rc == HG_OK; ++kind) {
        const int nv = kind == 0 ? hgk::hydro_variant_count() : hgk::thermal_variant_count();
        for (int v = 0; v < nv && rc == HG_OK; ++v) {
            if (kind == 0) c->hydro_variant = v; else c->thermal_variant = v;
            rc = fused_step(c, rain, thermal, smooth, false);   // warm-up
            if (rc == HG_OK) rc = resolve_timing(c);
            c->stats.clear();
            for (uint32_t s2 = 0; s2 < steps && rc == HG_OK; ++s2) rc = fused_step(c, rain, thermal, smooth, false);
            if (rc == HG_OK) rc = resolve_timing(c);
            const char* name = kind == 0 ? "fused_hydro" : "fused_thermal";
            for (auto& st : c->stats)
                if (st.name == name && st.launches && st.total_ms / st.launches < best[kind]) {
                    best[kind] = st.total_ms / st.launches;
                    pick[kind] = v;
                }
        }
        if (kind == 0) c->hydro_variant = pick[0]; else c->thermal_variant = pick[1];
    }
    // one-kernel steps: each variant's step time against the two kernels'
    int sweep_pick = -1;
    {
        double best_step = (thermal || smooth) ? best[0] + best[1] : best[0];
        for (int v = 0; v < hgk::sweep_variant_count() && rc == HG_OK; ++v) {
            c->sweep_variant = v;
            rc = fused_step(c, rain, thermal, smooth, false);   // warm-up
            if (rc == HG_OK) rc = resolve_timing(c);
            c->stats.clear();
            for (uint32_t s2 = 0; s2 < steps && rc == HG_OK; ++s2) rc = fused_step(c, rain, thermal, smooth, false);
            if (rc == HG_OK) rc = resolve_timing(c);
            for (auto& st : c->stats)
                if (st.name == "step_sweep" && st.launches && st.total_ms / st.launches < best_step) {
                    best_step = st.total_ms / st.launches;
                    sweep_pick = v;
                }
        }
        c->sweep_variant = sweep_pick;
    }
    c->stats = saved_stats;
    c->timing = saved_timing;
    for (int p = 0; p < HG_NUM_STATE_PLANES; ++p)
        (void)hipMemcpyAsync(c->cur[p], bk + p * pn, pn * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(backup);
    c->step = step0;
    if (rc != HG_OK) {
        c->hydro_variant = h0;
        c->thermal_variant = t0;
        c->sweep_variant = s0;
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[{c->n, c->arch}] = {pick[0], pick[1], sweep_pick};
    }
    if (hydro_best) *hydro_best = pick[0];
    if (thermal_best) *thermal_best = pick[1];
    return HG_OK;
}

int hg_reset_timing(hg_ctx* c) {
    if (!c) return HG_ERR_INVALID;
    DeviceGuard guard(c->device);
    const int rc = resolve_timing(c);
    if (rc != HG_OK) return rc;
    c->stats.clear();
    return HG_OK;
}

}  // extern "C"
