// hg_context.hpp — the context object behind hg_ctx* and the launch helpers
// shared by the runtime translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <initializer_list>
#include <string>
#include <utility>
#include <vector>

#include "hg_kernels.hpp"

using hgk::KParams;
using hgk::Planes;

struct KernelStat {
    std::string name;
    uint64_t launches = 0;
    double total_ms = 0.0;
};

struct PendingTiming {
    hipEvent_t start, end;
    int stat;
};

struct hg_ctx {
    int device = 0;
    int n = 0;
    hipStream_t stream = nullptr;
    void* block = nullptr;           // one allocation for all plane buffers
    float* cur[HG_NUM_PLANES] = {};
    float* alt[HG_NUM_PLANES] = {};
    unsigned* minmax = nullptr;      // 2 words for the init reduction
    hg_erosion_settings es{};
    hg_rain_settings rs{};
    int rain_mode = HG_RAIN_OFF;
    uint64_t seed = 0;
    uint64_t step = 0;
    KParams kp{};
    hg_droplet_settings ds{};
    uint32_t particle_iter = 0;
    int droplet_order = 0;           // hg_set_droplet_order
    uint32_t* drop_order = nullptr;  // slot -> droplet id (drop_cap words), then the tile offsets
    size_t drop_cap = 0;
    double* dev_exported = nullptr;  // cumulative droplet statistics (device)
    unsigned long long* dev_dsteps = nullptr;
    int hydro_variant = 0;           // fused kernel tile variants (see hg_autotune)
    int thermal_variant = 0;
    int sweep_variant = 0;           // one-kernel step variant; -1: the hydro and thermal kernels
    int xcd_order = 2;               // fused kernels' tile order (hg_set_xcd_order)
    float rain_time = 0.0f;          // HG_RAIN_REFERENCE time seed
    float map_max_height = 256.0f;   // HG_RAIN_REFERENCE mountain scale (src/settings.hpp:30)
    hg_ref_particle* refp = nullptr; // the reference particle mode's particles (hg_ref_particles_init)
    uint32_t refp_count = 0;
    int refp_dims[2] = {0, 0};       // hmap_dims of the reference particle mode (0 = n)
    unsigned long long* dev_giveups = nullptr;  // in the block's tail
    // hg_adjoint_rollout's workspace, recompute stream (default priority) and
    // its events, kept for the next call until hg_destroy
    void* adj_block = nullptr;
    size_t adj_bytes = 0;
    hipStream_t adj_stream = nullptr;
    hipEvent_t adj_ev[5] = {};
    int adj_fuse = HG_ADJ_FUSE_EROSION_FLUX | HG_ADJ_FUSE_SMOOTH_THERMAL1 | HG_ADJ_FUSE_FAR_THERMAL0 |
                   HG_ADJ_FUSE_TRANSPORT_BESIDE;   // hg_set_adjoint_fusion
    std::string arch;
    std::string err;
    // timing
    bool timing = false;
    std::vector<KernelStat> stats;
    std::vector<PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    // pinned staging for device-to-host reads (two chunks, allocated on the
    // first read): copies into pageable memory straight from the device ran
    // at ~5 GB/s
    void* pin[2] = {};
    hipEvent_t pin_ev[2] = {};
    float* rgba_dev = nullptr;       // 4 n^2 floats for hg_read's texels (allocated on first use)
};

namespace hgctx {
int fail(hg_ctx* c, int code, const std::string& msg);
int hip_fail(hg_ctx* c, hipError_t e, const char* what);
size_t plane_n(const hg_ctx* c);
Planes planes_of(float* const* p);
int resolve_timing(hg_ctx* c);
int stat_index(hg_ctx* c, const char* name);
hipEvent_t take_event(hg_ctx* c);
int run_pass(hg_ctx* c, int pass);
int fused_step(hg_ctx* c, bool rain, bool thermal, bool smooth, bool keep_velocity);
// n_floats from device `src` to host `dst` through the context's pinned chunks
// (ordered after the work on the context's stream; returns when dst is filled)
int read_pinned(hg_ctx* c, const float* src, float* dst, size_t n_floats);

// Launch helper: brackets the launch with events when timing is on.
template <class F>
int timed_launch(hg_ctx* c, const char* name, F&& launch) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (!e0 || !e1) return fail(c, HG_ERR_HIP, "hipEventCreate failed");
        const hipError_t er = hipEventRecord(e0, c->stream);
        if (er != hipSuccess) return hip_fail(c, er, "hipEventRecord");
    }
    const hipError_t e = launch();
    if (e != hipSuccess) return hip_fail(c, e, name);
    if (c->timing) {
        const hipError_t er = hipEventRecord(e1, c->stream);
        if (er != hipSuccess) return hip_fail(c, er, "hipEventRecord");
        c->pending.push_back(PendingTiming{e0, e1, stat_index(c, name)});
        if (c->pending.size() > 4096) {
            const int rc = resolve_timing(c);
            if (rc != HG_OK) return rc;
        }
    }
    return HG_OK;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};
}  // namespace hgctx

#define HG_HIP(ctx, call)                                                 \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return hgctx::hip_fail((ctx), e_, #call);   \
    } while (0)
