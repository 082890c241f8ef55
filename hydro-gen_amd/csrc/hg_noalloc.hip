// hg_noalloc.hip — droplet iterations on caller-owned planes, allocating
// nothing.
//
// hg_step_particle runs the droplet kernel on a context's planes and adds each
// workgroup's exported mass and droplet steps to counters the context
// allocated.  hg_droplets_noalloc runs the same kernel on planes the caller
// owns; each workgroup writes its two sums to a fixed device array of this
// library (`scratch`, 2^14 entries of 16 B = 256 KiB, one copy per GPU), and
// k_droplet_totals adds them in a fixed order into the caller's outputs.  No
// call here allocates, frees or clears memory through the runtime: the
// array's address is looked up once per device (hipGetSymbolAddress), and
// every write to it is a kernel's.  Calls that overlap on one GPU share the
// array and race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "hg_context.hpp"

namespace {

constexpr uint32_t kScratchBlocks = 1u << 14;   // droplets per iteration <= 256 * 2^14 = 2^22
static_assert(256ull * kScratchBlocks == HG_DROPLETS_NOALLOC_MAX, "scratch holds one partial per workgroup");

__device__ __align__(256) hgk::DropletPartial scratch[kScratchBlocks];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device (the host-side name is
// not a device address).
int scratch_on_current_device(hgk::DropletPartial** out) {
    int dev = -1;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hgctx::hip_fail(nullptr, e, "hg_droplets_noalloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return hgctx::fail(nullptr, HG_ERR_UNSUPPORTED, "hg_droplets_noalloc: device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t eg = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (eg != hipSuccess || !p) return hgctx::hip_fail(nullptr, eg, "hg_droplets_noalloc: hipGetSymbolAddress");
        g_scratch_addr[dev] = p;
    }
    *out = static_cast<hgk::DropletPartial*>(g_scratch_addr[dev]);
    return HG_OK;
}

const char* settings_problem(const hg_droplet_settings* s) {
    const float v[] = {s->inertia, s->capacity, s->deposition, s->erosion, s->evaporation, s->gravity,
                       s->min_slope, s->init_water, s->init_speed};
    for (float x : v)
        if (!std::isfinite(x) || x < 0.0f) return "parameters must be finite and >= 0";
    if (s->inertia > 1.0f || s->evaporation > 1.0f) return "inertia and evaporation must be <= 1";
    if (s->brush_radius < 1 || s->brush_radius > 6) return "brush_radius must be in [1, 6]";
    if (s->max_lifetime > 100000) return "max_lifetime too large";
    return nullptr;
}

}  // namespace

extern "C" int hg_droplets_noalloc(float* rock, float* dirt, uint32_t n, const hg_droplet_settings* s,
                                   uint32_t first_iteration, uint32_t iterations, void* stream,
                                   double* exported_mass, uint64_t* droplet_steps) {
    using hgctx::fail;
    if (!rock || !dirt || !s) return fail(nullptr, HG_ERR_INVALID, "hg_droplets_noalloc: null pointer");
    if (n < 8 || n > 32768) return fail(nullptr, HG_ERR_INVALID, "hg_droplets_noalloc: n must be in [8, 32768]");
    if (const char* why = settings_problem(s)) return fail(nullptr, HG_ERR_INVALID, std::string("hg_droplets_noalloc: ") + why);
    if (s->count > HG_DROPLETS_NOALLOC_MAX)
        return fail(nullptr, HG_ERR_INVALID,
                    "hg_droplets_noalloc: count exceeds HG_DROPLETS_NOALLOC_MAX (4194304 droplets per iteration), "
                    "the size of the library's fixed partial-sum array");
    hgk::DropletPartial* partials = nullptr;
    int rc = scratch_on_current_device(&partials);
    if (rc != HG_OK) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hgk::DropletParams p;
    p.count = s->count; p.max_lifetime = s->max_lifetime;
    p.inertia = s->inertia; p.capacity = s->capacity; p.deposition = s->deposition;
    p.erosion = s->erosion; p.evaporation = s->evaporation; p.gravity = s->gravity;
    p.min_slope = s->min_slope; p.init_water = s->init_water; p.init_speed = s->init_speed;
    p.seed_lo = uint32_t(s->seed); p.seed_hi = uint32_t(s->seed >> 32);
    const hgk::Brush brush = hgk::make_brush(s->brush_radius);
    const uint32_t nblocks = (s->count + 255) / 256;
    const bool totals = exported_mass || droplet_steps;
    if (iterations == 0 && totals) {
        const hipError_t e = hgk::launch_droplet_totals(partials, 0, exported_mass, droplet_steps, false, st);
        if (e != hipSuccess) return hgctx::hip_fail(nullptr, e, "hg_droplets_noalloc: totals launch");
    }
    for (uint32_t it = 0; it < iterations; ++it) {
        hipError_t e = hgk::launch_droplets_partials(rock, dirt, int(n), p, brush, first_iteration + it, partials, st);
        if (e != hipSuccess) return hgctx::hip_fail(nullptr, e, "hg_droplets_noalloc: droplet launch");
        if (totals) {
            e = hgk::launch_droplet_totals(partials, nblocks, exported_mass, droplet_steps, it > 0, st);
            if (e != hipSuccess) return hgctx::hip_fail(nullptr, e, "hg_droplets_noalloc: totals launch");
        }
    }
    return HG_OK;
}
