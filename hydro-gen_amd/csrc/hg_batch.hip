// hg_batch.hip — many independent terrains of uneven sizes stepped together.
//
// The reference erodes one terrain per GL context (src/erosion.cpp:158-200);
// a batch holds T terrains (square, sides multiple of 8) with their own
// coefficients, advanced in lockstep.  One step is one persistent launch of
// the hydro kernel and one of the thermal kernel over a work-list of the tiles
// of all terrains; hg_batch_capture records nsteps such steps in a hipGraph.
//
// Work-list order: natural (terrain, tile row, tile column), or cut into 8
// contiguous chunks interleaved so that list position t is handled by a
// workgroup on XCD t mod 8 (workgroups are dealt to the 8 XCDs round robin),
// which keeps neighbouring tiles, and the halo rows they share, in one XCD's
// L2.  Both orders measure the same on MI355X; natural is the default.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "hg_kernels.hpp"

using hgk::BatchTerrain;
using hgk::KParams;

struct hg_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int count = 0;
    int cu_count = 0;
    std::vector<int> n;
    std::vector<size_t> offset;         // float offset of terrain t's plane block
    void* block = nullptr;              // 2 buffers x 9 planes per terrain
    void* scratch = nullptr;            // per-pass baseline: velocity + thermal planes (largest terrain)
    std::vector<hg_erosion_settings> es;
    std::vector<BatchTerrain> terr;     // host copy of the descriptors
    BatchTerrain* dev_terr = nullptr;
    uint32_t* dev_step = nullptr;
    uint32_t cur = 0;                   // bit p: index of plane p's current buffer
    uint64_t step = 0;
    hg_rain_settings rs{};
    int rain_mode = HG_RAIN_OFF;
    uint64_t seed = 0;
    int hydro_variant = 0, thermal_variant = 1;
    size_t total_floats = 0;            // size of `block`
    bool persistent = false;            // false: one workgroup per work-list entry
    bool xcd_order = true;              // interleave the work-list for the 8 XCDs
    uint32_t* dev_tiles_h = nullptr;
    uint32_t* dev_tiles_t = nullptr;
    int ntiles_h = 0, ntiles_t = 0;
    // the work-lists in natural (terrain, row, column) order, and each
    // terrain's first entry and count in them: per-terrain launches
    uint32_t* dev_nat_h = nullptr;
    uint32_t* dev_nat_t = nullptr;
    std::vector<int> first_h, count_h, first_t, count_t;
    int grid_h = 0, grid_t = 0;
    bool tiles_dirty = true, terr_dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t graph_steps = 0;
    uint32_t graph_cur = 0;             // buffer selection baked into the graph's launches
    std::string err;
};

namespace {

thread_local std::string g_batch_create_error;

int bfail(hg_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg; else g_batch_create_error = msg;
    return code;
}

#define BH(b, call)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return bfail((b), HG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Guard {
    int prev = -1;
    explicit Guard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~Guard() {
        int c = -1;
        if (prev >= 0 && hipGetDevice(&c) == hipSuccess && c != prev) (void)hipSetDevice(prev);
    }
};

size_t pitch_floats(int n) { return hgk::plane_stride_bytes(size_t(n) * n) / sizeof(float); }

KParams kparams(const hg_erosion_settings& s, const hg_batch* b) {
    KParams k{};
    k.Kc = s.Kc; k.Kconv = s.Kconv; k.Ke = s.Ke; k.E = s.energy_kept; k.G = s.G; k.dt = s.d_t;
    for (int l = 0; l < 2; ++l) {
        k.Ks[l] = s.Ks[l]; k.Kd[l] = s.Kd[l]; k.Kalpha[l] = s.Kalpha[l]; k.Kspeed[l] = s.Kspeed[l];
        k.tan_alpha[l] = std::tan(s.Kalpha[l]);
        k.thr_diag[l] = std::sqrt(2.0f) * k.tan_alpha[l];
    }
    const float m = s.Kspeed[1] * s.d_t;
    k.smooth_m = m < 0.0f ? 0.0f : (m > 1.0f ? 1.0f : m);
    k.evap = 1.0f - s.Ke * s.d_t;
    k.rain_amount = b->rs.amount;
    k.rain_mode = b->rain_mode;
    k.seed_lo = uint32_t(b->seed);
    k.seed_hi = uint32_t(b->seed >> 32);
    k.store_nt = hgk::stream_stores(b->total_floats * sizeof(float));   // every terrain, both buffers
    return k;
}

float* plane_ptr(const hg_batch* b, int t, int p, int which) {
    const size_t pf = pitch_floats(b->n[t]);
    return static_cast<float*>(b->block) + b->offset[t] + (size_t(2 * p + which)) * pf;
}

int upload_terr(hg_batch* b) {
    if (!b->terr_dirty) return HG_OK;
    for (int t = 0; t < b->count; ++t) {
        BatchTerrain& T = b->terr[t];
        for (int p = 0; p < HG_NUM_STATE_PLANES; ++p)
            for (int w = 0; w < 2; ++w) T.plane[p][w] = plane_ptr(b, t, p, w);
        T.n = b->n[t];
        T.k = kparams(b->es[t], b);
    }
    BH(b, hipMemcpyAsync(b->dev_terr, b->terr.data(), sizeof(BatchTerrain) * b->count,
                         hipMemcpyHostToDevice, b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    b->terr_dirty = false;
    return HG_OK;
}

// Work-list of all tiles for a TX x TY tiling in natural order; `first` and
// `count` receive each terrain's range in it.
std::vector<uint32_t> natural_list(const hg_batch* b, int tx, int ty, std::vector<int>* first,
                                   std::vector<int>* count) {
    std::vector<uint32_t> nat;
    for (int t = 0; t < b->count; ++t) {
        if (first) first->push_back(int(nat.size()));
        const int nbx = (b->n[t] + tx - 1) / tx, nby = (b->n[t] + ty - 1) / ty;
        for (int y = 0; y < nby; ++y)
            for (int x = 0; x < nbx; ++x) nat.push_back((uint32_t(t) << 24) | (uint32_t(y) << 12) | uint32_t(x));
        if (count) count->push_back(nbx * nby);
    }
    return nat;
}

// XCD-interleaved work-list of all tiles for a TX x TY tiling.
std::vector<uint32_t> work_list(const hg_batch* b, int tx, int ty) {
    std::vector<uint32_t> nat = natural_list(b, tx, ty, nullptr, nullptr);
    if (!b->xcd_order) return nat;
    const size_t N = nat.size(), chunk = (N + 7) / 8;
    std::vector<uint32_t> out(chunk * 8, 0xffffffffu);
    for (size_t u = 0; u < N; ++u) out[(u % chunk) * 8 + u / chunk] = nat[u];
    return out;
}

int build_tiles(hg_batch* b) {
    if (!b->tiles_dirty) return HG_OK;
    int tx, ty, h, nt;
    hgk::hydro_variant_info(b->hydro_variant, &tx, &ty, &h, &nt);
    std::vector<uint32_t> lh = work_list(b, tx, ty);
    hgk::thermal_variant_info(b->thermal_variant, &tx, &ty, &h, &nt);
    std::vector<uint32_t> lt = work_list(b, tx, ty);
    if (b->dev_tiles_h) (void)hipFree(b->dev_tiles_h);
    if (b->dev_tiles_t) (void)hipFree(b->dev_tiles_t);
    b->dev_tiles_h = b->dev_tiles_t = nullptr;
    {
        void* ph = nullptr;
        void* pt = nullptr;
        BH(b, hipMalloc(&ph, lh.size() * sizeof(uint32_t)));
        b->dev_tiles_h = static_cast<uint32_t*>(ph);
        BH(b, hipMalloc(&pt, lt.size() * sizeof(uint32_t)));
        b->dev_tiles_t = static_cast<uint32_t*>(pt);
    }
    BH(b, hipMemcpy(b->dev_tiles_h, lh.data(), lh.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    BH(b, hipMemcpy(b->dev_tiles_t, lt.data(), lt.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    b->ntiles_h = int(lh.size());
    b->ntiles_t = int(lt.size());
    {
        b->first_h.clear(); b->count_h.clear(); b->first_t.clear(); b->count_t.clear();
        hgk::hydro_variant_info(b->hydro_variant, &tx, &ty, &h, &nt);
        std::vector<uint32_t> nh = natural_list(b, tx, ty, &b->first_h, &b->count_h);
        hgk::thermal_variant_info(b->thermal_variant, &tx, &ty, &h, &nt);
        std::vector<uint32_t> ntl = natural_list(b, tx, ty, &b->first_t, &b->count_t);
        if (b->dev_nat_h) (void)hipFree(b->dev_nat_h);
        if (b->dev_nat_t) (void)hipFree(b->dev_nat_t);
        b->dev_nat_h = b->dev_nat_t = nullptr;
        void* ph = nullptr;
        void* pt = nullptr;
        BH(b, hipMalloc(&ph, nh.size() * sizeof(uint32_t)));
        b->dev_nat_h = static_cast<uint32_t*>(ph);
        BH(b, hipMalloc(&pt, ntl.size() * sizeof(uint32_t)));
        b->dev_nat_t = static_cast<uint32_t*>(pt);
        BH(b, hipMemcpy(b->dev_nat_h, nh.data(), nh.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        BH(b, hipMemcpy(b->dev_nat_t, ntl.data(), ntl.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    // persistent grids: resident workgroups, a multiple of 8 (one per XCD slot)
    b->grid_h = b->cu_count * hgk::batch_hydro_blocks_per_cu(b->hydro_variant);
    b->grid_t = b->cu_count * hgk::batch_thermal_blocks_per_cu(b->thermal_variant);
    b->grid_h = std::max(8, b->grid_h / 8 * 8);
    b->grid_t = std::max(8, b->grid_t / 8 * 8);
    if (!b->persistent) {
        b->grid_h = b->ntiles_h;
        b->grid_t = b->ntiles_t;
    }
    b->tiles_dirty = false;
    return HG_OK;
}

void drop_graph(hg_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_steps = 0;
}

// Enqueue the launches of `nsteps` fused steps starting at offset 0 from
// *dev_step, then advance *dev_step.  Updates b->cur.  One launch of each
// fused kernel per step covers every terrain; with `per_terrain` each terrain
// gets launches of its own (the same kernels on its range of the natural
// work-list), the per-terrain pattern the batch is measured against.
int enqueue_steps(hg_batch* b, uint32_t nsteps, uint32_t flags, bool per_terrain = false) {
    const bool rain = (flags & HG_STEP_RAIN) && b->rain_mode != HG_RAIN_OFF;
    const bool thermal = (flags & HG_STEP_THERMAL) != 0, smooth = (flags & HG_STEP_SMOOTH) != 0;
    const bool nts = hgk::stream_stores(b->total_floats * sizeof(float)) != 0;   // every terrain, both buffers
    for (uint32_t s = 0; s < nsteps; ++s) {
        if (per_terrain) {
            for (int t = 0; t < b->count; ++t)
                BH(b, hgk::launch_batch_hydro(b->hydro_variant, b->dev_terr, b->dev_nat_h + b->first_h[t],
                                              b->count_h[t], b->count_h[t], b->cur, b->dev_step, s, rain, nts,
                                              b->stream));
        } else {
            BH(b, hgk::launch_batch_hydro(b->hydro_variant, b->dev_terr, b->dev_tiles_h, b->ntiles_h, b->grid_h,
                                          b->cur, b->dev_step, s, rain, nts, b->stream));
        }
        b->cur ^= (1u << HG_NUM_STATE_PLANES) - 1u;
        if (thermal || smooth) {
            if (per_terrain) {
                for (int t = 0; t < b->count; ++t)
                    BH(b, hgk::launch_batch_thermal(b->thermal_variant, b->dev_terr, b->dev_nat_t + b->first_t[t],
                                                    b->count_t[t], b->count_t[t], b->cur, thermal, smooth, nts,
                                                    b->stream));
            } else {
                BH(b, hgk::launch_batch_thermal(b->thermal_variant, b->dev_terr, b->dev_tiles_t, b->ntiles_t,
                                                b->grid_t, b->cur, thermal, smooth, nts, b->stream));
            }
            b->cur ^= (1u << HG_P_ROCK) | (1u << HG_P_DIRT);
        }
    }
    BH(b, hgk::launch_step_advance(b->dev_step, nsteps, b->stream));
    return HG_OK;
}

// Per-terrain launches (the baseline the batch is measured against): for each
// step, each terrain in turn gets the single-terrain fused kernels or, with
// per_pass, one launch per reference pass.
int per_terrain_steps(hg_batch* b, uint32_t nsteps, uint32_t flags, bool per_pass) {
    const bool rain = (flags & HG_STEP_RAIN) && b->rain_mode != HG_RAIN_OFF;
    const bool thermal = (flags & HG_STEP_THERMAL) != 0, smooth = (flags & HG_STEP_SMOOTH) != 0;
    int maxn = 0;
    for (int v : b->n) maxn = std::max(maxn, v);
    const size_t spf = pitch_floats(maxn);
    if (per_pass && !b->scratch) {
        void* p = nullptr;
        BH(b, hipMalloc(&p, spf * sizeof(float) * 2 * (HG_NUM_PLANES - HG_NUM_STATE_PLANES)));
        b->scratch = p;
    }
    for (uint32_t s = 0; s < nsteps; ++s) {
        const uint32_t step = uint32_t(b->step + s);
        uint32_t cur_after = b->cur;
        for (int t = 0; t < b->count; ++t) {
            const int n = b->n[t];
            const KParams& k = b->terr[t].k;
            uint32_t cur = b->cur;   // 20-bit mask: state planes + scratch planes
            auto ptr = [&](int p, int which) -> float* {
                if (p < HG_NUM_STATE_PLANES) return plane_ptr(b, t, p, which);
                return static_cast<float*>(b->scratch) + (size_t(2 * (p - HG_NUM_STATE_PLANES) + which)) * spf;
            };
            if (!per_pass) {
                hgk::HydroIO io;
                const float** in[9] = {&io.r, &io.d, &io.b, &io.fl, &io.fr, &io.ft, &io.fb, &io.s0, &io.s1};
                float** out[9] = {&io.ro, &io.dout, &io.bo, &io.flo, &io.fro, &io.fto, &io.fbo, &io.s0o, &io.s1o};
                for (int p = 0; p < 9; ++p) {
                    const int c = (cur >> p) & 1;
                    *in[p] = ptr(p, c);
                    *out[p] = ptr(p, c ^ 1);
                }
                io.vx = io.vy = io.vz = nullptr;
                BH(b, hgk::launch_fused_hydro(b->hydro_variant, io, n, n, step, uint32_t(t), rain, 2, k,
                                              b->stream));
                cur ^= (1u << HG_NUM_STATE_PLANES) - 1u;
                if (thermal || smooth) {
                    hgk::ThermIO tio;
                    const int cr = cur & 1, cd = (cur >> 1) & 1;
                    tio.r = ptr(HG_P_ROCK, cr); tio.ro = ptr(HG_P_ROCK, cr ^ 1);
                    tio.d = ptr(HG_P_DIRT, cd); tio.dout = ptr(HG_P_DIRT, cd ^ 1);
                    BH(b, hgk::launch_fused_thermal(b->thermal_variant, tio, n, n, thermal, smooth, 2, k,
                                                    b->stream));
                    cur ^= (1u << HG_P_ROCK) | (1u << HG_P_DIRT);
                }
                cur_after = cur;
                continue;
            }
            auto run = [&](int pass, std::initializer_list<int> produced) -> int {
                hgk::Planes in, out;
                for (int p = 0; p < HG_NUM_PLANES; ++p) {
                    const int c = (cur >> p) & 1;
                    in.p[p] = ptr(p, c);
                    out.p[p] = ptr(p, c ^ 1);
                }
                hgk::KParams kk = k;
                BH(b, hgk::launch_pass(pass, in, out, n, n, step, uint32_t(t), kk, b->stream));
                for (int p : produced) cur ^= 1u << p;
                return HG_OK;
            };
            int rc = HG_OK;
            if (rain) rc = run(HG_PASS_RAIN, {HG_P_WATER});
            if (rc == HG_OK)
                rc = run(HG_PASS_FLUX, {HG_P_WATER, HG_P_FLUX_L, HG_P_FLUX_R, HG_P_FLUX_T, HG_P_FLUX_B,
                                        HG_P_VEL_U, HG_P_VEL_V, HG_P_VEL_DEPTH});
            if (rc == HG_OK) rc = run(HG_PASS_EROSION, {HG_P_ROCK, HG_P_DIRT, HG_P_SED_ROCK, HG_P_SED_DIRT});
            if (rc == HG_OK) rc = run(HG_PASS_TRANSPORT, {HG_P_SED_ROCK, HG_P_SED_DIRT, HG_P_WATER});
            if (thermal) {
                const std::initializer_list<int> tp = {HG_P_THERM_L, HG_P_THERM_R, HG_P_THERM_T, HG_P_THERM_B,
                                                       HG_P_THERM_LT, HG_P_THERM_RT, HG_P_THERM_LB, HG_P_THERM_RB};
                if (rc == HG_OK) rc = run(HG_PASS_THERMAL_FLUX_0, tp);
                if (rc == HG_OK) rc = run(HG_PASS_THERMAL_TRANSPORT_0, {HG_P_ROCK});
                if (rc == HG_OK) rc = run(HG_PASS_THERMAL_FLUX_1, tp);
                if (rc == HG_OK) rc = run(HG_PASS_THERMAL_TRANSPORT_1, {HG_P_DIRT});
            }
            if (rc == HG_OK && smooth) rc = run(HG_PASS_SMOOTH, {HG_P_ROCK, HG_P_DIRT});
            if (rc != HG_OK) return rc;
            cur_after = cur;
        }
        // every terrain went through the same swaps; keep only the state bits,
        // the scratch planes are re-used from either buffer next step
        b->cur = cur_after & ((1u << HG_NUM_STATE_PLANES) - 1u);
    }
    return HG_OK;
}

}  // namespace

extern "C" {

int hg_batch_create(int device, int count, const uint32_t* sizes, hg_batch** out) {
    if (!out) return bfail(nullptr, HG_ERR_INVALID, "hg_batch_create: out is NULL");
    *out = nullptr;
    if (count < 1 || count > 255 || !sizes)
        return bfail(nullptr, HG_ERR_INVALID, "hg_batch_create: count must be in [1, 255]");
    for (int t = 0; t < count; ++t)
        if (sizes[t] < 8 || sizes[t] > 16384 || sizes[t] % 8)
            return bfail(nullptr, HG_ERR_INVALID, "hg_batch_create: sizes must be multiples of 8 in [8, 16384]");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return bfail(nullptr, HG_ERR_HIP, "hg_batch_create: no HIP device");
    if (device < 0 || device >= ndev) return bfail(nullptr, HG_ERR_INVALID, "hg_batch_create: bad device");
    Guard g(device);
    hg_batch* b = new hg_batch();
    b->device = device;
    b->count = count;
    size_t total = 0;
    for (int t = 0; t < count; ++t) {
        b->n.push_back(int(sizes[t]));
        b->offset.push_back(total);
        total += pitch_floats(int(sizes[t])) * 2 * HG_NUM_STATE_PLANES;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) prop.multiProcessorCount = 256;
    b->cu_count = prop.multiProcessorCount;
    auto cleanup = [&](const std::string& m, int code) {
        if (b->block) (void)hipFree(b->block);
        if (b->dev_terr) (void)hipFree(b->dev_terr);
        if (b->dev_step) (void)hipFree(b->dev_step);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return bfail(nullptr, code, m);
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup("hipStreamCreate failed", HG_ERR_HIP);
    b->total_floats = total;
    if (hipMalloc(&b->block, total * sizeof(float)) != hipSuccess)
        return cleanup("hipMalloc of the batch state failed", HG_ERR_OOM);
    if (hipMalloc(&b->dev_terr, sizeof(BatchTerrain) * count) != hipSuccess ||
        hipMalloc(&b->dev_step, 256) != hipSuccess)
        return cleanup("hipMalloc failed", HG_ERR_OOM);
    if (hipMemsetAsync(b->block, 0, total * sizeof(float), b->stream) != hipSuccess ||
        hipMemsetAsync(b->dev_step, 0, 256, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return cleanup("hipMemset failed", HG_ERR_HIP);
    hg_erosion_settings d{};
    d.Kc = 0.2f; d.Kalpha[0] = 1.3f; d.Kalpha[1] = 0.6f; d.Kconv = 0.001f;
    d.Ks[0] = 0.03f; d.Ks[1] = 0.09f; d.Kd[0] = 0.01f; d.Kd[1] = 0.03f; d.Ke = 0.03f;
    d.energy_kept = 1.0f; d.Kspeed[0] = 0.5f; d.Kspeed[1] = 2.0f; d.G = 1.0f; d.d_t = 0.001f;
    b->es.assign(count, d);
    b->terr.resize(count);
    b->rs.amount = 0.01f; b->rs.mountain_thresh = 0.55f; b->rs.mountain_multip = 0.05f;
    b->rs.period = 512; b->rs.drops = 0.02f;
    *out = b;
    return HG_OK;
}

int hg_batch_destroy(hg_batch* b) {
    if (!b) return HG_OK;
    Guard g(b->device);
    (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->dev_tiles_h) (void)hipFree(b->dev_tiles_h);
    if (b->dev_tiles_t) (void)hipFree(b->dev_tiles_t);
    if (b->dev_nat_h) (void)hipFree(b->dev_nat_h);
    if (b->dev_nat_t) (void)hipFree(b->dev_nat_t);
    if (b->scratch) (void)hipFree(b->scratch);
    (void)hipFree(b->block);
    (void)hipFree(b->dev_terr);
    (void)hipFree(b->dev_step);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return HG_OK;
}

const char* hg_batch_last_error(const hg_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int hg_batch_count(const hg_batch* b, int* count) {
    if (!b || !count) return HG_ERR_INVALID;
    *count = b->count;
    return HG_OK;
}

int hg_batch_set_erosion(hg_batch* b, int terrain, const hg_erosion_settings* s) {
    if (!b) return HG_ERR_INVALID;
    if (!s || terrain < -1 || terrain >= b->count) return bfail(b, HG_ERR_INVALID, "hg_batch_set_erosion: bad argument");
    if (!(s->d_t > 0.0f) || !std::isfinite(s->d_t) || !std::isfinite(s->Kc) || s->Kalpha[0] < 0.0f ||
        s->Kalpha[1] < 0.0f || !(s->Kalpha[0] < 1.5707963f) || !(s->Kalpha[1] < 1.5707963f))
        return bfail(b, HG_ERR_INVALID, "hg_batch_set_erosion: invalid settings");
    for (int t = 0; t < b->count; ++t)
        if (terrain < 0 || t == terrain) b->es[t] = *s;
    b->terr_dirty = true;
    drop_graph(b);
    return HG_OK;
}

int hg_batch_set_rain(hg_batch* b, const hg_rain_settings* s, int mode, uint64_t seed) {
    if (!b) return HG_ERR_INVALID;
    if (!s || mode < HG_RAIN_OFF || mode > HG_RAIN_UNIFORM || !std::isfinite(s->amount) || s->amount < 0.0f)
        return bfail(b, HG_ERR_INVALID, "hg_batch_set_rain: bad argument");
    b->rs = *s;
    b->rain_mode = mode;
    b->seed = seed;
    b->terr_dirty = true;
    drop_graph(b);
    return HG_OK;
}

int hg_batch_init_fbm(hg_batch* b, int terrain, const hg_fbm_params* p) {
    if (!b) return HG_ERR_INVALID;
    if (!p || terrain < 0 || terrain >= b->count || p->octaves < 1 || p->octaves > 32 || !(p->height_hi >= p->height_lo))
        return bfail(b, HG_ERR_INVALID, "hg_batch_init_fbm: bad argument");
    Guard g(b->device);
    const int n = b->n[terrain];
    const size_t pn = size_t(n) * n;
    for (int pl = 0; pl < HG_NUM_STATE_PLANES; ++pl)
        for (int w = 0; w < 2; ++w)
            if (!(pl == HG_P_ROCK && w == int((b->cur >> pl) & 1)))
                BH(b, hgk::launch_fill(plane_ptr(b, terrain, pl, w), pn, 0.0f, b->stream));
    uint8_t perm[256];
    hgk::fbm_permutation(p->seed, perm);
    float* rock = plane_ptr(b, terrain, HG_P_ROCK, int(b->cur & 1));
    BH(b, hgk::launch_fbm(rock, n, n, perm, p->octaves, p->lacunarity, p->gain, p->base_freq, b->stream));
    unsigned* mm = reinterpret_cast<unsigned*>(b->dev_step) + 16;
    BH(b, hgk::launch_minmax(rock, pn, mm, b->stream));
    BH(b, hgk::launch_rescale(rock, pn, mm, p->height_lo, p->height_hi, b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    return HG_OK;
}

int hg_batch_set_tiles(hg_batch* b, int hydro_variant, int thermal_variant) {
    if (!b) return HG_ERR_INVALID;
    if (hydro_variant >= hgk::hydro_variant_count() || thermal_variant >= hgk::thermal_variant_count())
        return bfail(b, HG_ERR_INVALID, "hg_batch_set_tiles: no such variant");
    if (hydro_variant >= 0) b->hydro_variant = hydro_variant;
    if (thermal_variant >= 0) b->thermal_variant = thermal_variant;
    b->tiles_dirty = true;
    drop_graph(b);
    return HG_OK;
}

int hg_batch_autotune(hg_batch* b, uint32_t steps, uint32_t flags, int* hydro_best, int* thermal_best) {
    if (!b) return HG_ERR_INVALID;
    if (flags & ~(HG_STEP_RAIN | HG_STEP_THERMAL | HG_STEP_SMOOTH))
        return bfail(b, HG_ERR_INVALID, "hg_batch_autotune: only rain/thermal/smooth flags");
    if (steps == 0) steps = 4;
    Guard g(b->device);
    int rc = upload_terr(b);
    if (rc != HG_OK) return rc;
    // the trials advance the state: back up both buffers of every plane, the
    // buffer selection and the device step counter
    void* backup = nullptr;
    BH(b, hipMalloc(&backup, b->total_floats * sizeof(float) + 256));
    struct Free {
        void* p;
        ~Free() { (void)hipFree(p); }
    } free_backup{backup};
    char* bk = static_cast<char*>(backup);
    BH(b, hipMemcpyAsync(bk, b->block, b->total_floats * sizeof(float), hipMemcpyDeviceToDevice, b->stream));
    BH(b, hipMemcpyAsync(bk + b->total_floats * sizeof(float), b->dev_step, 256, hipMemcpyDeviceToDevice,
                         b->stream));
    const uint32_t cur0 = b->cur;
    const uint64_t step0 = b->step;
    const int h0 = b->hydro_variant, t0 = b->thermal_variant;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BH(b, hipEventCreate(&e0));
    BH(b, hipEventCreate(&e1));
    int pick[2] = {h0, t0};
    for (int kind = 0; kind < 2 && rc == HG_OK; ++kind) {
        const int nv = kind == 0 ? hgk::hydro_variant_count() : hgk::thermal_variant_count();
        float best = 1e30f;
        for (int v = 0; v < nv && rc == HG_OK; ++v) {
            if (kind == 0) b->hydro_variant = v; else b->thermal_variant = v;
            b->tiles_dirty = true;
            rc = build_tiles(b);
            if (rc == HG_OK) rc = enqueue_steps(b, 1, flags);            // warm-up
            if (rc == HG_OK && hipEventRecord(e0, b->stream) != hipSuccess) rc = bfail(b, HG_ERR_HIP, "event");
            if (rc == HG_OK) rc = enqueue_steps(b, steps, flags);
            if (rc == HG_OK && (hipEventRecord(e1, b->stream) != hipSuccess ||
                                hipEventSynchronize(e1) != hipSuccess))
                rc = bfail(b, HG_ERR_HIP, "event");
            float ms = 0.0f;
            if (rc == HG_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best) {
                best = ms;
                pick[kind] = v;
            }
        }
        if (kind == 0) b->hydro_variant = pick[0]; else b->thermal_variant = pick[1];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipMemcpyAsync(b->block, bk, b->total_floats * sizeof(float), hipMemcpyDeviceToDevice, b->stream);
    (void)hipMemcpyAsync(b->dev_step, bk + b->total_floats * sizeof(float), 256, hipMemcpyDeviceToDevice,
                         b->stream);
    (void)hipStreamSynchronize(b->stream);
    b->cur = cur0;
    b->step = step0;
    b->tiles_dirty = true;
    drop_graph(b);
    if (rc != HG_OK) {
        b->hydro_variant = h0;
        b->thermal_variant = t0;
        return rc;
    }
    if (hydro_best) *hydro_best = pick[0];
    if (thermal_best) *thermal_best = pick[1];
    return HG_OK;
}

int hg_batch_set_xcd_order(hg_batch* b, int on) {
    if (!b) return HG_ERR_INVALID;
    b->xcd_order = on != 0;
    b->tiles_dirty = true;
    drop_graph(b);
    return HG_OK;
}

int hg_batch_set_persistent(hg_batch* b, int persistent) {
    if (!b) return HG_ERR_INVALID;
    b->persistent = persistent != 0;
    b->tiles_dirty = true;
    drop_graph(b);
    return HG_OK;
}

int hg_batch_step(hg_batch* b, uint32_t nsteps, uint32_t flags) {
    if (!b) return HG_ERR_INVALID;
    const uint32_t known = HG_STEP_RAIN | HG_STEP_THERMAL | HG_STEP_SMOOTH | HG_STEP_PER_PASS | HG_STEP_PER_TERRAIN;
    if (flags & ~known) return bfail(b, HG_ERR_INVALID, "hg_batch_step: unknown flag bits");
    if ((flags & HG_STEP_PER_PASS) && !(flags & HG_STEP_PER_TERRAIN))
        return bfail(b, HG_ERR_INVALID, "hg_batch_step: HG_STEP_PER_PASS needs HG_STEP_PER_TERRAIN");
    Guard g(b->device);
    int rc = upload_terr(b);
    if (rc == HG_OK) rc = build_tiles(b);
    if (rc != HG_OK) return rc;
    if (flags & HG_STEP_PER_TERRAIN) {
        rc = per_terrain_steps(b, nsteps, flags, (flags & HG_STEP_PER_PASS) != 0);
        if (rc != HG_OK) return rc;
        BH(b, hgk::launch_step_advance(b->dev_step, nsteps, b->stream));
    } else {
        rc = enqueue_steps(b, nsteps, flags);
        if (rc != HG_OK) return rc;
    }
    b->step += nsteps;
    return HG_OK;
}

int hg_batch_capture(hg_batch* b, uint32_t nsteps, uint32_t flags) {
    if (!b) return HG_ERR_INVALID;
    if (nsteps == 0 || (nsteps % 2) != 0)
        return bfail(b, HG_ERR_INVALID, "hg_batch_capture: nsteps must be even and > 0 (buffers return to their start every 2 steps)");
    if (flags & ~(HG_STEP_RAIN | HG_STEP_THERMAL | HG_STEP_SMOOTH | HG_STEP_PER_TERRAIN))
        return bfail(b, HG_ERR_INVALID, "hg_batch_capture: only rain/thermal/smooth/per-terrain flags");
    Guard g(b->device);
    int rc = upload_terr(b);
    if (rc == HG_OK) rc = build_tiles(b);
    if (rc != HG_OK) return rc;
    drop_graph(b);
    const uint32_t cur0 = b->cur;
    BH(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_steps(b, nsteps, flags, (flags & HG_STEP_PER_TERRAIN) != 0);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &graph);
    b->cur = cur0;   // capture does not execute anything
    if (rc != HG_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, HG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = graph;
    BH(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_steps = nsteps;
    b->graph_cur = cur0;
    return HG_OK;
}

int hg_batch_run(hg_batch* b) {
    if (!b) return HG_ERR_INVALID;
    if (!b->exec) return bfail(b, HG_ERR_STATE, "hg_batch_run: nothing captured (or settings changed since)");
    // The captured launches name the plane buffers that were current at
    // capture time.  An odd number of eager steps since then swapped them.
    if (b->cur != b->graph_cur)
        return bfail(b, HG_ERR_STATE,
                     "hg_batch_run: the buffers were swapped by an odd number of hg_batch_step calls since "
                     "hg_batch_capture; step once more or capture again");
    Guard g(b->device);
    BH(b, hipGraphLaunch(b->exec, b->stream));
    b->step += b->graph_steps;   // buffers are back at their start: graph_steps is even
    return HG_OK;
}

int hg_batch_sync(hg_batch* b) {
    if (!b) return HG_ERR_INVALID;
    Guard g(b->device);
    BH(b, hipStreamSynchronize(b->stream));
    return HG_OK;
}

int hg_batch_get_step(const hg_batch* b, uint64_t* step) {
    if (!b || !step) return HG_ERR_INVALID;
    *step = b->step;
    return HG_OK;
}

int hg_batch_read_plane(hg_batch* b, int terrain, int plane, float* dst, size_t n_floats) {
    if (!b) return HG_ERR_INVALID;
    if (terrain < 0 || terrain >= b->count || plane < 0 || plane >= HG_NUM_STATE_PLANES || !dst ||
        n_floats != size_t(b->n[terrain]) * b->n[terrain])
        return bfail(b, HG_ERR_INVALID, "hg_batch_read_plane: bad argument");
    Guard g(b->device);
    const float* src = plane_ptr(b, terrain, plane, int((b->cur >> plane) & 1));
    BH(b, hipMemcpyAsync(dst, src, n_floats * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    return HG_OK;
}

int hg_batch_write_plane(hg_batch* b, int terrain, int plane, const float* src, size_t n_floats) {
    if (!b) return HG_ERR_INVALID;
    if (terrain < 0 || terrain >= b->count || plane < 0 || plane >= HG_NUM_STATE_PLANES || !src ||
        n_floats != size_t(b->n[terrain]) * b->n[terrain])
        return bfail(b, HG_ERR_INVALID, "hg_batch_write_plane: bad argument");
    Guard g(b->device);
    float* dst = plane_ptr(b, terrain, plane, int((b->cur >> plane) & 1));
    BH(b, hipMemcpyAsync(dst, src, n_floats * sizeof(float), hipMemcpyHostToDevice, b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    return HG_OK;
}

}  // extern "C"
