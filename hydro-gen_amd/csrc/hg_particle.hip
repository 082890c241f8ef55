// hg_particle.hip — droplet (particle) erosion.
//
// One thread per droplet; every droplet of an iteration runs concurrently and
// changes the height planes with float atomic adds (global_atomic_add_f32),
// so a droplet sees the others' changes in whatever order the memory system
// applies them.  Height reads are relaxed atomic loads.  Erosion takes from a
// cell with a returning atomic add and gives back what was not there (take()),
// so concurrent droplets never drive a layer below 0.  Mass bookkeeping: every
// amount taken from the map is added to the droplet's load and every amount
// deposited is subtracted from it; sediment carried over the map edge is
// summed in `exported`, so sum(rock + dirt) + exported is conserved up to
// FP32 rounding.
//
// Model (Beyer-style droplets; the reference's particle mode,
// glsl/particle.glsl:64-136 and glsl/particle_erosion.glsl:22-128, is a
// different model): see hg_droplet_settings in include/hg.h.
#include <algorithm>
#include <cmath>

#include "hg_kernels.hpp"

namespace hgk {

namespace {

__device__ __forceinline__ void atomic_addf(float* p, float v) { unsafeAtomicAdd(p, v); }

// Height reads are relaxed device-scope atomic loads: they are served by L2,
// where the atomic adds land, and never by a stale line of the non-coherent
// per-CU vector cache (a droplet must see its own earlier erosion).
__device__ __forceinline__ float ld(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float height(const float* R, const float* D, size_t i) { return ld(R + i) + ld(D + i); }

// Removes up to `want` from *p, never more than it holds, and returns what
// was removed.  The amount a relaxed load says is there is subtracted with a
// returning atomic add; if other droplets took from the cell in between, the
// height it replaced is smaller, and the part that was not there is added
// back.  Each take is then at most max(0, what the cell held at its atomic
// add, less the give-backs still outstanding), so the takes of all droplets
// from one cell never exceed what it held and it never ends below 0 in exact
// arithmetic; in FP32 the subtraction and the give-back round, which can
// leave a residue of a few ulp of the amounts involved below 0.
__device__ __forceinline__ float take(float* p, float want) {
    const float est = fminf(fmaxf(ld(p), 0.0f), want);
    if (!(est > 0.0f)) return 0.0f;
    const float old = atomicAdd(p, -est);
    const float got = fminf(fmaxf(old, 0.0f), est);
    if (got < est) atomic_addf(p, est - got);
    return got;
}

// Cell of a position inside [0, n-1]^2; n-1 itself maps to cell n-2 (with
// fraction 1) so the 2x2 stencil stays in the grid.
__device__ __forceinline__ int cell_of(float p, int n) { return min(int(p), n - 2); }

// Bilinear deposit of `amount` into dirt around (px, py).
__device__ __forceinline__ void deposit(float* D, int n, float px, float py, float amount) {
    const int x = cell_of(px, n), y = cell_of(py, n);
    const float u = px - float(x), v = py - float(y);
    const size_t i = size_t(y) * n + x;
    atomic_addf(D + i, amount * (1.0f - u) * (1.0f - v));
    atomic_addf(D + i + 1, amount * u * (1.0f - v));
    atomic_addf(D + i + n, amount * (1.0f - u) * v);
    atomic_addf(D + i + n + 1, amount * u * v);
}

// Start position of droplet `id` in an iteration: uniform over [0, n-1)^2.
__device__ __forceinline__ void droplet_start(uint32_t id, uint32_t iteration, const DropletParams& p, int n,
                                              float& px, float& py) {
    uint32_t o[4];
    philox4x32_10(id, iteration, 0x64726f70u, 0u, p.seed_lo, p.seed_hi, o);
    const float span = float(n - 1);
    px = u24(o[0]) * span;
    py = u24(o[1]) * span;
}

// PARTIALS = false: the workgroup's exported mass and droplet steps are added
// atomically to *exported_out and *steps_out.  PARTIALS = true: they are
// written to partials[blockIdx.x], for k_droplet_totals to sum in a fixed
// order.
template <bool PARTIALS>
__global__ __launch_bounds__(256) void k_droplets(float* __restrict__ R, float* __restrict__ D, int n,
                                                  DropletParams p, Brush brush, uint32_t iteration,
                                                  const uint32_t* __restrict__ order, double* exported_out,
                                                  unsigned long long* steps_out,
                                                  DropletPartial* partials) {
    __shared__ double s_exp[256];
    __shared__ unsigned s_steps[256];
    const uint32_t slot = blockIdx.x * 256 + threadIdx.x;
    double exported = 0.0;
    unsigned steps = 0;
    if (slot < p.count) {
        const uint32_t id = order ? order[slot] : slot;
        const float span = float(n - 1);
        float px, py;
        droplet_start(id, iteration, p, n, px, py);
        float dx = 0.0f, dy = 0.0f, speed = p.init_speed, water = p.init_water, sed = 0.0f;
        bool alive = true;
        for (uint32_t life = 0; life < p.max_lifetime; ++life) {
            ++steps;
            const int cx = cell_of(px, n), cy = cell_of(py, n);
            const float u = px - float(cx), v = py - float(cy);
            const size_t c = size_t(cy) * n + cx;
            const float h00 = height(R, D, c), h10 = height(R, D, c + 1);
            const float h01 = height(R, D, c + n), h11 = height(R, D, c + n + 1);
            const float gx = (h10 - h00) * (1.0f - v) + (h11 - h01) * v;
            const float gy = (h01 - h00) * (1.0f - u) + (h11 - h10) * u;
            const float h = h00 * (1.0f - u) * (1.0f - v) + h10 * u * (1.0f - v) + h01 * (1.0f - u) * v +
                            h11 * u * v;
            dx = dx * p.inertia - gx * (1.0f - p.inertia);
            dy = dy * p.inertia - gy * (1.0f - p.inertia);
            // correctly rounded sqrt and divide (as on the host): a droplet's
            // path is sensitive to the last bit of its direction.  Plain `/`
            // and sqrtf lower to the IEEE sequences; HIP's __fsqrt_rn lowers to
            // the native 3-ulp sqrt.
            const float len = sqrtf(dx * dx + dy * dy);
            if (len > 0.0f) {
                dx = dx / len;
                dy = dy / len;
            }
            const float ox = px, oy = py;
            if (dx == 0.0f && dy == 0.0f) {       // not moving: drop the load here
                deposit(D, n, px, py, sed);
                sed = 0.0f;
                alive = false;
                break;
            }
            px += dx;
            py += dy;
            if (!(px >= 0.0f && py >= 0.0f && px < span && py < span)) {   // over the edge (or NaN)
                exported += double(sed);
                sed = 0.0f;
                alive = false;
                break;
            }
            const int nx2 = cell_of(px, n), ny2 = cell_of(py, n);
            const float u2 = px - float(nx2), v2 = py - float(ny2);
            const size_t c2 = size_t(ny2) * n + nx2;
            const float hn = height(R, D, c2) * (1.0f - u2) * (1.0f - v2) + height(R, D, c2 + 1) * u2 * (1.0f - v2) +
                             height(R, D, c2 + n) * (1.0f - u2) * v2 + height(R, D, c2 + n + 1) * u2 * v2;
            const float dh = hn - h;
            const float cap = fmaxf(-dh, p.min_slope) * speed * water * p.capacity;
            if (sed > cap || dh > 0.0f) {
                const float amount = dh > 0.0f ? fminf(dh, sed) : (sed - cap) * p.deposition;
                sed -= amount;
                deposit(D, n, ox, oy, amount);
            } else {
                const float amount = fminf((cap - sed) * p.erosion, -dh);
                const bool interior = cx >= brush.r && cy >= brush.r && cx < n - brush.r && cy < n - brush.r;
                float wsum = 1.0f;
                if (!interior) {
                    wsum = 0.0f;
                    for (int t = 0; t < brush.count; ++t) {
                        const int x = cx + brush.dx[t], y = cy + brush.dy[t];
                        if (x >= 0 && y >= 0 && x < n && y < n) wsum += brush.w[t];
                    }
                }
                const float scale = interior ? amount : amount / wsum;
                for (int t = 0; t < brush.count; ++t) {
                    const int x = cx + brush.dx[t], y = cy + brush.dy[t];
                    if (!interior && !(x >= 0 && y >= 0 && x < n && y < n)) continue;
                    const size_t i = size_t(y) * n + x;
                    const float want = scale * brush.w[t];
                    const float td = take(D + i, want);
                    const float tr = take(R + i, want - td);
                    sed += td + tr;
                }
            }
            speed = sqrtf(fmaxf(0.0f, speed * speed - dh * p.gravity));
            water = water * (1.0f - p.evaporation);
        }
        if (alive) deposit(D, n, px, py, sed);
    }
    s_exp[threadIdx.x] = exported;
    s_steps[threadIdx.x] = steps;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_exp[threadIdx.x] += s_exp[threadIdx.x + s];
            s_steps[threadIdx.x] += s_steps[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if constexpr (PARTIALS) {
            partials[blockIdx.x].exported = s_exp[0];
            partials[blockIdx.x].steps = s_steps[0];
        } else {
            if (s_exp[0] != 0.0) unsafeAtomicAdd(exported_out, s_exp[0]);
            atomicAdd(steps_out, (unsigned long long)s_steps[0]);
        }
    }
}

// ---- droplet ordering (counting sort by start tile) --------------------------

__device__ __forceinline__ uint32_t start_tile(uint32_t id, uint32_t iteration, const DropletParams& p, int n,
                                               int side, int ntx) {
    float px, py;
    droplet_start(id, iteration, p, n, px, py);
    return uint32_t(int(py) / side) * uint32_t(ntx) + uint32_t(int(px) / side);
}

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* a, uint32_t count) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) a[i] = 0u;
}

__global__ __launch_bounds__(256) void k_drop_count(DropletParams p, int n, int side, int ntx, uint32_t iteration,
                                                    uint32_t* tiles) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id < p.count) atomicAdd(&tiles[start_tile(id, iteration, p, n, side, ntx)], 1u);
}

// Exclusive prefix sum of tiles[0, count) in place, one workgroup of 1024
// threads: each thread sums a contiguous run, the run totals are scanned in
// LDS, then each run is rewritten as offsets.
__global__ __launch_bounds__(1024) void k_drop_scan(uint32_t* tiles, uint32_t count) {
    __shared__ uint32_t s[1024];
    const uint32_t per = (count + 1023) / 1024;
    const uint32_t b = threadIdx.x * per, e = min(b + per, count);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += tiles[i];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {          // inclusive Hillis-Steele scan
        const uint32_t v = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - sum;                      // exclusive
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t c = tiles[i];
        tiles[i] = run;
        run += c;
    }
}

// Places droplet id at its sorted position (its tile's next free offset).
__global__ __launch_bounds__(256) void k_drop_scatter(DropletParams p, int n, int side, int ntx, uint32_t iteration,
                                                      uint32_t* tiles, uint32_t* order) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id >= p.count) return;
    order[atomicAdd(&tiles[start_tile(id, iteration, p, n, side, ntx)], 1u)] = id;
}

// Sums partials[0, nblocks) in a fixed order (thread t: blocks t, t + 256, ...;
// then a fixed tree) into *exported_out and *steps_out (either may be null),
// replacing them, or adding to them with `accumulate`.
__global__ __launch_bounds__(256) void k_droplet_totals(const DropletPartial* __restrict__ partials,
                                                        uint32_t nblocks, double* exported_out,
                                                        uint64_t* steps_out, int accumulate) {
    __shared__ double s_exp[256];
    __shared__ unsigned long long s_steps[256];
    double e = 0.0;
    unsigned long long st = 0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256) {
        e += partials[b].exported;
        st += partials[b].steps;
    }
    s_exp[threadIdx.x] = e;
    s_steps[threadIdx.x] = st;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            s_exp[threadIdx.x] += s_exp[threadIdx.x + h];
            s_steps[threadIdx.x] += s_steps[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (exported_out) *exported_out = (accumulate ? *exported_out : 0.0) + s_exp[0];
        if (steps_out) *steps_out = (accumulate ? *steps_out : 0ull) + uint64_t(s_steps[0]);
    }
}

}  // namespace

Brush make_brush(int r) {
    // weights 1 - dist/r over cells with dist < r, normalised in double and
    // rounded to float once (the oracle builds the same table)
    Brush b{};
    b.r = r;
    double raw[HG_BRUSH_MAX];
    double sum = 0;
    for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx) {
            const double d = std::sqrt(double(dx * dx + dy * dy));
            if (d < r && b.count < HG_BRUSH_MAX) {
                b.dx[b.count] = int8_t(dx);
                b.dy[b.count] = int8_t(dy);
                raw[b.count] = 1.0 - d / r;
                sum += raw[b.count];
                ++b.count;
            }
        }
    for (int t = 0; t < b.count; ++t) b.w[t] = float(raw[t] / sum);
    return b;
}

hipError_t launch_droplets(float* rock, float* dirt, int n, const DropletParams& p, const Brush& brush,
                           uint32_t iteration, const uint32_t* order, double* exported,
                           unsigned long long* steps, hipStream_t s) {
    if (p.count == 0) return hipSuccess;
    k_droplets<false><<<(p.count + 255) / 256, 256, 0, s>>>(rock, dirt, n, p, brush, iteration, order, exported,
                                                             steps, nullptr);
    return hipGetLastError();
}

int droplet_tile_side(int n) { return std::max(16, (n + 127) / 128); }

int droplet_tile_count(int n) {
    const int side = droplet_tile_side(n), nt = (n + side - 1) / side;
    return nt * nt;
}

hipError_t launch_droplet_order(int n, const DropletParams& p, uint32_t iteration, uint32_t* tiles,
                                uint32_t* order, hipStream_t s) {
    if (p.count == 0) return hipSuccess;
    const int side = droplet_tile_side(n), ntx = (n + side - 1) / side;
    const uint32_t nt = uint32_t(ntx) * uint32_t(ntx), g = (p.count + 255) / 256;
    k_zero_u32<<<(nt + 255) / 256, 256, 0, s>>>(tiles, nt);
    k_drop_count<<<g, 256, 0, s>>>(p, n, side, ntx, iteration, tiles);
    k_drop_scan<<<1, 1024, 0, s>>>(tiles, nt);
    k_drop_scatter<<<g, 256, 0, s>>>(p, n, side, ntx, iteration, tiles, order);
    return hipGetLastError();
}

hipError_t launch_droplets_partials(float* rock, float* dirt, int n, const DropletParams& p, const Brush& brush,
                                    uint32_t iteration, DropletPartial* partials, hipStream_t s) {
    if (p.count == 0) return hipSuccess;
    k_droplets<true><<<(p.count + 255) / 256, 256, 0, s>>>(rock, dirt, n, p, brush, iteration, nullptr, nullptr,
                                                            nullptr, partials);
    return hipGetLastError();
}

hipError_t launch_droplet_totals(const DropletPartial* partials, uint32_t nblocks, double* exported,
                                 uint64_t* steps, bool accumulate, hipStream_t s) {
    k_droplet_totals<<<1, 256, 0, s>>>(partials, nblocks, exported, steps, accumulate ? 1 : 0);
    return hipGetLastError();
}

}  // namespace hgk
