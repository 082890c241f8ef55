// hg_refparticle.hip — the reference's own particle mode.
//
// One Erosion::dispatch_particle (src/erosion.cpp:132-156) is four launches:
//  * k_refp_move: one thread per particle, glsl/particle.glsl:64-136 (spawn,
//    terrain normal and momentum sampled bilinearly, velocity and volume
//    update, capacity, kill tests).  The spawn range and the border test
//    use Map_settings_data.hmap_dims, which the reference leaves at
//    (1024, 1024) whatever the map size; here it is the grid's n unless
//    hg_set_ref_particle_dims sets it.  Particles outside the grid read 0
//    and their corner updates are dropped, like imageLoad/imageStore
//    outside an image;
//  * k_refp_erode: one thread per particle, glsl/particle_erosion.glsl:22-128,
//    the 4 corners of the particle's quad in the shader's order;
//  * the fused thermal kernel (both thermal layers and the grid smoothing of
//    rock and dirt), as for the grid step;
//  * k_refp_smooth_extra: the particle branch of glsl/smoothing.glsl:76-95
//    (momentum map and water display), per cell.
//
// The shader serialises the corner updates at a pixel with a per-pixel
// spinlock (glsl/particle_erosion.glsl:78-90).  Here every update of a pixel
// is an atomic operation of its own instead: the pure additions (deposits,
// the water display, the momentum accumulation) are float atomic adds, and a
// layer's dissolve, whose clamp at 0 reads the height, is one compare-and-swap
// loop on that layer.  The outcome is that of some serial order of the
// updates, as with the lock, except that a corner's dirt and rock updates are
// not one transaction: another particle can touch the pixel's rock between
// them.  A CAS loop that keeps failing gives up after a fixed number of tries
// and adds without the clamp (counted; the host reads the count).
//
// Arithmetic is FP32 in the shaders' order; division and square root are the
// correctly rounded ones and mix(a, b, t) is a + t (b - a), so that one
// particle on its own reproduces the FP32 oracle bit for bit.  The spawn
// hash's sin is hash_sinf below, whose operations a host compiles alike, so
// spawns reproduce it too.
#include <cmath>

#include "hg_kernels.hpp"

namespace hgk {

namespace {

constexpr int kCasTries = 4096;

__device__ __forceinline__ float fract_(float x) { return x - floorf(x); }

// sin of an FP32 argument for the spawn hash, evaluated in FP64 and rounded
// once: x = k pi/2 + r with pi/2 split into three parts (two of 24
// significant bits, so k * part is exact for |k| < 2^29, i.e. |x| < 8e8), the
// Taylor polynomials of sin r and cos r on |r| <= pi/4 to degree 15 and 16,
// and the quadrant of k.  The hash multiplies sin by 1e4 before taking the
// fraction, so any last-bit difference moves a spawn; GLSL leaves sin's
// precision to the implementation, and this one runs the same operations on
// the device and on a host (no contraction: -ffp-contract=off).
__device__ __forceinline__ float hash_sinf(float xf) {
    const double x = double(xf);
    const double k = rint(x * 0.6366197723675814);
    const double r = ((x - k * 1.570796251296997) - k * 7.549789415861596e-08) - k * 5.390302858158119e-15;
    const double z = r * r;
    const double sn = r * (1.0 + z * (-1.0 / 6.0 + z * (1.0 / 120.0 + z * (-1.0 / 5040.0 + z * (1.0 / 362880.0 +
                      z * (-1.0 / 39916800.0 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0))))))));
    const double cs = 1.0 + z * (-0.5 + z * (1.0 / 24.0 + z * (-1.0 / 720.0 + z * (1.0 / 40320.0 +
                      z * (-1.0 / 3628800.0 + z * (1.0 / 479001600.0 + z * (-1.0 / 87178291200.0 +
                      z * (1.0 / 20922789888000.0))))))));
    const int q = int(static_cast<long long>(k) & 3);
    const double v = q == 0 ? sn : q == 1 ? cs : q == 2 ? -sn : -cs;
    return float(v);
}

// rand(vec2 p), glsl/particle.glsl:41-44
__device__ __forceinline__ float rp_rand(float x, float y) {
    return fract_(1e4f * hash_sinf(17.0f * x + y * 0.1f) * (0.1f + fabsf(hash_sinf(y * 13.0f + x))));
}

// img_bilinear (glsl/img_interpolation.glsl:3-22) of one plane; texels
// outside the image read 0.
__device__ __forceinline__ float rp_tex(const float* p, int n, int x, int y) {
    return (x >= 0 && y >= 0 && x < n && y < n) ? p[size_t(y) * n + x] : 0.0f;
}
__device__ __forceinline__ float rp_bilinear(const float* p, int n, float x, float y) {
    const int ix = int(x), iy = int(y);
    const float fx = fract_(x), fy = fract_(y);
    const float a = rp_tex(p, n, ix, iy), b = rp_tex(p, n, ix + 1, iy);
    const float c = rp_tex(p, n, ix, iy + 1), d = rp_tex(p, n, ix + 1, iy + 1);
    const float v1 = a + fx * (b - a);
    const float v2 = c + fx * (d - c);
    return v1 + fy * (v2 - v1);
}

// Correctly rounded division and square root: plain `/` and sqrtf lower to
// the IEEE sequences (HIP's __fsqrt_rn lowers to the native 3-ulp sqrt).
__device__ __forceinline__ float fdiv(float a, float b) { return a / b; }
__device__ __forceinline__ float fsqrt(float a) { return sqrtf(a); }

__global__ __launch_bounds__(256) void k_refp_move(hg_ref_particle* __restrict__ parts, uint32_t count,
                                                   const float* __restrict__ R, const float* __restrict__ D,
                                                   const float* __restrict__ B, const float* __restrict__ MX,
                                                   const float* __restrict__ MY, int n, RefParticleParams k,
                                                   float time, int should_rain) {
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= count) return;
    hg_ref_particle p = parts[id];
    if (p.iters == 0 && !should_rain) return;
    for (int i = 0; i < 2; ++i)
        if (p.sediment[i] < 0.0f || p.iters == 0) p.sediment[i] = 0.0f;
    if (p.iters == 0 || p.to_kill) {                                             // :71-86
        const float ax = fract_(time * 1.37f) * 1000.0f, ay = fract_(time * 7.21f) * 1000.0f;
        p.position[0] = fdiv(rp_rand(ax, float(id)) * float(k.dims_x - 4.0f), 1.0f) + 2.0f;
        p.position[1] = fdiv(rp_rand(ay, float(id) + 3.14f) * float(k.dims_y - 4.0f), 1.0f) + 2.0f;
        p.to_kill = 0u;
        p.velocity[0] = 0.0f;
        p.velocity[1] = 0.0f;
        p.volume = k.init_volume;
        if (!should_rain) return;          // p.iters = 0; return: nothing is written back
        p.iters = 1;
    }
    const float px = p.position[0], py = p.position[1];
    // get_terr_normal (:51-63)
    const float rr = rp_bilinear(R, n, px + 1.0f, py), rg = rp_bilinear(D, n, px + 1.0f, py);
    const float lr = rp_bilinear(R, n, px - 1.0f, py), lg = rp_bilinear(D, n, px - 1.0f, py);
    const float br = rp_bilinear(R, n, px, py - 1.0f), bg = rp_bilinear(D, n, px, py - 1.0f);
    const float tr = rp_bilinear(R, n, px, py + 1.0f), tg = rp_bilinear(D, n, px, py + 1.0f);
    const float dx = rr + rg - lr - lg;
    const float dz = tr + tg - br - bg;
    const float cx = dx * 2.0f - 0.0f * dz, cy = 0.0f * 0.0f - 2.0f * 2.0f, cz = 2.0f * dz - dx * 0.0f;
    const float cl = fsqrt(cx * cx + cy * cy + cz * cz);
    const float nx = fdiv(cx, cl), ny = fdiv(cy, cl), nz = fdiv(cz, cl);
    const float mx = rp_bilinear(MX, n, px, py), my = rp_bilinear(MY, n, px, py);
    const float water = rp_bilinear(B, n, px, py);
    float vx = p.velocity[0], vy = p.velocity[1];
    vx = vx - fdiv(k.dt * nx, p.volume) * k.G;                                   // :101
    vy = vy - fdiv(k.dt * nz, p.volume) * k.G;
    const float mlen = fsqrt(mx * mx + my * my);
    float vlen = fsqrt(vx * vx + vy * vy);
    if (mlen > 0.0f && vlen > 0.0f) {                                            // :103-105
        const float d = fdiv(mx, mlen) * fdiv(vx, vlen) + fdiv(my, mlen) * fdiv(vy, vlen);
        const float t = fdiv(k.inertia * d, p.volume + 1e5f * water);
        vx = vx + t * mx;
        vy = vy + t * my;
    }
    vlen = fsqrt(vx * vx + vy * vy);
    if (vlen > 1.0f) {                                                           // :108-110
        vx = fdiv(vx, vlen);
        vy = fdiv(vy, vlen);
    }
    float qx = px + k.dt * vx, qy = py + k.dt * vy;                              // :113
    if (qx <= 1.0f || qy <= 1.0f || qx * 1.0f >= float(k.dims_x - 2) || qy * 1.0f >= float(k.dims_y - 2)) {   // :114-124
        qx = px;
        qy = py;
        vx = 0.0f;
        vy = 0.0f;
        p.to_kill = 1u;
    }
    const float fr = 1.0f - k.dt * k.friction * ny;                              // :125
    vx = vx * fr;
    vy = vy * fr;
    p.volume = p.volume - k.dt * k.Ke;                                           // :126
    const float sin_a = fsqrt(1.0f - ny * ny);                                   // :129
    vlen = fsqrt(vx * vx + vy * vy);
    const float m02 = 0.02f < sin_a ? sin_a : 0.02f;                             // GLSL max(0.02, sin_a)
    const float sc = k.Kc * p.volume * vlen * m02;
    p.sc = 0.0f < sc ? sc : 0.0f;                                                // :131
    p.iters += 1;
    if (p.volume <= k.min_volume || vlen < k.min_velocity || uint32_t(p.iters) >= k.ttl) p.to_kill = 1u;
    p.position[0] = qx;
    p.position[1] = qy;
    p.velocity[0] = vx;
    p.velocity[1] = vy;
    parts[id] = p;
}

__device__ __forceinline__ void atomic_addf(float* p, float v) { unsafeAtomicAdd(p, v); }

// h' = max-with-clamp of h - eroded (h' = 0 when h - eroded < 0) as one atomic
// update of *p.  Returns the height it replaced; `t` = that height - eroded.
__device__ float dissolve(float* p, float eroded, float& t, unsigned long long* giveups) {
    unsigned int* u = reinterpret_cast<unsigned int*>(p);
    unsigned int old = __hip_atomic_load(u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int tries = 0; tries < kCasTries; ++tries) {
        const float h = __uint_as_float(old);
        const float tt = h - eroded;
        const float nv = tt < 0.0f ? 0.0f : tt;
        unsigned int expected = old;
        if (__hip_atomic_compare_exchange_strong(u, &expected, __float_as_uint(nv), __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            t = tt;
            return h;
        }
        old = expected;
    }
    const float h = atomicAdd(p, -eroded);
    t = h - eroded;
    atomicAdd(giveups, 1ull);
    return h;
}

__global__ __launch_bounds__(256) void k_refp_erode(hg_ref_particle* __restrict__ parts, uint32_t count,
                                                    float* R, float* D, float* B, float* MZ, float* MW, int n,
                                                    RefParticleParams k, unsigned long long* giveups) {
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= count) return;
    hg_ref_particle part = parts[id];
    if (part.iters == 0) return;
    const float px = part.position[0] * 1.0f, py = part.position[1] * 1.0f;
    const int bx = int(px), by = int(py);
    const int qx[4] = {bx, bx + 1, bx + 1, bx}, qy[4] = {by, by, by + 1, by + 1};
    const float fx = fract_(px), fy = fract_(py);
    const float ox[4] = {1.0f - fx, fx, fx, 1.0f - fx}, oy[4] = {1.0f - fy, 1.0f - fy, fy, fy};
    const float old_sed[2] = {part.sediment[0], part.sediment[1]};
    float* layer[2] = {R, D};
#pragma unroll 1
    for (int c = 0; c < 4; ++c) {
        if (qx[c] < 0 || qy[c] < 0 || qx[c] >= n || qy[c] >= n) continue;   // imageStore outside: dropped
        const size_t i = size_t(qy[c]) * n + qx[c];
        const float multipl = ox[c] * oy[c];
        float cap = 0.0f;
#pragma unroll
        for (int l = 1; l >= 0; --l) {
            float* h = layer[l] + i;
            if (part.to_kill) {                                                  // :36-42
                const float sed = old_sed[l] * multipl;
                atomic_addf(h, sed);
                part.sediment[l] = part.sediment[l] - sed;
                continue;
            }
            const float Kls = k.dt * k.Ks[l], Kld = k.dt * k.Kd[l];
            const float cs = part.sc - cap;
            const float cc = 0.0f < cs ? cs : 0.0f;
            float s1 = old_sed[l];
            if (cc > s1) {                                                       // :52-63
                const float eroded = multipl * Kls * (cc - s1);
                s1 = s1 + eroded;
                float t;
                const float old_terr = dissolve(h, eroded, t, giveups);
                if (t < 0.0f) {
                    s1 = s1 + t;
                    cap = cap + old_terr;
                } else {
                    part.sediment[l] = s1;
                    break;
                }
            } else {                                                             // :65-69
                const float deposit = multipl * Kld * (s1 - cc);
                s1 = s1 - deposit;
                atomic_addf(h, deposit);
            }
            part.sediment[l] = s1;
        }
        const float conv = part.sediment[0] * k.Kconv * k.dt;                   // :71-75
        part.sediment[1] = part.sediment[1] + conv;
        part.sediment[0] = part.sediment[0] - conv;
        atomic_addf(B + i, 1e-5f * part.volume * multipl);                       // :79
        atomic_addf(MZ + i, part.volume * part.velocity[0] * multipl);          // :81
        atomic_addf(MW + i, part.volume * part.velocity[1] * multipl);
    }
    parts[id] = part;
}

// glsl/smoothing.glsl:76-95 on the interior cells (border cells return before it).
__global__ __launch_bounds__(256) void k_refp_smooth_extra(float* B, float* MX, float* MY, float* MZ, float* MW,
                                                           int n, float cnt) {
    const size_t cells = size_t(n) * n;
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= cells) return;
    const int y = int(i / size_t(n)), x = int(i - size_t(y) * n);
    if (x == 0 || y == 0 || x == n - 1 || y == n - 1) return;
    const float fm0 = 1.0f - (1e-12f * cnt), fb0 = 1.0f - (8e-8f * cnt);
    const float fm = fminf(fmaxf(fm0, 0.0f), 1.0f), fb = fminf(fmaxf(fb0, 0.0f), 1.0f);
    float mx = MX[i] * fm, my = MY[i] * fm;
    mx = mx + (1e-12f * cnt) * MZ[i];
    my = my + (1e-12f * cnt) * MW[i];
    MZ[i] = 0.0f;
    MW[i] = 0.0f;
    float b = B[i] * fb;
    if (b < 1e-6f) b = 0.0f;
    if (fsqrt(mx * mx + my * my) < 1e-12f) {
        mx = 0.0f;
        my = 0.0f;
    }
    MX[i] = mx;
    MY[i] = my;
    B[i] = b;
}

}  // namespace

static_assert(sizeof(hg_ref_particle) == 48, "Particle is 48 bytes in std430");

hipError_t launch_ref_particle_move(hg_ref_particle* parts, uint32_t count, const float* rock, const float* dirt,
                                    const float* water, const float* mx, const float* my, int n,
                                    const RefParticleParams& k, float time, int should_rain, hipStream_t s) {
    if (count == 0) return hipSuccess;
    k_refp_move<<<(count + 255) / 256, 256, 0, s>>>(parts, count, rock, dirt, water, mx, my, n, k, time, should_rain);
    return hipGetLastError();
}

hipError_t launch_ref_particle_erode(hg_ref_particle* parts, uint32_t count, float* rock, float* dirt, float* water,
                                     float* mz, float* mw, int n, const RefParticleParams& k,
                                     unsigned long long* giveups, hipStream_t s) {
    if (count == 0) return hipSuccess;
    k_refp_erode<<<(count + 255) / 256, 256, 0, s>>>(parts, count, rock, dirt, water, mz, mw, n, k, giveups);
    return hipGetLastError();
}

hipError_t launch_ref_particle_smooth_extra(float* water, float* mx, float* my, float* mz, float* mw, int n,
                                            uint32_t particle_count, hipStream_t s) {
    const size_t cells = size_t(n) * n;
    k_refp_smooth_extra<<<unsigned((cells + 255) / 256), 256, 0, s>>>(water, mx, my, mz, mw, n,
                                                                        float(particle_count));
    return hipGetLastError();
}

}  // namespace hgk
