// hg_init.hip — simplex-fBm heightmap initialisation: seeded-permutation
// 2-D simplex fBm, min-max rescaled into the rock layer.  This replaces the
// reference's World::gen_heightmap dispatch (src/state.cpp:116-147) for
// terrains defined as "simplex fBm with a seeded permutation".
//
// Noise definition (shared with the CPU oracle, written independently there):
//  * permutation of 0..255 from a Philox4x32-10 driven Fisher-Yates shuffle
//    (built on the host, passed by value);
//  * Perlin's 2-D simplex lattice: skew (x+y)*(sqrt3-1)/2, unskew
//    (i+j)*(3-sqrt3)/6, corner kernel (0.5-r^2)^4 times the dot product with
//    one of 8 gradients (+-1,+-1), (+-1,0), (0,+-1) picked by
//    perm[(i + perm[j]) & 255] & 7, total scaled by 70;
//  * fBm = sum over octaves o of gain^o * simplex(x*f_o, y*f_o),
//    f_o = base_freq * lacunarity^o.
#include "hg_kernels.hpp"

namespace hgk {

namespace {

struct Perm {
    uint8_t p[256];
};

__device__ __forceinline__ float grad_dot(int h, float x, float y) {
    switch (h & 7) {
        case 0: return x + y;
        case 1: return -x + y;
        case 2: return x - y;
        case 3: return -x - y;
        case 4: return x;
        case 5: return -x;
        case 6: return y;
        default: return -y;
    }
}

__device__ __forceinline__ float corner(const Perm& P, int gi, int gj, float x, float y) {
    const float t = 0.5f - x * x - y * y;
    if (!(t > 0.0f)) return 0.0f;
    const int h = P.p[(gi + P.p[gj & 255]) & 255];
    const float t2 = t * t;
    return t2 * t2 * grad_dot(h, x, y);
}

__device__ float simplex2(const Perm& P, float x, float y) {
    const float F2 = 0.36602540378443864676f;
    const float G2 = 0.21132486540518711775f;
    const float s = (x + y) * F2;
    const float fi = floorf(x + s), fj = floorf(y + s);
    const float t = (fi + fj) * G2;
    const float x0 = x - (fi - t), y0 = y - (fj - t);
    const int i1 = x0 > y0 ? 1 : 0, j1 = 1 - i1;
    const float x1 = x0 - float(i1) + G2, y1 = y0 - float(j1) + G2;
    const float x2 = x0 - 1.0f + 2.0f * G2, y2 = y0 - 1.0f + 2.0f * G2;
    const int i = int(fi) & 255, j = int(fj) & 255;
    const float n = corner(P, i, j, x0, y0) + corner(P, i + i1, j + j1, x1, y1) +
                    corner(P, i + 1, j + 1, x2, y2);
    return 70.0f * n;
}

__global__ __launch_bounds__(256) void k_fbm(float* out, int nx, int ny, Perm P, int octaves,
                                             float lacunarity, float gain, float base_freq) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= nx || y >= ny) return;
    float sum = 0.0f, amp = 1.0f, f = base_freq;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * simplex2(P, float(x) * f, float(y) * f);
        amp *= gain;
        f *= lacunarity;
    }
    out[size_t(y) * nx + x] = sum;
}

// Order-preserving map of a float to an unsigned int.
__device__ __forceinline__ unsigned ord(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// mm[0] = ord(min), mm[1] = ord(max); mm must hold {0xffffffff, 0} on entry.
__global__ __launch_bounds__(256) void k_minmax(const float* v, size_t n, unsigned* mm) {
    __shared__ unsigned smin[256], smax[256];
    unsigned lo = 0xffffffffu, hi = 0u;
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
        const unsigned o = ord(v[i]);
        lo = min(lo, o);
        hi = max(hi, o);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = min(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMin(&mm[0], smin[0]);
        atomicMax(&mm[1], smax[0]);
    }
}

__global__ __launch_bounds__(256) void k_rescale(float* v, size_t n, const unsigned* mm, float lo,
                                                 float hi) {
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const float mn = unord(mm[0]), mx = unord(mm[1]);
    const float span = mx - mn;
    v[i] = span > 0.0f ? lo + (hi - lo) * ((v[i] - mn) / span) : lo;
}

__global__ __launch_bounds__(256) void k_fill(float* p, size_t n, float value) {
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = value;
}

__global__ void k_mm_reset(unsigned* mm) {
    if (threadIdx.x == 0) { mm[0] = 0xffffffffu; mm[1] = 0u; }
}

// ---------------------------------------------------------------------------
// Reference heightmap (glsl/heightmap.glsl:52-158) and its noise
// (glsl/simplex_noise.glsl:579-676): the Hugo Elias integer hash, gradient
// noise with quintic interpolation and analytic derivatives (noised), its fBm
// (perlfbm) and the derivative-damped fBm (erosion_perlfbm).  Every operation
// follows the shader's order in FP32 (int32 arithmetic wraps), and exp() is
// gen_expf below, whose operations a host compiles alike, so a host FP32
// restatement can agree bit for bit.
// ---------------------------------------------------------------------------

// exp(x) in FP32 for the exp mask (glsl/heightmap.glsl:43-44): Cody and
// Waite's argument reduction x = k ln2 + r, with ln2 split into a 12-bit head
// (k * head is exact for |k| < 2^12) and a tail, so |r| <= ln2 / 2 + a few
// ulp; exp(r) by its degree-7 Taylor polynomial in Horner form (the first
// omitted term, r^8 / 8!, is below 2^-27); then the exact scaling by 2^k.
// Within 2 ulp of exp over the normal range.  GLSL leaves exp's precision to
// the implementation; this one is written so that the same operations run on
// the device and on a host (no contraction: -ffp-contract=off).
__device__ __forceinline__ float gen_expf(float x) {
    if (!(x < 88.5f)) return x != x ? x : __builtin_huge_valf();   // NaN, or overflow
    if (x < -87.0f) return 0.0f;                                   // below the normal range
    const float k = rintf(x * 1.44269504088896341f);
    const float r = (x - k * 0.693145751953125f) - k * 1.42860682030941723e-6f;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return ldexpf(p, int(k));
}
struct MapParams {
    float max_height, max_dirt, height_mult, seed, persistance, lacunarity, scale, uplift_scale,
        domain_warp_scale, terrace_scale;
    int octaves, domain_warp, terrace;
    unsigned mask_round, mask_exp, mask_power, mask_slope, uplift;
};

__device__ __forceinline__ void ehash(int px, int py, float& hx, float& hy) {
    uint32_t nx = uint32_t(px) * 3u + uint32_t(py) * 311u;       // 2D -> 1D
    uint32_t ny = uint32_t(px) * 37u + uint32_t(py) * 113u;
    nx = (nx << 13) ^ nx;
    ny = (ny << 13) ^ ny;
    nx = nx * (nx * nx * 15731u + 789221u) + 1376312589u;        // Hugo Elias
    ny = ny * (ny * ny * 15731u + 789221u) + 1376312589u;
    const float den = float(0x0fffffff);
    hx = -1.0f + (2.0f * float(int(nx & 0x0fffffffu))) / den;
    hy = -1.0f + (2.0f * float(int(ny & 0x0fffffffu))) / den;
}

// value, d/dx, d/dy
__device__ __forceinline__ void noised(float px, float py, float& v, float& dx, float& dy) {
    const float flx = floorf(px), fly = floorf(py);
    const int ix = int(flx), iy = int(fly);
    const float fx = px - flx, fy = py - fly;
    const float ux = ((fx * fx) * fx) * (fx * (fx * 6.0f - 15.0f) + 10.0f);
    const float uy = ((fy * fy) * fy) * (fy * (fy * 6.0f - 15.0f) + 10.0f);
    const float dux = ((30.0f * fx) * fx) * (fx * (fx - 2.0f) + 1.0f);
    const float duy = ((30.0f * fy) * fy) * (fy * (fy - 2.0f) + 1.0f);
    float gax, gay, gbx, gby, gcx, gcy, gdx, gdy;
    ehash(ix, iy, gax, gay);
    ehash(ix + 1, iy, gbx, gby);
    ehash(ix, iy + 1, gcx, gcy);
    ehash(ix + 1, iy + 1, gdx, gdy);
    const float va = gax * fx + gay * fy;
    const float vb = gbx * (fx - 1.0f) + gby * fy;
    const float vc = gcx * fx + gcy * (fy - 1.0f);
    const float vd = gdx * (fx - 1.0f) + gdy * (fy - 1.0f);
    const float k4 = ((va - vb) - vc) + vd;
    v = ((va + ux * (vb - va)) + uy * (vc - va)) + (ux * uy) * k4;
    dx = (((gax + ux * (gbx - gax)) + uy * (gcx - gax)) + (ux * uy) * (((gax - gbx) - gcx) + gdx)) +
         dux * ((uy * k4 + vb) - va);
    dy = (((gay + ux * (gby - gay)) + uy * (gcy - gay)) + (ux * uy) * (((gay - gby) - gcy) + gdy)) +
         duy * ((ux * k4 + vc) - va);
}

__device__ float perlfbm(float x, float y, const MapParams& m) {
    x = x + m.seed * 100.0f;
    y = y + m.seed * 100.0f;
    float result = 0.0f, amplitude = 1.0f, frequency = 1.0f, maximum = amplitude;
    for (int o = 0; o < m.octaves; ++o) {
        float v, dx, dy;
        noised((x * frequency) * m.scale, (y * frequency) * m.scale, v, dx, dy);
        result = result + ((v + 1.0f) / 2.0f) * amplitude;
        frequency = frequency * m.lacunarity;
        amplitude = amplitude * m.persistance;
        maximum = maximum + amplitude;
    }
    return result / maximum;
}

__device__ float erosion_perlfbm(float x, float y, const MapParams& m) {
    x = x + m.seed * 100.0f;
    y = y + m.seed * 100.0f;
    float result = 0.0f, amplitude = 1.0f, frequency = 1.5f, maximum = amplitude;
    float sx = 0.0f, sy = 0.0f;
    for (int o = 0; o < m.octaves; ++o) {
        float v, dx, dy;
        noised((x * frequency) * m.scale, (y * frequency) * m.scale, v, dx, dy);
        sx = sx + dx;
        sy = sy + dy;
        const float nv = (v + 1.0f) / 2.0f;
        result = result + (nv * amplitude) / (1.0f + (sx * sx + sy * sy));
        frequency = frequency * m.lacunarity;
        amplitude = amplitude * m.persistance;
        maximum = maximum + amplitude;
    }
    return result / maximum;
}

// gln_sfbm with turbulence and ridge (the uplift field)
__device__ float ridged_sfbm(float x, float y, const MapParams& m) {
    x = x + m.seed * 100.0f;
    y = y + m.seed * 100.0f;
    const float scale = m.scale / m.uplift_scale;
    float result = 0.0f, amplitude = 1.0f, frequency = 1.0f, maximum = amplitude;
    for (int o = 0; o < m.octaves; ++o) {
        float nv = ref_simplex((x * frequency) * scale, (y * frequency) * scale);
        nv = 1.0f - fabsf(nv);
        result = result + nv * amplitude;
        frequency = frequency * m.lacunarity;
        amplitude = amplitude * m.persistance;
        maximum = maximum + amplitude;
    }
    return result / maximum;
}

__global__ __launch_bounds__(256) void k_gen_heightmap(float* rock, float* dirt, int n, MapParams m) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= n || y >= n) return;
    const float fx = float(x), fy = float(y);
    const float uvx = fx / float(n), uvy = fy / float(n);
    float distx = 1.0f, disty = 1.0f;
    if (m.domain_warp != 0) {
        distx = perlfbm(fx + 2.3f, fy + 2.9f, m);
        disty = perlfbm(fx - 3.1f, fy - 4.3f, m);
        if (m.domain_warp == 2) {
            const float wx = fx + m.domain_warp_scale * distx, wy = fy + m.domain_warp_scale * disty;
            distx = perlfbm(wx - 5.7f, wy + 27.9f, m);
            disty = perlfbm(wx + 11.5f, wy - 23.7f, m);
        }
    }
    float val = erosion_perlfbm(fx + m.domain_warp_scale * distx, fy + m.domain_warp_scale * disty, m);
    float hm = m.height_mult;
    if (m.uplift != 0) {
        const float up = ridged_sfbm(fx - 7.3f, fy + 19.9f, m);
        if (m.mask_exp != 0) hm += 2.5f;
        hm += 1.0f;
        val *= up;
    }
    if (m.mask_round != 0) {
        if (m.mask_exp != 0) hm += 16.0f;
        const float ax = uvx - 0.5f, ay = uvy - 0.5f;
        const float v = 1.0f - ((ax * ax + ay * ay) + 0.75f);
        val = val * fmaxf(0.0f, v * 1.0f);
    }
    if (m.mask_exp != 0) {
        hm += 2.0f;
        val = (val * (gen_expf(val) - 1.0f)) / 1.718f;
    }
    if (m.mask_power != 0) {
        hm += 1.25f;
        if (m.mask_exp != 0) hm += 2.0f;
        const float b = val + 0.5f;
        val = val * ((((b * b) * b - 0.125f) / 3.25f) * 0.55f + 0.45f);
    }
    if (m.mask_slope != 0) val = 0.25f * val + 0.75f * ((val * uvx) * uvy);
    val = val + float(4u * m.mask_round) * val;
    if (m.terrace > 0) {
        const float step = 1.0f / (float(m.terrace) * hm);
        const float lol = floorf(val / step);
        val = ((m.terrace_scale * lol) * step) + val * (1.0f - m.terrace_scale);
    }
    const size_t i = size_t(y) * size_t(n) + size_t(x);
    rock[i] = fminf(m.max_height, (val * m.max_height) * hm);
    dirt[i] = (perlfbm(fx + 13.7f, fy + 27.1f, m) + 1.5f) * m.max_dirt;
}

}  // namespace

// Philox-driven Fisher-Yates permutation, identical to the oracle's.
void fbm_permutation(uint64_t seed, uint8_t perm[256]) {
    for (int i = 0; i < 256; ++i) perm[i] = uint8_t(i);
    const uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
    for (int i = 255; i >= 1; --i) {
        uint32_t o[4];
        philox4x32_10(uint32_t(i), 0x5eedu, 0u, 0u, k0, k1, o);
        const int j = int(o[0] % uint32_t(i + 1));
        const uint8_t t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
    }
}

hipError_t launch_fbm(float* out, int nx, int ny, const uint8_t* perm_host, int octaves,
                      float lacunarity, float gain, float base_freq, hipStream_t s) {
    Perm P;
    for (int i = 0; i < 256; ++i) P.p[i] = perm_host[i];
    const dim3 grid((nx + 63) / 64, (ny + 3) / 4);
    k_fbm<<<grid, 256, 0, s>>>(out, nx, ny, P, octaves, lacunarity, gain, base_freq);
    return hipGetLastError();
}

hipError_t launch_minmax(const float* v, size_t n, unsigned* mm_dev, hipStream_t s) {
    k_mm_reset<<<1, 64, 0, s>>>(mm_dev);
    size_t blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    k_minmax<<<unsigned(blocks), 256, 0, s>>>(v, n, mm_dev);
    return hipGetLastError();
}

hipError_t launch_rescale(float* v, size_t n, const unsigned* mm_dev, float lo, float hi,
                          hipStream_t s) {
    k_rescale<<<unsigned((n + 255) / 256), 256, 0, s>>>(v, n, mm_dev, lo, hi);
    return hipGetLastError();
}

hipError_t launch_gen_heightmap(float* rock, float* dirt, int n, const hg_map_settings& ms, hipStream_t s) {
    MapParams m;
    m.max_height = ms.max_height; m.max_dirt = ms.max_dirt; m.height_mult = ms.height_mult; m.seed = ms.seed;
    m.persistance = ms.persistance; m.lacunarity = ms.lacunarity; m.scale = ms.scale;
    m.uplift_scale = ms.uplift_scale; m.domain_warp_scale = ms.domain_warp_scale;
    m.terrace_scale = ms.terrace_scale; m.octaves = ms.octaves; m.domain_warp = ms.domain_warp;
    m.terrace = ms.terrace; m.mask_round = ms.mask_round; m.mask_exp = ms.mask_exp;
    m.mask_power = ms.mask_power; m.mask_slope = ms.mask_slope; m.uplift = ms.uplift;
    const dim3 grid((n + 63) / 64, (n + 3) / 4);
    k_gen_heightmap<<<grid, 256, 0, s>>>(rock, dirt, n, m);
    return hipGetLastError();
}

// RGBA32F texel i = (a[i], b[i], c[i], d[i]) for channel reads; a null plane
// reads 0, and with `total` the fourth component is (a + b) + c, the HEIGHT
// channel's total height.
__global__ __launch_bounds__(256) void k_interleave_rgba(float4* out, const float* a, const float* b,
                                                         const float* c, const float* d, size_t n, int total) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
        const float x = a ? a[i] : 0.0f, y = b ? b[i] : 0.0f, z = c ? c[i] : 0.0f;
        const float w = total ? (x + y) + z : (d ? d[i] : 0.0f);
        out[i] = make_float4(x, y, z, w);
    }
}

hipError_t launch_interleave_rgba(float* out, const float* const planes[4], size_t n, bool total, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, 8192));
    k_interleave_rgba<<<grid, 256, 0, s>>>(reinterpret_cast<float4*>(out), planes[0], planes[1], planes[2],
                                           planes[3], n, total ? 1 : 0);
    return hipGetLastError();
}

// The inverse for channel writes: plane k (when not null) = component k.
__global__ __launch_bounds__(256) void k_deinterleave_rgba(const float4* in, float* a, float* b, float* c,
                                                           float* d, size_t n) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
        const float4 v = in[i];
        if (a) a[i] = v.x;
        if (b) b[i] = v.y;
        if (c) c[i] = v.z;
        if (d) d[i] = v.w;
    }
}

hipError_t launch_deinterleave_rgba(const float* in, float* const planes[4], size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, 8192));
    k_deinterleave_rgba<<<grid, 256, 0, s>>>(reinterpret_cast<const float4*>(in), planes[0], planes[1], planes[2],
                                             planes[3], n);
    return hipGetLastError();
}

hipError_t launch_fill(float* p, size_t n, float value, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_fill<<<unsigned((n + 255) / 256), 256, 0, s>>>(p, n, value);
    return hipGetLastError();
}

}  // namespace hgk
