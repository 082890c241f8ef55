// hg_init.hip — config heightmap initialisation: seeded-permutation 2-D
// simplex fBm, min-max rescaled into the rock layer.  This replaces the
// reference's World::gen_heightmap dispatch (src/state.cpp:116-147) for the
// configs' simplex-fBm terrains.
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

hipError_t launch_fill(float* p, size_t n, float value, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_fill<<<unsigned((n + 255) / 256), 256, 0, s>>>(p, n, value);
    return hipGetLastError();
}

}  // namespace hgk
