// hg_oracle.hpp — CPU restatement of hydro-gen's erosion hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
// against; nothing in the product links, loads or calls it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// Every pass is a literal restatement of the reference GLSL compute shader it
// names (file:line of ger0/hydro-gen), templated on the arithmetic type T so
// the same code runs in double (the FP64 oracle), float (the CPU baseline and
// ulp-level checks) and a dual-number type (forward-mode derivatives).
// Each pass is Jacobi: it reads the input state and writes a new one, like
// the reference's Tex_pair ping-pong (src/shaderprogram.cpp:51-109).
//
// Decisions where the reference is undefined or silent (documented, not
// pinned by running the reference, which cannot run here):
//  * out-of-range imageLoad/texelFetch reads 0 (GL robust access);
//  * smoothstep(1e-3, 5e-4, x) uses the textbook formula despite edge0 > edge1;
//  * the stored "total" channel is recomputed as rock+dirt+water wherever a
//    pass reads it (the reference stores it after every pass).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace orc {

// ---------------------------------------------------------------------------
// Arithmetic helpers.  `val()` strips a dual number to its value so that
// comparisons (the branches of the reference) follow the primal computation.
// ---------------------------------------------------------------------------
inline double val(double x) { return x; }
inline float val(float x) { return x; }
template <int N> struct Dual;
template <int N> double val(const Dual<N>& x);
template <int K> struct Shadow;
template <int K> double val(const Shadow<K>& x);

template <class T> inline T tmax(T a, T b) { return val(a) < val(b) ? b : a; }   // GLSL max(a,b)
template <class T> inline T tmin(T a, T b) { return val(b) < val(a) ? b : a; }   // GLSL min(a,b)
template <class T> inline T tclamp(T x, T lo, T hi) { return tmin(tmax(x, lo), hi); }

using std::sqrt;
using std::atan;
using std::floor;
using std::fabs;
using std::exp;

// exp of the reference generator's exp mask.  FP64: the library exp.  FP32:
// the method the device kernel uses (Cody and Waite's reduction
// x = k ln2 + r with a 12-bit head of ln2, the degree-7 Taylor polynomial of
// exp(r), exact scaling by 2^k), restated here; pinned against FP64 exp.
inline double gen_exp(double x) { return std::exp(x); }
inline float gen_exp(float x) {
    if (!(x < 88.5f)) return x != x ? x : HUGE_VALF;
    if (x < -87.0f) return 0.0f;
    const float k = std::rint(x * 1.44269504088896341f);
    const float r = (x - k * 0.693145751953125f) - k * 1.42860682030941723e-6f;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return std::ldexp(p, int(k));
}

// FP32 literal 999999999999.0 used by the reference for out-of-grid heights
// (glsl/hydro_flux.glsl:45-52, glsl/thermal_erosion.glsl:20-26).
constexpr float OOB_HEIGHT_F = 999999999999.0f;

// ---------------------------------------------------------------------------
// Settings: Erosion_data (glsl/bindings.glsl:39-60) in type T.
// ---------------------------------------------------------------------------
template <class T> struct Settings {
    uint32_t particle_count = 0;
    T Kc, Kalpha[2], Kconv, Ks[2], Kd[2], Ke, E, Kspeed[2], G, dt;
};

// Per-cell state, structure of arrays, index y*nx + x.  Planes 0..8 persist
// across steps; vx, vy, vz are the velocity texture; thermal C/D are the
// thermal outflow textures (glsl/thermal_erosion.glsl:107-108).
template <class T> struct State {
    int nx = 0, ny = 0;
    std::vector<T> r, d, b;            // rock, dirt, water  (heightmap .r .g .b)
    std::vector<T> fl, fr, ft, fb;     // flux (L, R, T, B)
    std::vector<T> s0, s1;             // sediment (rock, dirt)
    std::vector<T> vx, vy, vz;         // velocity (u, v, d1+d2)
    std::vector<T> tc[4], td[4];       // thermal cross (L,R,T,B) / diagonal (LT,RT,LB,RB)
    void resize(int x, int y) {
        nx = x; ny = y;
        size_t n = size_t(x) * size_t(y);
        for (auto* p : planes()) p->assign(n, T(0));
    }
    std::vector<std::vector<T>*> planes() {
        std::vector<std::vector<T>*> v = {&r, &d, &b, &fl, &fr, &ft, &fb, &s0, &s1, &vx, &vy, &vz};
        for (int k = 0; k < 4; ++k) v.push_back(&tc[k]);
        for (int k = 0; k < 4; ++k) v.push_back(&td[k]);
        return v;
    }
    // Scratch planes for Jacobi outputs; a pass writes into scratch and then
    // swaps the planes it produced into place.
    std::vector<T> tmp[8];
    std::vector<T>& scratch(int k) {
        tmp[k].resize(size_t(nx) * size_t(ny));
        return tmp[k];
    }
    inline size_t at(int x, int y) const { return size_t(y) * size_t(nx) + size_t(x); }
    inline bool in(int x, int y) const { return x >= 0 && y >= 0 && x < nx && y < ny; }
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based RNG of
// the config rain and the fBm permutation shuffle.
// ---------------------------------------------------------------------------
struct Philox {
    static void round(uint32_t c[4], const uint32_t k[2]) {
        uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
        uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        uint32_t o[4] = {hi1 ^ c[1] ^ k[0], lo1, hi0 ^ c[3] ^ k[1], lo0};
        std::memcpy(c, o, sizeof(o));
    }
    static void gen(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
        uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
        uint32_t k[2] = {key[0], key[1]};
        for (int r = 0; r < 10; ++r) {
            if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
            round(c, k);
        }
        std::memcpy(out, c, 4 * sizeof(uint32_t));
    }
};

// Config rain, uniform mode: u in [0,1) with 24 bits, exact in FP32.
inline double rain_uniform_u(uint64_t seed, uint64_t step, uint32_t terrain, int x, int y) {
    uint32_t ctr[4] = {uint32_t(x) >> 2, uint32_t(y), uint32_t(step), terrain};
    uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint32_t o[4];
    Philox::gen(ctr, key, o);
    return double(o[x & 3] >> 8) * (1.0 / 16777216.0);
}

// ---------------------------------------------------------------------------
// Simplex-fBm heightmap init: seeded-permutation 2-D simplex fBm (this
// project's own init for "simplex fBm (seeded permutation)" terrains; the reference
// init is glsl/heightmap.glsl:52-158 and is a different noise).
//  * permutation: Fisher-Yates over 0..255 driven by Philox4x32-10 with
//    key = seed, ctr = {i, 0x5eed, 0, 0}, j = word0 % (i+1), i = 255..1;
//  * simplex: Perlin's 2-D simplex grid (skew F2 = (sqrt3-1)/2, unskew
//    G2 = (3-sqrt3)/6), radial kernel (0.5 - r^2)^4, 8 gradients
//    (+-1,+-1), (+-1,0), (0,+-1) chosen by perm[(i + perm[j]) & 255] & 7,
//    sum scaled by 70;
//  * fBm: sum_o gain^o * simplex(x * f_o, y * f_o), f_o = base_freq * lacunarity^o.
// ---------------------------------------------------------------------------
inline void fbm_permutation(uint64_t seed, uint8_t perm[256]) {
    for (int i = 0; i < 256; ++i) perm[i] = uint8_t(i);
    uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    for (int i = 255; i >= 1; --i) {
        uint32_t ctr[4] = {uint32_t(i), 0x5eedu, 0u, 0u}, o[4];
        Philox::gen(ctr, key, o);
        int j = int(o[0] % uint32_t(i + 1));
        std::swap(perm[i], perm[j]);
    }
}

template <class T> struct Simplex2 {
    const uint8_t* perm;
    static T grad_dot(int h, T x, T y) {
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
    T corner(int gi, int gj, T x, T y) const {
        T t = T(0.5) - x * x - y * y;
        if (!(val(t) > 0)) return T(0);
        int h = perm[(gi + perm[gj & 255]) & 255];
        T t2 = t * t;
        return t2 * t2 * grad_dot(h, x, y);
    }
    T operator()(T x, T y) const {
        const T F2 = T(0.36602540378443864676);  // (sqrt(3)-1)/2
        const T G2 = T(0.21132486540518711775);  // (3-sqrt(3))/6
        T s = (x + y) * F2;
        T fi = floor(x + s), fj = floor(y + s);
        T t = (fi + fj) * G2;
        T x0 = x - (fi - t), y0 = y - (fj - t);
        int i1 = val(x0) > val(y0) ? 1 : 0, j1 = 1 - i1;
        T x1 = x0 - T(i1) + G2, y1 = y0 - T(j1) + G2;
        T x2 = x0 - T(1) + T(2) * G2, y2 = y0 - T(1) + T(2) * G2;
        int i = int(val(fi)) & 255, j = int(val(fj)) & 255;
        T n = corner(i, j, x0, y0) + corner(i + i1, j + j1, x1, y1) + corner(i + 1, j + 1, x2, y2);
        return T(70) * n;
    }
};

template <class T>
T fbm_value(const Simplex2<T>& sx, int x, int y, int octaves, T lacunarity, T gain, T base_freq) {
    T sum = T(0), amp = T(1), f = base_freq;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * sx(T(x) * f, T(y) * f);
        amp *= gain;
        f *= lacunarity;
    }
    return sum;
}

// Fill `out` (nx*ny) with the min-max rescaled fBm.
template <class T>
void fbm_init(std::vector<T>& out, int nx, int ny, uint64_t seed, int octaves, T lacunarity, T gain,
              T base_freq, T lo, T hi) {
    uint8_t perm[256];
    fbm_permutation(seed, perm);
    Simplex2<T> sx{perm};
    out.assign(size_t(nx) * ny, T(0));
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
            out[size_t(y) * nx + x] = fbm_value(sx, x, y, octaves, lacunarity, gain, base_freq);
    T mn = out[0], mx = out[0];
    for (auto& v : out) { mn = tmin(mn, v); mx = tmax(mx, v); }
    T span = mx - mn;
    for (auto& v : out) v = val(span) > 0 ? lo + (hi - lo) * ((v - mn) / span) : lo;
}

// ---------------------------------------------------------------------------
// Rain (config modes; the reference's glsl/rain.glsl:32-56 adds `incr` to
// water and recomputes the total).
// ---------------------------------------------------------------------------
enum { RAIN_OFF = 0, RAIN_CONSTANT = 1, RAIN_UNIFORM = 2 };

template <class T>
void pass_rain(State<T>& s, int mode, T amount, uint64_t seed, uint64_t step, uint32_t terrain = 0) {
    if (mode == RAIN_OFF) return;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < s.ny; ++y)
        for (int x = 0; x < s.nx; ++x) {
            T incr = mode == RAIN_CONSTANT ? amount
                                           : amount * T(rain_uniform_u(seed, step, terrain, x, y));
            s.b[s.at(x, y)] += incr;
        }
}

// ---------------------------------------------------------------------------
// Reference rain (glsl/rain.glsl:32-56), the mode the reference runs every
// `period` steps, with its wall-clock seed replaced by an explicit time:
//   r = max(0, sfbm(x, y)), sfbm = gln_sfbm (glsl/simplex_noise.glsl:419-456)
//   with seed fract(time * 1372.914227) * 1000, persistence 0.5, lacunarity
//   2, scale = drops, 8 octaves, over gln_simplex (:377-402);
//   incr = amount * r, plus mountain * multip * r / ((1 - thresh) * max_h)
//   where mountain = w - max_h * thresh > 0 (w = rock + dirt + water).
// GLSL semantics: mod(x, y) = x - y * floor(x / y), fract(x) = x - floor(x);
// the literals are FP32 constants.  pow(result, 1.0) in gln_sfbm is taken as
// result (GLSL leaves pow of a negative base undefined; max(0, .) follows).
// ---------------------------------------------------------------------------
template <class T> inline T glsl_mod(T x, T y) { return x - y * floor(x / y); }
template <class T> inline T glsl_fract(T x) { return x - floor(x); }

template <class T> T ref_simplex(T vx, T vy) {
    const T Cx = T(0.211324865405187f), Cy = T(0.366025403784439f), Cz = T(-0.577350269189626f),
            Cw = T(0.024390243902439f);
    const T d = vx * Cy + vy * Cy;                                    // dot(v, C.yy)
    T ix = floor(vx + d), iy = floor(vy + d);
    const T di = ix * Cx + iy * Cx;                                   // dot(i, C.xx)
    const T x0x = (vx - ix) + di, x0y = (vy - iy) + di;
    const T i1x = val(x0x) > val(x0y) ? T(1) : T(0), i1y = T(1) - i1x;
    const T x1x = (x0x + Cx) - i1x, x1y = (x0y + Cx) - i1y;           // x12.xy
    const T x2x = x0x + Cz, x2y = x0y + Cz;                           // x12.zw
    ix = glsl_mod(ix, T(289));
    iy = glsl_mod(iy, T(289));
    auto rnd = [](T q) { return glsl_mod((q * T(34) + T(1)) * q, T(289)); };   // gln_rand3
    const T p[3] = {rnd((rnd(iy) + ix)), rnd((rnd(iy + i1y) + ix) + i1x), rnd((rnd(iy + T(1)) + ix) + T(1))};
    const T gx[3] = {x0x, x1x, x2x}, gy[3] = {x0y, x1y, x2y};
    T acc = T(0);
    for (int c = 0; c < 3; ++c) {
        T m = tmax(T(0.5f) - (gx[c] * gx[c] + gy[c] * gy[c]), T(0));
        m = m * m;
        m = m * m;
        const T x = T(2) * glsl_fract(p[c] * Cw) - T(1);
        const T h = fabs(x) - T(0.5f);
        const T a0 = x - floor(x + T(0.5f));
        m = m * (T(1.79284291400159f) - T(0.85373472095314f) * (a0 * a0 + h * h));
        acc = acc + m * (a0 * gx[c] + h * gy[c]);                    // dot(m, g), x then y then z
    }
    return T(130) * acc;
}

template <class T> T ref_sfbm(T x, T y, T seed, T persistence, T lacunarity, T scale, int octaves) {
    x = x + seed * T(100);
    y = y + seed * T(100);
    T result = T(0), amplitude = T(1), frequency = T(1), maximum = amplitude;
    for (int o = 0; o < octaves; ++o) {
        const T px = (x * frequency) * scale, py = (y * frequency) * scale;
        result = result + ref_simplex(px, py) * amplitude;
        frequency = frequency * lacunarity;
        amplitude = amplitude * persistence;
        maximum = maximum + amplitude;
    }
    return result / maximum;
}

struct RefRain {
    float amount = 0.0f, mountain_thresh = 0.0f, mountain_multip = 0.0f, drops = 0.0f;
    float time = 0.0f, max_height = 0.0f;
};

template <class T> T ref_rain_incr(int x, int y, T w, const RefRain& rr) {
    const T seed = glsl_fract(T(rr.time) * T(1.372914227e3f)) * T(1000);
    const T r = tmax(T(0), ref_sfbm(T(x), T(y), seed, T(0.5f), T(2), T(rr.drops), 8));
    T incr = T(rr.amount) * r;
    const T mountain = w - T(rr.max_height) * T(rr.mountain_thresh);
    if (val(mountain) > 0)
        incr = incr + ((mountain * T(rr.mountain_multip)) * r) / ((T(1) - T(rr.mountain_thresh)) * T(rr.max_height));
    return incr;
}

template <class T> void pass_rain_reference(State<T>& s, const RefRain& rr) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < s.ny; ++y)
        for (int x = 0; x < s.nx; ++x) {
            const size_t i = s.at(x, y);
            const T w = s.r[i] + s.d[i] + s.b[i];
            s.b[i] = s.b[i] + ref_rain_incr(x, y, w, rr);
        }
}

// ---------------------------------------------------------------------------
// Reference heightmap generator (glsl/heightmap.glsl:52-158) and its noise
// (glsl/simplex_noise.glsl:579-676).  int32 arithmetic of the hash wraps; the
// literals are FP32 constants; GLSL's pow(x, 2) of a negative x (round mask)
// is x*x and pow(x, 3) (power mask) is x*x*x; pow(result, 1) of the fBms is
// result.
// ---------------------------------------------------------------------------
// Hugo Elias hash of an integer lattice point to a gradient in [-1, 1]^2.
template <class T> void elias_hash(int px, int py, T& gx, T& gy) {
    const uint32_t ux = uint32_t(px), uy = uint32_t(py);
    uint32_t n[2] = {ux * 3u + uy * 311u, ux * 37u + uy * 113u};
    T g[2];
    for (int c = 0; c < 2; ++c) {
        uint32_t v = (n[c] << 13) ^ n[c];
        v = v * (v * v * 15731u + 789221u) + 1376312589u;
        const int32_t lo = int32_t(v & 0x0fffffffu);
        // float(int) rounds to FP32 first, like the shader's vec2(ivec2)
        g[c] = T(-1.0f) + (T(2.0f) * T(float(lo))) / T(float(0x0fffffff));
    }
    gx = g[0];
    gy = g[1];
}

// Gradient noise with quintic interpolation: value and analytic gradient.
template <class T> void noised(T px, T py, T& v, T& dvx, T& dvy) {
    const T flx = floor(px), fly = floor(py);
    const int ix = int(val(flx)), iy = int(val(fly));
    const T fx = px - flx, fy = py - fly;
    auto quintic = [](T f) { return ((f * f) * f) * (f * (f * T(6) - T(15)) + T(10)); };
    auto dquintic = [](T f) { return ((T(30) * f) * f) * (f * (f - T(2)) + T(1)); };
    const T ux = quintic(fx), uy = quintic(fy), dux = dquintic(fx), duy = dquintic(fy);
    T ga[2], gb[2], gc[2], gd[2];
    elias_hash(ix, iy, ga[0], ga[1]);
    elias_hash(ix + 1, iy, gb[0], gb[1]);
    elias_hash(ix, iy + 1, gc[0], gc[1]);
    elias_hash(ix + 1, iy + 1, gd[0], gd[1]);
    const T va = ga[0] * fx + ga[1] * fy;
    const T vb = gb[0] * (fx - T(1)) + gb[1] * fy;
    const T vc = gc[0] * fx + gc[1] * (fy - T(1));
    const T vd = gd[0] * (fx - T(1)) + gd[1] * (fy - T(1));
    const T k4 = ((va - vb) - vc) + vd;
    v = ((va + ux * (vb - va)) + uy * (vc - va)) + (ux * uy) * k4;
    T d[2];
    const T uo[2] = {uy, ux}, du[2] = {dux, duy}, vbc[2] = {vb, vc};
    for (int c = 0; c < 2; ++c)
        d[c] = (((ga[c] + ux * (gb[c] - ga[c])) + uy * (gc[c] - ga[c])) +
                (ux * uy) * (((ga[c] - gb[c]) - gc[c]) + gd[c])) +
               du[c] * ((uo[c] * k4 + vbc[c]) - va);
    dvx = d[0];
    dvy = d[1];
}

struct MapCfg {
    float max_height, max_dirt, height_mult, seed, persistance, lacunarity, scale, uplift_scale,
        domain_warp_scale, terrace_scale;
    int octaves, domain_warp, terrace;
    unsigned mask_round, mask_exp, mask_power, mask_slope, uplift;
};

// perlfbm (damped = false) and erosion_perlfbm (damped = true)
template <class T> T grad_fbm(T x, T y, const MapCfg& m, bool damped) {
    x = x + T(m.seed) * T(100);
    y = y + T(m.seed) * T(100);
    T result = T(0), amplitude = T(1), frequency = damped ? T(1.5f) : T(1), maximum = amplitude;
    T sx = T(0), sy = T(0);
    for (int o = 0; o < m.octaves; ++o) {
        T v, dx, dy;
        noised((x * frequency) * T(m.scale), (y * frequency) * T(m.scale), v, dx, dy);
        const T nv = (v + T(1)) / T(2);
        if (damped) {
            sx = sx + dx;
            sy = sy + dy;
            result = result + (nv * amplitude) / (T(1) + (sx * sx + sy * sy));
        } else {
            result = result + nv * amplitude;
        }
        frequency = frequency * T(m.lacunarity);
        amplitude = amplitude * T(m.persistance);
        maximum = maximum + amplitude;
    }
    return result / maximum;
}

template <class T> T ridged_sfbm(T x, T y, const MapCfg& m) {
    x = x + T(m.seed) * T(100);
    y = y + T(m.seed) * T(100);
    const T scale = T(m.scale) / T(m.uplift_scale);
    T result = T(0), amplitude = T(1), frequency = T(1), maximum = amplitude;
    for (int o = 0; o < m.octaves; ++o) {
        const T nv = T(1) - fabs(ref_simplex((x * frequency) * scale, (y * frequency) * scale));
        result = result + nv * amplitude;
        frequency = frequency * T(m.lacunarity);
        amplitude = amplitude * T(m.persistance);
        maximum = maximum + amplitude;
    }
    return result / maximum;
}

template <class T> void gen_heightmap_cell(int x, int y, int n, const MapCfg& m, T& rock, T& dirt) {
    const T fx = T(float(x)), fy = T(float(y));
    const T uvx = fx / T(float(n)), uvy = fy / T(float(n));
    T distx = T(1), disty = T(1);
    if (m.domain_warp != 0) {
        distx = grad_fbm(fx + T(2.3f), fy + T(2.9f), m, false);
        disty = grad_fbm(fx - T(3.1f), fy - T(4.3f), m, false);
        if (m.domain_warp == 2) {
            const T wx = fx + T(m.domain_warp_scale) * distx, wy = fy + T(m.domain_warp_scale) * disty;
            distx = grad_fbm(wx - T(5.7f), wy + T(27.9f), m, false);
            disty = grad_fbm(wx + T(11.5f), wy - T(23.7f), m, false);
        }
    }
    T v = grad_fbm(fx + T(m.domain_warp_scale) * distx, fy + T(m.domain_warp_scale) * disty, m, true);
    T hm = T(m.height_mult);
    if (m.uplift != 0) {
        const T up = ridged_sfbm(fx - T(7.3f), fy + T(19.9f), m);
        if (m.mask_exp != 0) hm = hm + T(2.5f);
        hm = hm + T(1);
        v = v * up;
    }
    if (m.mask_round != 0) {
        if (m.mask_exp != 0) hm = hm + T(16);
        const T ax = uvx - T(0.5f), ay = uvy - T(0.5f);
        const T w = T(1) - ((ax * ax + ay * ay) + T(0.75f));
        v = v * tmax(T(0), w * T(1));
    }
    if (m.mask_exp != 0) {
        hm = hm + T(2);
        v = (v * (gen_exp(v) - T(1))) / T(1.718f);
    }
    if (m.mask_power != 0) {
        hm = hm + T(1.25f);
        if (m.mask_exp != 0) hm = hm + T(2);
        const T b = v + T(0.5f);
        v = v * ((((b * b) * b - T(0.125f)) / T(3.25f)) * T(0.55f) + T(0.45f));
    }
    if (m.mask_slope != 0) v = T(0.25f) * v + T(0.75f) * ((v * uvx) * uvy);
    v = v + T(float(4u * m.mask_round)) * v;
    if (m.terrace > 0) {
        const T step = T(1) / (T(float(m.terrace)) * hm);
        const T lol = floor(v / step);
        v = ((T(m.terrace_scale) * lol) * step) + v * (T(1) - T(m.terrace_scale));
    }
    rock = tmin(T(m.max_height), (v * T(m.max_height)) * hm);
    dirt = (grad_fbm(fx + T(13.7f), fy + T(27.1f), m, false) + T(1.5f)) * T(m.max_dirt);
}

template <class T> void gen_heightmap(std::vector<T>& rock, std::vector<T>& dirt, int n, const MapCfg& m) {
    rock.assign(size_t(n) * n, T(0));
    dirt.assign(size_t(n) * n, T(0));
#pragma omp parallel for schedule(static)
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) gen_heightmap_cell(x, y, n, m, rock[size_t(y) * n + x], dirt[size_t(y) * n + x]);
}

// ---------------------------------------------------------------------------
// hydro_flux: glsl/hydro_flux.glsl:77-166 (helpers :33-47).
// ---------------------------------------------------------------------------
template <class T> void pass_flux(State<T>& s, const Settings<T>& k) {
    const int nx = s.nx, ny = s.ny;
    const T A = T(1), L = T(1);
    std::vector<T>&ob = s.scratch(0), &ofl = s.scratch(1), &ofr = s.scratch(2), &oft = s.scratch(3),
                  &ofb = s.scratch(4), &ovx = s.scratch(5), &ovy = s.scratch(6), &ovz = s.scratch(7);
    auto wheight = [&](int x, int y) -> T {   // get_wheight, :33-39
        if (!s.in(x, y)) return T(OOB_HEIGHT_F);
        size_t i = s.at(x, y);
        return s.r[i] + s.d[i] + s.b[i];
    };
    auto flux = [&](int x, int y, int c) -> T {  // get_flux, :41-47; c = 0..3 = L,R,T,B
        if (!s.in(x, y)) return T(0);
        size_t i = s.at(x, y);
        return c == 0 ? s.fl[i] : c == 1 ? s.fr[i] : c == 2 ? s.ft[i] : s.fb[i];
    };
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            size_t i = s.at(x, y);
            T f[4] = {s.fl[i], s.fr[i], s.ft[i], s.fb[i]};
            T d1 = s.b[i];
            T w = s.r[i] + s.d[i] + s.b[i];
            T dh[4] = {w - wheight(x - 1, y), w - wheight(x + 1, y), w - wheight(x, y + 1),
                       w - wheight(x, y - 1)};                                     // :88-92
            T in[4] = {flux(x - 1, y, 1), flux(x + 1, y, 0), flux(x, y + 1, 3),
                       flux(x, y - 1, 2)};                                          // :94-98
            for (int c = 0; c < 4; ++c)                                             // :100-107
                f[c] = tmax(T(0), k.E * f[c] + k.dt * A * (k.G * dh[c]) / L);
            if (x <= 0) f[0] = T(0); else if (x >= nx - 1) f[1] = T(0);             // :110-119
            if (y <= 0) f[3] = T(0); else if (y >= ny - 1) f[2] = T(0);
            T sum_in = in[0] + in[1] + in[2] + in[3];
            T sum_out = f[0] + f[1] + f[2] + f[3];
            T K = tmin(T(1), (d1 * L * L) / (sum_out * k.dt));                      // :125
            if (val(sum_out) == 0) K = T(1);  // b/0 = inf or 0/0 = NaN: GLSL min returns 1
            for (int c = 0; c < 4; ++c) f[c] *= K;
            sum_out *= K;
            T d_volume = k.dt * (sum_in - sum_out);                                 // :133
            T d2 = tmax(T(0), d1 + (d_volume / (L * L)));
            ob[i] = d2;
            T vz = d1 + d2;                                                         // :144
            ovz[i] = vz;
            if (val(vz) > 0) {                                                      // :146-161
                ovx[i] = (flux(x - 1, y, 1) - flux(x, y, 0) + flux(x, y, 1) - flux(x + 1, y, 0)) / (L * vz);
                ovy[i] = (flux(x, y - 1, 2) - flux(x, y, 3) + flux(x, y, 2) - flux(x, y + 1, 3)) / (L * vz);
            } else {
                ovx[i] = T(0);
                ovy[i] = T(0);
            }
            ofl[i] = f[0]; ofr[i] = f[1]; oft[i] = f[2]; ofb[i] = f[3];
        }
    std::swap(s.b, ob); std::swap(s.fl, ofl); std::swap(s.fr, ofr); std::swap(s.ft, oft);
    std::swap(s.fb, ofb); std::swap(s.vx, ovx); std::swap(s.vy, ovy); std::swap(s.vz, ovz);
}

// ---------------------------------------------------------------------------
// hydro_erosion: glsl/hydro_erosion.glsl:23-92.
// ---------------------------------------------------------------------------
template <class T> T erosion_sin_slope(const State<T>& s, int x, int y) {
    // get_terr_normal, :23-35; out-of-grid imageLoad reads 0.
    auto rg = [&](int xx, int yy, int c) -> T {
        if (!s.in(xx, yy)) return T(0);
        size_t i = s.at(xx, yy);
        return c == 0 ? s.r[i] : s.d[i];
    };
    const T L = T(1);
    T dx = rg(x + 1, y, 0) + rg(x + 1, y, 1) - rg(x - 1, y, 0) - rg(x - 1, y, 1);
    T dz = rg(x, y + 1, 0) + rg(x, y + 1, 1) - rg(x, y - 1, 0) - rg(x, y - 1, 1);
    // normalize(cross((2L, dx, 0), (0, dz, 2L))) = (2L dx, -4L^2, 2L dz) / |.|
    T cx = dx * (T(2) * L) - T(0) * dz;
    T cy = T(0) * T(0) - (T(2) * L) * (T(2) * L);
    T cz = (T(2) * L) * dz - dx * T(0);
    T len = sqrt(cx * cx + cy * cy + cz * cz);
    T ny = cy / len;
    return fabs(sqrt(T(1) - ny * ny));                                             // :57
}

template <class T> void erosion_cell(T& r, T& d, T& s0, T& s1, T vx, T vy, T vz, T sin_a,
                                     const Settings<T>& k) {
    // :45-52
    T dd = vz;
    T ero_vel = sqrt(vx * vx + vy * vy);
    if (val(dd) < double(1e-3f)) {
        dd = tmax(T(5e-4f), dd);
        T t = tclamp((dd - T(1e-3f)) / (T(5e-4f) - T(1e-3f)), T(0), T(1));
        T sm = t * t * (T(3) - T(2) * t);
        ero_vel = ero_vel * (T(1) - sm) + T(0) * sm;
    }
    T cap = T(0);
    T* h[2] = {&r, &d};
    T* sd[2] = {&s0, &s1};
    for (int i = 1; i >= 0; --i) {                                                  // :58-83
        T Kls = k.dt * k.Ks[i];
        T Kld = k.dt * k.Kd[i];
        T c = tmax(T(0), k.Kc * tmax(T(0.02f), sin_a) * ero_vel - cap);
        if (val(c) > val(*sd[i])) {
            T old = *h[i];
            T delta = Kls * (c - *sd[i]);
            *h[i] -= delta;
            *sd[i] += delta;
            if (val(*h[i]) < 0) {
                *sd[i] += *h[i];
                *h[i] = T(0);
                cap += old;
            } else {
                break;
            }
        } else {
            T delta = Kld * (*sd[i] - c);
            *h[i] += delta;
            *sd[i] -= delta;
        }
    }
    T conv = s0 * k.Kconv * k.dt;                                                   // :84-88
    s1 += conv;
    s0 -= conv;
}

template <class T> void pass_erosion(State<T>& s, const Settings<T>& k) {
    std::vector<T>&orr = s.scratch(0), &od = s.scratch(1), &os0 = s.scratch(2), &os1 = s.scratch(3);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < s.ny; ++y)
        for (int x = 0; x < s.nx; ++x) {
            size_t i = s.at(x, y);
            T sin_a = erosion_sin_slope(s, x, y);
            T r = s.r[i], d = s.d[i], s0 = s.s0[i], s1 = s.s1[i];
            erosion_cell(r, d, s0, s1, s.vx[i], s.vy[i], s.vz[i], sin_a, k);
            orr[i] = r; od[i] = d; os0[i] = s0; os1[i] = s1;
        }
    std::swap(s.r, orr); std::swap(s.d, od); std::swap(s.s0, os0); std::swap(s.s1, os1);
}

// ---------------------------------------------------------------------------
// sediment_transport: glsl/sediment_transport.glsl:66-93 with img_bilinear
// (glsl/img_interpolation.glsl:3-22) and get_lerp_sed clamping (:26-30).
// ---------------------------------------------------------------------------
template <class T> T mixv(T a, T b, T t) { return a * (T(1) - t) + b * t; }

template <class T> void bilinear2(const State<T>& s, const std::vector<T>& p0, const std::vector<T>& p1,
                                  T bx, T by, T& o0, T& o1) {
    int ix = int(val(bx)), iy = int(val(by));          // ivec2(sample_pos): truncation
    T fx = bx - floor(bx), fy = by - floor(by);         // fract
    auto fetch = [&](const std::vector<T>& p, int x, int y) -> T {
        return s.in(x, y) ? p[s.at(x, y)] : T(0);        // robust texelFetch reads 0
    };
    for (int c = 0; c < 2; ++c) {
        const std::vector<T>& p = c == 0 ? p0 : p1;
        T v1 = mixv(fetch(p, ix, iy), fetch(p, ix + 1, iy), fx);
        T v2 = mixv(fetch(p, ix, iy + 1), fetch(p, ix + 1, iy + 1), fx);
        (c == 0 ? o0 : o1) = mixv(v1, v2, fy);
    }
}

template <class T> void pass_transport(State<T>& s, const Settings<T>& k) {
    std::vector<T>&os0 = s.scratch(0), &os1 = s.scratch(1), &ob = s.scratch(2);
    const T xmax = T(s.nx - 1), ymax = T(s.ny - 1);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < s.ny; ++y)
        for (int x = 0; x < s.nx; ++x) {
            size_t i = s.at(x, y);
            T bx = T(x) - s.vx[i] * k.dt;                                           // :70
            T by = T(y) - s.vy[i] * k.dt;
            bx = tclamp(bx, T(0), xmax);                                            // :26-30
            by = tclamp(by, T(0), ymax);
            bilinear2(s, s.s0, s.s1, bx, by, os0[i], os1[i]);
            ob[i] = s.b[i] * (T(1) - k.Ke * k.dt);                                  // :75
        }
    std::swap(s.s0, os0); std::swap(s.s1, os1); std::swap(s.b, ob);
}

// ---------------------------------------------------------------------------
// thermal_erosion: glsl/thermal_erosion.glsl:20-115 (layer `layer`).
// Directions in the reference's order: cross L,R,T,B; diagonal LT,RT,LB,RB.
// ---------------------------------------------------------------------------
static const int TDX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
static const int TDY[8] = {0, 0, 1, -1, 1, 1, -1, -1};

template <class T> void pass_thermal_flux(State<T>& s, const Settings<T>& k, int layer) {
    std::vector<T>* o[8];
    for (int kk = 0; kk < 8; ++kk) o[kk] = &s.scratch(kk);
    const T Lc = T(1);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < s.ny; ++y)
        for (int x = 0; x < s.nx; ++x) {
            size_t i = s.at(x, y);
            T t[2] = {s.r[i], s.d[i]};
            T dh[8];
            for (int kk = 0; kk < 8; ++kk) dh[kk] = T(0);
            for (int l = 0; l <= layer; ++l)                                        // :32-44
                for (int kk = 0; kk < 8; ++kk) {
                    int xx = x + TDX[kk], yy = y + TDY[kk];
                    T nb = s.in(xx, yy) ? (l == 0 ? s.r[s.at(xx, yy)] : s.d[s.at(xx, yy)])
                                        : T(OOB_HEIGHT_F);
                    dh[kk] += t[l] - nb;
                }
            T H = T(0);                                                             // :46-58
            for (int kk = 0; kk < 8; ++kk)
                if (val(dh[kk]) > val(H)) H = dh[kk];
            H = tmin(t[layer], H);
            bool mark[8];
            T bk = T(0), sharp = T(1);
            for (int kk = 0; kk < 8; ++kk) {                                        // :63-91
                mark[kk] = false;
                T bb = dh[kk];
                if (!(val(bb) > 0)) continue;
                T dist = Lc;
                if (kk >= 4) dist *= T(sqrtf(2.0f));
                T alph = atan(bb / dist);
                T Ka = k.Kalpha[layer];
                if (val(alph) > val(Ka)) {
                    T newsh = T(1) + alph - Ka;
                    if (val(newsh) > val(sharp)) sharp = newsh;
                    bk += bb;
                    mark[kk] = true;
                }
            }
            sharp *= sharp * sharp;                                                 // :92
            T S = k.dt * k.Kspeed[layer] * sharp * Lc * H / T(2);                   // :95
            for (int kk = 0; kk < 8; ++kk) {                                        // :98-106
                (*o[kk])[i] = mark[kk] ? S * dh[kk] / bk : T(0);
            }
        }
    for (int kk = 0; kk < 4; ++kk) std::swap(s.tc[kk], *o[kk]);
    for (int kk = 0; kk < 4; ++kk) std::swap(s.td[kk], *o[kk + 4]);
}

// thermal_transport: glsl/thermal_transport.glsl:146-180 (gather_inflow :31-52).
template <class T> void pass_thermal_transport(State<T>& s, int layer) {
    std::vector<T>& oh = s.scratch(0);
    const std::vector<T>& h = layer == 0 ? s.r : s.d;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < s.ny; ++y)
        for (int x = 0; x < s.nx; ++x) {
            size_t i = s.at(x, y);
            auto c = [&](int xx, int yy, int comp) -> T {
                return s.in(xx, yy) ? s.tc[comp][s.at(xx, yy)] : T(0);
            };
            auto dg = [&](int xx, int yy, int comp) -> T {
                return s.in(xx, yy) ? s.td[comp][s.at(xx, yy)] : T(0);
            };
            T in = T(0);
            in += c(x - 1, y, 1);      // from the left cell, its R outflow
            in += c(x + 1, y, 0);      // from the right cell, its L
            in += c(x, y + 1, 3);      // from the top cell, its B
            in += c(x, y - 1, 2);      // from the bottom cell, its T
            in += dg(x - 1, y + 1, 3); // LT cell, its RB
            in += dg(x + 1, y + 1, 2); // RT cell, its LB
            in += dg(x - 1, y - 1, 1); // LB cell, its RT
            in += dg(x + 1, y - 1, 0); // RB cell, its LT
            T sum = T(0);
            for (int kk = 0; kk < 4; ++kk) sum -= s.tc[kk][i];
            for (int kk = 0; kk < 4; ++kk) sum -= s.td[kk][i];
            sum += in;
            oh[i] = h[i] + sum;
        }
    std::swap(layer == 0 ? s.r : s.d, oh);
}

// ---------------------------------------------------------------------------
// smoothing, grid branch (particle_count == 0): glsl/smoothing.glsl:22-74,97-103.
// ---------------------------------------------------------------------------
// Extremum flags of interior cell (x, y) for the rock (er) and surface (eg)
// channels (glsl/smoothing.glsl:40-73).
template <class T> void smooth_flags_at(const State<T>& s, int x, int y, bool& er, bool& eg) {
    const size_t i = s.at(x, y);
    const T tr = s.r[i], tg = s.d[i];
    const size_t il = s.at(x - 1, y), ir = s.at(x + 1, y), it = s.at(x, y + 1), ib = s.at(x, y - 1);
    const T lr = s.r[il], lg = s.d[il], rr = s.r[ir], rg = s.d[ir];
    const T t_r = s.r[it], t_g = s.d[it], br = s.r[ib], bg = s.d[ib];
    T dlr = tr - lr, dlg = tg - lg; dlg += dlr;                                 // :40-48
    T drr = tr - rr, drg = tg - rg; drg += drr;
    T dtr = tr - t_r, dtg = tg - t_g; dtg += dtr;
    T dbr = tr - br, dbg = tg - bg; dbg += dbr;
    const T g_hdiff = fabs((dlg + drg + dtg + dbg) / T(4));                     // :50-55
    const T r_hdiff = fabs((dlr + drr + dtr + dbr) / T(4));
    const T xcr = dlr * drr, xcg = dlg * drg;                                   // :60-61
    const T ycr = dtr * dbr, ycg = dtg * dbg;
    er = ((val(-dlr) > val(r_hdiff) || val(-drr) > val(r_hdiff)) && val(xcr) > 0) ||   // :63
         ((val(-dtr) > val(r_hdiff) || val(-dbr) > val(r_hdiff)) && val(ycr) > 0);
    eg = ((val(-dlg) > val(g_hdiff) || val(-drg) > val(g_hdiff)) && val(xcg) > 0) ||   // :69
         ((val(-dtg) > val(g_hdiff) || val(-dbg) > val(g_hdiff)) && val(ycg) > 0);
}

template <class T> void pass_smooth(State<T>& s, const Settings<T>& k) {
    std::vector<T>&orr = s.scratch(0), &od = s.scratch(1);
    const int nx = s.nx, ny = s.ny;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            size_t i = s.at(x, y);
            if (x == 0 || y == 0 || x == nx - 1 || y == ny - 1) {                   // :29-34
                orr[i] = s.r[i]; od[i] = s.d[i];
                continue;
            }
            bool er, eg;
            smooth_flags_at(s, x, y, er, eg);
            const T tr = s.r[i], tg = s.d[i];
            T nr = tr, ng = tg;
            if (er) nr = (tr + s.r[s.at(x - 1, y)] + s.r[s.at(x + 1, y)] + s.r[s.at(x, y + 1)] +   // :63-67
                          s.r[s.at(x, y - 1)]) / T(5);
            if (eg) ng = (tg + s.d[s.at(x - 1, y)] + s.d[s.at(x + 1, y)] + s.d[s.at(x, y + 1)] +   // :69-73
                          s.d[s.at(x, y - 1)]) / T(5);
            T m = tclamp(k.Kspeed[1] * k.dt, T(0), T(1));                           // :75
            orr[i] = m * nr + (T(1) - m) * s.r[i];                                  // :100
            od[i] = m * ng + (T(1) - m) * s.d[i];
        }
    std::swap(s.r, orr); std::swap(s.d, od);
}

// ---------------------------------------------------------------------------
// One GRID step: Erosion::dispatch_grid (src/erosion.cpp:158-200) with
// run_thermal_erosion (:103-121), preceded by the config rain.
// ---------------------------------------------------------------------------
enum { PASS_RAIN = 0, PASS_FLUX, PASS_EROSION, PASS_TRANSPORT, PASS_TFLUX0, PASS_TTRANS0, PASS_TFLUX1,
       PASS_TTRANS1, PASS_SMOOTH };
enum { STEP_RAIN = 1, STEP_THERMAL = 2, STEP_SMOOTH = 4 };

struct RainCfg {
    int mode = RAIN_OFF;
    double amount = 0;
    uint64_t seed = 0;
    uint32_t terrain = 0;
};

template <class T>
void run_pass(State<T>& s, const Settings<T>& k, const RainCfg& rc, uint64_t step, int pass) {
    switch (pass) {
        case PASS_RAIN: pass_rain(s, rc.mode, T(rc.amount), rc.seed, step, rc.terrain); break;
        case PASS_FLUX: pass_flux(s, k); break;
        case PASS_EROSION: pass_erosion(s, k); break;
        case PASS_TRANSPORT: pass_transport(s, k); break;
        case PASS_TFLUX0: pass_thermal_flux(s, k, 0); break;
        case PASS_TTRANS0: pass_thermal_transport(s, 0); break;
        case PASS_TFLUX1: pass_thermal_flux(s, k, 1); break;
        case PASS_TTRANS1: pass_thermal_transport(s, 1); break;
        case PASS_SMOOTH: pass_smooth(s, k); break;
        default: break;
    }
}

template <class T>
void grid_step(State<T>& s, const Settings<T>& k, const RainCfg& rc, uint64_t step, unsigned flags) {
    if (flags & STEP_RAIN) run_pass(s, k, rc, step, PASS_RAIN);
    run_pass(s, k, rc, step, PASS_FLUX);
    run_pass(s, k, rc, step, PASS_EROSION);
    run_pass(s, k, rc, step, PASS_TRANSPORT);
    if (flags & STEP_THERMAL)
        for (int p = PASS_TFLUX0; p <= PASS_TTRANS1; ++p) run_pass(s, k, rc, step, p);
    if (flags & STEP_SMOOTH) run_pass(s, k, rc, step, PASS_SMOOTH);
}

// ---------------------------------------------------------------------------
// Droplet (particle) erosion, a mass-conserving Beyer-style model (H. Beyer,
// "Implementation of a method for hydraulic erosion", 2015), restated here as
// a serial FP64 model: droplets are simulated one after another, in id order.
// The reference's own particle mode (glsl/particle.glsl:64-136,
// glsl/particle_erosion.glsl:22-128) is a different, non-mass-conserving model.
//  * start: Philox4x32-10(ctr = {id, iteration, 0x64726f70, 0}, key = seed),
//    pos = u24(word k) * (N - 1) for k = 0 (x), 1 (y);
//  * each lifetime step: bilinear height h and gradient g at pos; dir =
//    dir*inertia - g*(1-inertia), normalised; pos += dir; stop when dir = 0
//    (deposit everything at pos) or pos leaves [0, N-1)^2 (sediment leaves the
//    map and is counted as exported); dh = h(pos_new) - h;
//    cap = max(-dh, min_slope)*speed*water*capacity;
//    if sed > cap or dh > 0: deposit min(dh, sed) (uphill) or
//    (sed-cap)*deposition at the old position, bilinear weights, into dirt;
//    else: erode min((cap-sed)*erosion, -dh) through the brush (weights
//    1 - dist/r over cells with dist < r around the old cell, renormalised
//    over in-map cells), taking from dirt first, then rock, never below 0;
//    speed = sqrt(max(0, speed^2 - dh*gravity)); water *= 1 - evaporation;
//  * after the last lifetime step the remaining sediment is deposited at pos.
// Terrain height is rock + dirt.  Returns the exported mass.
// ---------------------------------------------------------------------------
struct DropletParams {
    uint32_t count, max_lifetime;
    double inertia, capacity, deposition, erosion, evaporation, gravity, min_slope;
    int brush_radius;
    double init_water, init_speed;
    uint64_t seed;
};

struct BrushTap { int dx, dy; double w; };

// Brush taps: weights 1 - dist/r over cells with dist < r, normalised (double).
inline std::vector<BrushTap> droplet_brush(int r) {
    std::vector<BrushTap> taps;
    double sum = 0;
    for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx) {
            const double d = std::sqrt(double(dx * dx + dy * dy));
            if (d < r) { taps.push_back({dx, dy, 1.0 - d / r}); sum += 1.0 - d / r; }
        }
    for (auto& t : taps) t.w /= sum;
    return taps;
}

template <class T>
double droplets(std::vector<T>& rock, std::vector<T>& dirt, int n, const DropletParams& p, uint32_t iteration) {
    const auto brush = droplet_brush(p.brush_radius);
    auto H = [&](int x, int y) -> T { const size_t i = size_t(y) * n + x; return rock[i] + dirt[i]; };
    // cell of a position in [0, n-1]; n-1 maps to cell n-2 with fraction 1
    auto cell = [&](T q) { return std::min(int(val(q)), n - 2); };
    auto deposit = [&](T px, T py, T amount) {
        const int x = cell(px), y = cell(py);
        const T u = px - T(x), v = py - T(y);
        dirt[size_t(y) * n + x] += amount * (T(1) - u) * (T(1) - v);
        dirt[size_t(y) * n + x + 1] += amount * u * (T(1) - v);
        dirt[size_t(y + 1) * n + x] += amount * (T(1) - u) * v;
        dirt[size_t(y + 1) * n + x + 1] += amount * u * v;
    };
    const uint32_t key[2] = {uint32_t(p.seed), uint32_t(p.seed >> 32)};
    double exported = 0;
    for (uint32_t id = 0; id < p.count; ++id) {
        uint32_t ctr[4] = {id, iteration, 0x64726f70u, 0u}, o[4];
        Philox::gen(ctr, key, o);
        T px = T(double(o[0] >> 8) * (1.0 / 16777216.0)) * T(n - 1);
        T py = T(double(o[1] >> 8) * (1.0 / 16777216.0)) * T(n - 1);
        T dx = T(0), dy = T(0), speed = T(p.init_speed), water = T(p.init_water), sed = T(0);
        bool alive = true;
        for (uint32_t life = 0; life < p.max_lifetime; ++life) {
            const int cx = cell(px), cy = cell(py);
            const T u = px - T(cx), v = py - T(cy);
            const T h00 = H(cx, cy), h10 = H(cx + 1, cy), h01 = H(cx, cy + 1), h11 = H(cx + 1, cy + 1);
            const T gx = (h10 - h00) * (T(1) - v) + (h11 - h01) * v;
            const T gy = (h01 - h00) * (T(1) - u) + (h11 - h10) * u;
            const T h = h00 * (T(1) - u) * (T(1) - v) + h10 * u * (T(1) - v) + h01 * (T(1) - u) * v + h11 * u * v;
            dx = dx * T(p.inertia) - gx * (T(1) - T(p.inertia));
            dy = dy * T(p.inertia) - gy * (T(1) - T(p.inertia));
            const T len = sqrt(dx * dx + dy * dy);
            if (val(len) > 0) { dx = dx / len; dy = dy / len; }
            const T ox = px, oy = py;
            if (val(dx) == 0 && val(dy) == 0) { deposit(px, py, sed); sed = T(0); alive = false; break; }
            px = px + dx;
            py = py + dy;
            if (!(val(px) >= 0 && val(py) >= 0 && val(px) < n - 1 && val(py) < n - 1)) {
                exported += double(val(sed));
                sed = T(0);
                alive = false;
                break;
            }
            const int nx2 = cell(px), ny2 = cell(py);
            const T u2 = px - T(nx2), v2 = py - T(ny2);
            const T hn = H(nx2, ny2) * (T(1) - u2) * (T(1) - v2) + H(nx2 + 1, ny2) * u2 * (T(1) - v2) +
                         H(nx2, ny2 + 1) * (T(1) - u2) * v2 + H(nx2 + 1, ny2 + 1) * u2 * v2;
            const T dh = hn - h;
            const T cap = tmax(-dh, T(p.min_slope)) * speed * water * T(p.capacity);
            if (val(sed) > val(cap) || val(dh) > 0) {
                const T amount = val(dh) > 0 ? tmin(dh, sed) : (sed - cap) * T(p.deposition);
                sed -= amount;
                deposit(ox, oy, amount);
            } else {
                const T amount = tmin((cap - sed) * T(p.erosion), -dh);
                // interior cells use the normalised weights as they are; at the
                // map edge they are renormalised over the in-map cells
                const bool interior = cx >= p.brush_radius && cy >= p.brush_radius &&
                                      cx < n - p.brush_radius && cy < n - p.brush_radius;
                T scale = amount;
                if (!interior) {
                    T wsum = T(0);
                    for (const auto& b : brush) {
                        const int x = cx + b.dx, y = cy + b.dy;
                        if (x >= 0 && y >= 0 && x < n && y < n) wsum += T(b.w);
                    }
                    scale = amount / wsum;
                }
                for (const auto& b : brush) {
                    const int x = cx + b.dx, y = cy + b.dy;
                    if (!(x >= 0 && y >= 0 && x < n && y < n)) continue;
                    const size_t i = size_t(y) * n + x;
                    const T want = scale * T(b.w);
                    const T td = tmin(tmax(dirt[i], T(0)), want);
                    const T tr = tmin(tmax(rock[i], T(0)), want - td);
                    dirt[i] -= td;
                    rock[i] -= tr;
                    sed += td + tr;
                }
            }
            speed = sqrt(tmax(T(0), speed * speed - dh * T(p.gravity)));
            water = water * (T(1) - T(p.evaporation));
        }
        if (alive) deposit(px, py, sed);
    }
    return exported;
}

// ---------------------------------------------------------------------------
// The reference's own particle mode, one dispatch_particle
// (src/erosion.cpp:132-156): movement (glsl/particle.glsl:64-136), erosion
// (glsl/particle_erosion.glsl:22-128), thermal erosion of both layers, and
// smoothing with its particle branch (glsl/smoothing.glsl:22-103).  Serial:
// the movement of every particle, then the erosion of every particle in id
// order (the shader's per-pixel spinlock serialises corner updates in some
// order; id order is one of them).  WORLD_SCALE = 1, SED_LAYERS = 2.
//
// Spelled-out choices where GLSL leaves freedom: mix(a, b, t) = a + t (b - a)
// (as in the grid passes); normalize(v) = v / length(v), length = IEEE sqrt;
// texelFetch outside the texture reads 0; the `time` uniform is an explicit
// argument.  The momentum map (rgba) is four planes mx, my, mz, mw.
// ---------------------------------------------------------------------------
template <class T> struct RefParticle {
    T sc = T(0);
    int32_t iters = 0;
    T px = T(0), py = T(0), vx = T(0), vy = T(0), volume = T(0);
    T sed[2] = {T(0), T(0)};
    bool to_kill = false;
};

template <class T> struct RefParticleParams {
    uint32_t particle_count = 0, ttl = 0;
    T Kc, Kconv, Ks[2], Kd[2], Ke, G, dt, init_volume, friction, inertia, min_volume, min_velocity;
    // Map_settings_data.hmap_dims (glsl/bindings.glsl:75), which the shader
    // reads for its spawn range and border test; 0 = the grid's n.
    int dims_x = 0, dims_y = 0;
};

template <class T> inline T glsl_fract_t(T x) { return x - T(std::floor(val(x))); }

// sin of the spawn hash.  FP64: the library sin.  FP32: the method the
// device kernel uses (FP64 evaluation of an FP32 argument: reduction by a
// three-part pi/2, Taylor polynomials of sin and cos to degree 15 and 16,
// the quadrant; one rounding to FP32), restated here; pinned against FP64 sin.
inline double hash_sin(double x) { return std::sin(x); }
inline float hash_sin(float xf) {
    const double x = double(xf);
    const double k = std::rint(x * 0.6366197723675814);
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
template <class T> T ref_particle_rand(T x, T y) {
    using std::fabs;
    return glsl_fract_t(T(1e4) * hash_sin(T(17.0) * x + y * T(0.1)) * (T(0.1) + fabs(hash_sin(y * T(13.0) + x))));
}

// img_bilinear (glsl/img_interpolation.glsl:3-22) of one plane.
template <class T> T ref_bilinear(const std::vector<T>& img, int n, T x, T y) {
    const int ix = int(val(x)), iy = int(val(y));
    const T fx = glsl_fract_t(x), fy = glsl_fract_t(y);
    auto f = [&](int a, int b) -> T {
        return (a >= 0 && b >= 0 && a < n && b < n) ? img[size_t(b) * n + a] : T(0);
    };
    const T v1 = f(ix, iy) + fx * (f(ix + 1, iy) - f(ix, iy));
    const T v2 = f(ix, iy + 1) + fx * (f(ix + 1, iy + 1) - f(ix, iy + 1));
    return v1 + fy * (v2 - v1);
}

// Movement of one particle (glsl/particle.glsl:64-136).  Returns false when
// the shader returns without writing the particle back.
template <class T>
bool ref_particle_move(RefParticle<T>& pw, uint32_t id, const State<T>& s, const std::vector<T> mom[4],
                       const RefParticleParams<T>& k, T time, bool should_rain) {
    using std::sqrt;
    const int n = s.nx;
    const int dx_ = k.dims_x > 0 ? k.dims_x : n, dy_ = k.dims_y > 0 ? k.dims_y : n;   // hmap_dims
    RefParticle<T> p = pw;
    if (p.iters == 0 && !should_rain) return false;
    for (int i = 0; i < 2; ++i)
        if (val(p.sed[i]) < 0 || p.iters == 0) p.sed[i] = T(0);
    if (p.iters == 0 || p.to_kill) {
        const T ax = glsl_fract_t(time * T(1.37)) * T(1000.0), ay = glsl_fract_t(time * T(7.21)) * T(1000.0);
        p.px = ref_particle_rand(ax, T(double(id))) * T(double(dx_) - 4.0) / T(1.0) + T(2.0);
        p.py = ref_particle_rand(ay, T(double(id)) + T(3.14)) * T(double(dy_) - 4.0) / T(1.0) + T(2.0);
        p.to_kill = false;
        p.vx = T(0); p.vy = T(0);
        p.volume = k.init_volume;
        if (should_rain) {
            p.iters = 1;
        } else {
            return false;     // p.iters = 0; return (the particle is not written back)
        }
    }
    // get_terr_normal (:51-63): rock + dirt sampled one cell away
    auto rg = [&](T x, T y, T& r, T& g) { r = ref_bilinear(s.r, n, x, y); g = ref_bilinear(s.d, n, x, y); };
    T rr, rgn, lr, lg, br, bg, tr, tg;
    rg(p.px + T(1.0), p.py, rr, rgn);
    rg(p.px - T(1.0), p.py, lr, lg);
    rg(p.px, p.py - T(1.0), br, bg);
    rg(p.px, p.py + T(1.0), tr, tg);
    const T dx = rr + rgn - lr - lg;
    const T dz = tr + tg - br - bg;
    // normalize(cross((2, dx, 0), (0, dz, 2))) = normalize((2 dx, -4, 2 dz))
    const T cx = dx * T(2.0) - T(0.0) * dz, cy = T(0.0) * T(0.0) - T(2.0) * T(2.0), cz = T(2.0) * dz - dx * T(0.0);
    const T cl = sqrt(cx * cx + cy * cy + cz * cz);
    const T nx = cx / cl, ny = cy / cl, nz = cz / cl;
    const T mx = ref_bilinear(mom[0], n, p.px, p.py), my = ref_bilinear(mom[1], n, p.px, p.py);
    const T water = ref_bilinear(s.b, n, p.px, p.py);
    p.vx = p.vx - (k.dt * nx) / p.volume * k.G;                                   // :101
    p.vy = p.vy - (k.dt * nz) / p.volume * k.G;
    const T mlen = sqrt(mx * mx + my * my);
    T vlen = sqrt(p.vx * p.vx + p.vy * p.vy);
    if (val(mlen) > 0 && val(vlen) > 0) {                                         // :103-105
        const T d = (mx / mlen) * (p.vx / vlen) + (my / mlen) * (p.vy / vlen);
        const T t = k.inertia * d / (p.volume + T(1e5) * water);
        p.vx = p.vx + t * mx;
        p.vy = p.vy + t * my;
    }
    vlen = sqrt(p.vx * p.vx + p.vy * p.vy);
    if (val(vlen) > 1.0) { p.vx = p.vx / vlen; p.vy = p.vy / vlen; }             // :108-110
    const T ox = p.px, oy = p.py;
    p.px = p.px + k.dt * p.vx;                                                    // :113
    p.py = p.py + k.dt * p.vy;
    if (val(p.px) <= 1 || val(p.py) <= 1 || val(p.px * T(1.0)) >= double(dx_ - 2) ||
        val(p.py * T(1.0)) >= double(dy_ - 2)) {                                   // :114-124
        p.px = ox; p.py = oy;
        p.vx = T(0); p.vy = T(0);
        p.to_kill = true;
    }
    const T fr = T(1.0) - k.dt * k.friction * ny;                                 // :125
    p.vx = p.vx * fr;
    p.vy = p.vy * fr;
    p.volume = p.volume - k.dt * k.Ke;                                            // :126
    const T sin_a = sqrt(T(1.0) - ny * ny);                                       // :129
    vlen = sqrt(p.vx * p.vx + p.vy * p.vy);
    p.sc = tmax(T(0.0), k.Kc * p.volume * vlen * tmax(T(0.02), sin_a));         // :131
    p.iters += 1;
    if (val(p.volume) <= val(k.min_volume) || val(vlen) < val(k.min_velocity) ||
        uint32_t(p.iters) >= k.ttl)                                               // :134-139
        p.to_kill = true;
    pw = p;
    return true;
}

// Erosion of one particle at the 4 corners of its quad, in the shader's
// corner order (glsl/particle_erosion.glsl:102-128, erode_layers :22-75).
template <class T>
void ref_particle_erode(RefParticle<T>& part, State<T>& s, std::vector<T> mom[4], const RefParticleParams<T>& k) {
    if (part.iters == 0) return;
    const int n = s.nx;
    const int bx = int(val(part.px * T(1.0))), by = int(val(part.py * T(1.0)));
    const int qx[4] = {bx, bx + 1, bx + 1, bx}, qy[4] = {by, by, by + 1, by + 1};
    const T fx = glsl_fract_t(part.px * T(1.0)), fy = glsl_fract_t(part.py * T(1.0));
    const T ox[4] = {T(1.0) - fx, fx, fx, T(1.0) - fx}, oy[4] = {T(1.0) - fy, T(1.0) - fy, fy, fy};
    const T old_sed[2] = {part.sed[0], part.sed[1]};
    for (int c = 0; c < 4; ++c) {
        if (qx[c] < 0 || qy[c] < 0 || qx[c] >= n || qy[c] >= n) continue;   // outside the image: no store
        const size_t i = size_t(qy[c]) * n + qx[c];
        T terr[2] = {s.r[i], s.d[i]};
        const T multipl = ox[c] * oy[c];
        T cap = T(0.0);
        for (int l = 1; l >= 0; --l) {
            if (part.to_kill) {
                const T sed = old_sed[l] * multipl;
                terr[l] = terr[l] + sed;
                part.sed[l] = part.sed[l] - sed;
                continue;
            }
            const T Kls = k.dt * k.Ks[l], Kld = k.dt * k.Kd[l];
            const T cc = tmax(T(0.0), part.sc - cap);
            T s1 = old_sed[l];
            const T old_terr = terr[l];
            if (val(cc) > val(s1)) {
                const T eroded = multipl * Kls * (cc - s1);
                s1 = s1 + eroded;
                terr[l] = terr[l] - eroded;
                if (val(terr[l]) < 0) {
                    s1 = s1 + terr[l];
                    terr[l] = T(0);
                    cap = cap + old_terr;
                } else {
                    part.sed[l] = s1;
                    break;
                }
            } else {
                const T deposit = multipl * Kld * (s1 - cc);
                s1 = s1 - deposit;
                terr[l] = terr[l] + deposit;
            }
            part.sed[l] = s1;
        }
        const T conv = part.sed[0] * k.Kconv * k.dt;
        part.sed[1] = part.sed[1] + conv;
        part.sed[0] = part.sed[0] - conv;
        s.r[i] = terr[0];
        s.d[i] = terr[1];
        s.b[i] = s.b[i] + T(1e-5) * part.volume * multipl;
        mom[2][i] = mom[2][i] + part.volume * part.vx * multipl;
        mom[3][i] = mom[3][i] + part.volume * part.vy * multipl;
    }
}

// The particle branch of the smoothing pass (glsl/smoothing.glsl:76-95): the
// momentum map and the water display, per interior cell.  Border cells return
// before it (:29-34): their water is kept, and their momentum is never written
// (erosion only touches corners of quads inside [1, n-2], so it stays 0).
template <class T> void ref_particle_smooth_extra(State<T>& s, std::vector<T> mom[4], uint32_t particle_count) {
    using std::sqrt;
    const int n = s.nx;
    const T cnt = T(double(float(particle_count)));
    const T fm = tclamp(T(1.0) - (T(1e-12) * cnt), T(0.0), T(1.0));
    const T fb = tclamp(T(1.0) - (T(8e-8) * cnt), T(0.0), T(1.0));
    for (int y = 1; y < n - 1; ++y)
        for (int x = 1; x < n - 1; ++x) {
            const size_t i = size_t(y) * n + x;
            T mx = mom[0][i] * fm, my = mom[1][i] * fm;
            mx = mx + (T(1e-12) * cnt) * mom[2][i];
            my = my + (T(1e-12) * cnt) * mom[3][i];
            mom[2][i] = T(0); mom[3][i] = T(0);
            T b = s.b[i] * fb;
            if (val(b) < 1e-6) b = T(0);
            if (val(sqrt(mx * mx + my * my)) < 1e-12) { mx = T(0); my = T(0); }
            mom[0][i] = mx; mom[1][i] = my;
            s.b[i] = b;
        }
}

// One dispatch_particle (src/erosion.cpp:132-156): movement of every
// particle, erosion of every particle (id order), thermal erosion of both
// layers and smoothing (grid branch for rock/dirt, particle branch for water
// and momentum).  `thermal` / `smooth` switch those passes off for tests.
// The reference's per-pixel lock (glsl/particle_erosion.glsl:89-99) makes each
// particle's update of a corner one transaction, in an order the GPU picks;
// id order is one such order, and `reverse_order` (erosion in descending id
// order) another, so two runs bound how much the order alone moves a result.
template <class T>
void ref_particle_dispatch(State<T>& s, std::vector<T> mom[4], std::vector<RefParticle<T>>& parts,
                           const RefParticleParams<T>& pk, const Settings<T>& k, T time, bool should_rain,
                           bool thermal, bool smooth, bool reverse_order = false) {
    for (size_t id = 0; id < parts.size(); ++id) ref_particle_move(parts[id], uint32_t(id), s, mom, pk, time, should_rain);
    if (reverse_order)
        for (size_t id = parts.size(); id-- > 0;) ref_particle_erode(parts[id], s, mom, pk);
    else
        for (auto& p : parts) ref_particle_erode(p, s, mom, pk);
    if (thermal) {
        pass_thermal_flux(s, k, 0);
        pass_thermal_transport(s, 0);
        pass_thermal_flux(s, k, 1);
        pass_thermal_transport(s, 1);
    }
    if (smooth) {
        pass_smooth(s, k);
        ref_particle_smooth_extra(s, mom, pk.particle_count);
    }
}

}  // namespace orc
