// hg_cell.hpp — per-cell arithmetic of the erosion passes, shared by the
// per-pass kernels and the fused kernels so that both produce bit-identical
// results for the same inputs.
//
// Semantics follow the reference GLSL shaders (ger0/hydro-gen, glsl/*.glsl);
// each function cites the lines it restates.  Where the code is arranged for
// fewer instructions than the literal shader, the arithmetic differs from the
// shader's FP32 rounding only (never in meaning):
//  * slope: sin_a = sqrt((dx^2+dz^2)/(dx^2+dz^2+4)) is the closed form of
//    |cross((2,dx,0),(0,dz,2))| normalisation (hydro_erosion.glsl:23-35,57);
//  * talus test: D/d > tan(Kalpha) replaces atan(D/d) > Kalpha, and the
//    sharpness 1 + atan(max marked D/d) - Kalpha takes one atan per layer
//    instead of one per direction (thermal_erosion.glsl:63-91), since atan is
//    monotonic;
//  * thermal outflow S*D_k/bk is formed as (S/bk)*D_k, so a cell's outflows
//    are described by one scalar g = S/bk and its height differences;
//  * mix(a,b,t) = a*(1-t)+b*t is evaluated as a + t*(b-a);
//  * divisions by a per-cell value use the hardware reciprocal (1 ulp), square
//    roots the hardware square root (1 ulp), and
//    the thermal height change is accumulated per direction,
//    h += sum_k (in_k - out_k), instead of -sum(out) + sum(in).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hgk {

// Out-of-grid total height, FP32 literal 999999999999.0
// (hydro_flux.glsl:33-39, thermal_erosion.glsl:20-26).
#define HG_OOB_HEIGHT 999999999999.0f

// Kernel-side copy of the erosion coefficients with the per-layer thermal
// thresholds precomputed on the host.
struct KParams {
    float Kc, Kconv, Ke, E, G, dt;
    float Ks[2], Kd[2], Kalpha[2], Kspeed[2];
    float tan_alpha[2];       // tan(Kalpha[l])
    float thr_diag[2];        // sqrt(2) * tan(Kalpha[l])
    float smooth_m;           // clamp(Kspeed[1]*dt, 0, 1)
    float evap;               // 1 - Ke*dt
    float rain_amount;
    int rain_mode;            // HG_RAIN_*
    uint32_t seed_lo, seed_hi;
    // reference rain (HG_RAIN_REFERENCE): Rain_data fields, the explicit time
    // seed and the map's max height
    float rr_thresh, rr_multip, rr_drops, rr_time, rr_max_h;
    // fused kernels write their outputs with non-temporal stores (the step's
    // working set exceeds the last-level cache; see stream_stores)
    int store_nt;
};

// ---------------------------------------------------------------------------
// Reference rain (glsl/rain.glsl:32-56): simplex noise gln_simplex
// (glsl/simplex_noise.glsl:377-402), its fBm gln_sfbm (:419-456) and the rain
// increment with the mountain bonus.  FP32 with the GLSL definitions
// mod(x, y) = x - y floor(x / y) and fract(x) = x - floor(x); pow(result, 1)
// of gln_sfbm is taken as result.  The order of every operation follows the
// shader, so host (FP32) and device agree bit for bit.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ float glsl_mod(float x, float y) { return x - y * floorf(x / y); }
__host__ __device__ __forceinline__ float glsl_fract(float x) { return x - floorf(x); }

__device__ __forceinline__ float ref_simplex(float vx, float vy) {
    const float Cx = 0.211324865405187f, Cy = 0.366025403784439f, Cz = -0.577350269189626f,
                Cw = 0.024390243902439f;
    const float d = vx * Cy + vy * Cy;
    float ix = floorf(vx + d), iy = floorf(vy + d);
    const float di = ix * Cx + iy * Cx;
    const float x0x = (vx - ix) + di, x0y = (vy - iy) + di;
    const float i1x = x0x > x0y ? 1.0f : 0.0f, i1y = 1.0f - i1x;
    const float x1x = (x0x + Cx) - i1x, x1y = (x0y + Cx) - i1y;
    const float x2x = x0x + Cz, x2y = x0y + Cz;
    ix = glsl_mod(ix, 289.0f);
    iy = glsl_mod(iy, 289.0f);
    auto rnd = [](float q) { return glsl_mod((q * 34.0f + 1.0f) * q, 289.0f); };
    const float p[3] = {rnd(rnd(iy) + ix), rnd((rnd(iy + i1y) + ix) + i1x), rnd((rnd(iy + 1.0f) + ix) + 1.0f)};
    const float gx[3] = {x0x, x1x, x2x}, gy[3] = {x0y, x1y, x2y};
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float q = 0.5f - (gx[c] * gx[c] + gy[c] * gy[c]);
        float m = q < 0.0f ? 0.0f : q;
        m = m * m;
        m = m * m;
        const float x = 2.0f * glsl_fract(p[c] * Cw) - 1.0f;
        const float h = fabsf(x) - 0.5f;
        const float a0 = x - floorf(x + 0.5f);
        m = m * (1.79284291400159f - 0.85373472095314f * (a0 * a0 + h * h));
        acc = acc + m * (a0 * gx[c] + h * gy[c]);
    }
    return 130.0f * acc;
}

// Rain increment of cell (x, y) with total height w (rock + dirt + water).
__device__ __forceinline__ float ref_rain_incr(int x, int y, float w, const KParams& k) {
    const float seed = glsl_fract(k.rr_time * 1.372914227e3f) * 1000.0f;
    float vx = float(x) + seed * 100.0f, vy = float(y) + seed * 100.0f;
    float result = 0.0f, amplitude = 1.0f, frequency = 1.0f, maximum = amplitude;
    for (int o = 0; o < 8; ++o) {
        result = result + ref_simplex((vx * frequency) * k.rr_drops, (vy * frequency) * k.rr_drops) * amplitude;
        frequency = frequency * 2.0f;
        amplitude = amplitude * 0.5f;
        maximum = maximum + amplitude;
    }
    const float sf = result / maximum;
    const float r = sf < 0.0f ? 0.0f : sf;
    float incr = k.rain_amount * r;
    const float mountain = w - k.rr_max_h * k.rr_thresh;
    if (mountain > 0.0f) incr = incr + ((mountain * k.rr_multip) * r) / ((1.0f - k.rr_thresh) * k.rr_max_h);
    return incr;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).
// ---------------------------------------------------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = uint32_t(p1);
        uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = uint32_t(p0);
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__host__ __device__ inline float u24(uint32_t w) { return float(w >> 8) * (1.0f / 16777216.0f); }

// Hardware reciprocal (v_rcp_f32, 1 ulp); rcp(0) = +inf.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// Square root of the erosion pass: the hardware v_sqrt_f32 (1 ulp), as GLSL's
// sqrt is not correctly rounded either.
__device__ __forceinline__ float hsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// ---------------------------------------------------------------------------
// hydro_flux (glsl/hydro_flux.glsl:77-166) for one cell.
//  f[4]      : own old flux (L,R,T,B), replaced by the new flux
//  d1        : own water (after rain)
//  w         : own total height, wn[4] : neighbour totals (OOB = HG_OOB_HEIGHT)
//  fin[4]    : neighbours' old flux pointing at this cell
//              (F(p+L).R, F(p+R).L, F(p+T).B, F(p+B).T), 0 out of grid
//  wall      : bit0 x==0, bit1 x==N-1, bit2 y==0, bit3 y==N-1
// Returns the new water d2; writes velocity (vx, vy, vz).
// ---------------------------------------------------------------------------
// Intermediates of flux_cell kept for the adjoint.
struct FluxTape {
    float f1[4];      // outflow after the max(0, .) and the walls, before K
    unsigned act;     // bit i: outflow i was positive (not clamped, not a wall)
    float S;          // sum of f1
    float q;          // d1 / (S*dt); K = min(1, q)
    float arg;        // d1 + dt*(sum_in - sum_out), d2 = max(0, arg)
    float inv;        // 1 / (d1 + d2) when d1 + d2 > 0
};

__device__ __forceinline__ float flux_cell_t(float f[4], float d1, float w, const float wn[4],
                                             const float fin[4], unsigned wall, const KParams& k,
                                             float& vx, float& vy, float& vz, FluxTape& tp) {
    const float fo0 = f[0], fo1 = f[1], fo2 = f[2], fo3 = f[3];
    const float c = k.dt * k.G;  // dt*A*G/L with A = L = 1
    tp.act = 0u;
#pragma unroll
    for (int i = 0; i < 4; ++i) {                                                     // :100-107
        const float f0 = k.E * f[i] + c * (w - wn[i]);
        f[i] = fmaxf(0.0f, f0);
        if (f0 > 0.0f) tp.act |= 1u << i;
    }
    if (wall & 1u) { f[0] = 0.0f; tp.act &= ~1u; }                                   // :110-119
    else if (wall & 2u) { f[1] = 0.0f; tp.act &= ~2u; }
    if (wall & 4u) { f[3] = 0.0f; tp.act &= ~8u; }
    else if (wall & 8u) { f[2] = 0.0f; tp.act &= ~4u; }
#pragma unroll
    for (int i = 0; i < 4; ++i) tp.f1[i] = f[i];
    const float sum_in = (fin[0] + fin[1]) + (fin[2] + fin[3]);
    float sum_out = (f[0] + f[1]) + (f[2] + f[3]);
    tp.S = sum_out;
    // K = min(1, b/(sum_out*dt)); sum_out == 0 gives K = 1 (:125)
    const float q = d1 * rcp(sum_out * k.dt);
    tp.q = q;
    const float K = (q < 1.0f) ? q : 1.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) f[i] *= K;
    sum_out *= K;
    tp.arg = d1 + k.dt * (sum_in - sum_out);
    const float d2 = fmaxf(0.0f, tp.arg);                                             // :133-134
    vz = d1 + d2;                                                                      // :144
    tp.inv = 0.0f;
    if (vz > 0.0f) {                                                                   // :146-161
        const float inv = rcp(vz);
        tp.inv = inv;
        vx = ((fin[0] - fo0) + (fo1 - fin[1])) * inv;
        vy = ((fin[3] - fo3) + (fo2 - fin[2])) * inv;
    } else {
        vx = 0.0f;
        vy = 0.0f;
    }
    return d2;
}

__device__ __forceinline__ float flux_cell(float f[4], float d1, float w, const float wn[4],
                                           const float fin[4], unsigned wall, const KParams& k,
                                           float& vx, float& vy, float& vz) {
    FluxTape tp;
    return flux_cell_t(f, d1, w, wn, fin, wall, k, vx, vy, vz, tp);
}

// ---------------------------------------------------------------------------
// hydro_erosion (glsl/hydro_erosion.glsl:37-92) for one cell.
// hl, hr, ht, hb: rock+dirt of the L, R, T, B neighbours (0 out of grid).
// ---------------------------------------------------------------------------
// Intermediates of erosion_cell kept for the adjoint.  Layer cases:
// 0 dissolve, 1 dissolve that exhausted the layer, 2 deposit, 3 not reached.
struct ErosionTape {
    float ero0, sm, tt, sin_a, g2, dx, dz, cap0;
    float c1, c0;          // transport capacity seen by dirt and rock
    float s1c, s0c;        // sediment before the rock->dirt conversion
    int case1, case0;
    bool slow, t_inner;    // shallow-water damping applied; its t strictly in (0,1)
};

__device__ __forceinline__ void erosion_cell_t(float& r, float& d, float& s0, float& s1, float vx,
                                               float vy, float vz, float hl, float hr, float ht,
                                               float hb, const KParams& k, ErosionTape& tp) {
    float ero = hsqrt(vx * vx + vy * vy);                                              // :45-52
    tp.ero0 = ero;
    tp.slow = vz < 1e-3f;
    tp.sm = 0.0f;
    tp.tt = 0.0f;
    tp.t_inner = false;
    if (tp.slow) {
        const float dd = fmaxf(5e-4f, vz);
        float t = (dd - 1e-3f) * (1.0f / (5e-4f - 1e-3f));
        tp.t_inner = t > 0.0f && t < 1.0f && vz > 5e-4f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        tp.tt = t;
        tp.sm = t * t * (3.0f - 2.0f * t);
        ero = ero * (1.0f - tp.sm);
    }
    const float dx = hr - hl, dz = ht - hb;                                            // :23-35
    const float g2 = dx * dx + dz * dz;
    const float sin_a = hsqrt(g2 * rcp(g2 + 4.0f));                                    // :57
    tp.dx = dx; tp.dz = dz; tp.g2 = g2; tp.sin_a = sin_a;
    const float cap0 = k.Kc * fmaxf(0.02f, sin_a) * ero;
    tp.cap0 = cap0;
    // layer loop, dirt (1) then rock (0), :58-83
    float cap = 0.0f;
    bool stop = false;
    {
        const float c = fmaxf(0.0f, cap0 - cap);
        tp.c1 = c;
        if (c > s1) {
            const float old = d;
            const float delta = (k.dt * k.Ks[1]) * (c - s1);
            d -= delta;
            s1 += delta;
            if (d < 0.0f) { s1 += d; d = 0.0f; cap += old; tp.case1 = 1; }
            else { stop = true; tp.case1 = 0; }
        } else {
            const float delta = (k.dt * k.Kd[1]) * (s1 - c);
            d += delta;
            s1 -= delta;
            tp.case1 = 2;
        }
    }
    tp.case0 = 3;
    tp.c0 = 0.0f;
    if (!stop) {
        const float c = fmaxf(0.0f, cap0 - cap);
        tp.c0 = c;
        if (c > s0) {
            const float delta = (k.dt * k.Ks[0]) * (c - s0);
            r -= delta;
            s0 += delta;
            if (r < 0.0f) { s0 += r; r = 0.0f; tp.case0 = 1; }
            else tp.case0 = 0;
        } else {
            const float delta = (k.dt * k.Kd[0]) * (s0 - c);
            r += delta;
            s0 -= delta;
            tp.case0 = 2;
        }
    }
    tp.s0c = s0;
    tp.s1c = s1;
    const float conv = s0 * (k.Kconv * k.dt);                                          // :84-88
    s1 += conv;
    s0 -= conv;
}

__device__ __forceinline__ void erosion_cell(float& r, float& d, float& s0, float& s1, float vx,
                                             float vy, float vz, float hl, float hr, float ht,
                                             float hb, const KParams& k) {
    ErosionTape tp;
    erosion_cell_t(r, d, s0, s1, vx, vy, vz, hl, hr, ht, hb, k, tp);
}

// Semi-Lagrangian back-trace (sediment_transport.glsl:70 and :26-30): returns
// the clamped sample position.
__device__ __forceinline__ void back_trace(int x, int y, float vx, float vy, float dt, int nx, int ny,
                                           float& bx, float& by) {
    bx = fminf(fmaxf(float(x) - vx * dt, 0.0f), float(nx - 1));
    by = fminf(fmaxf(float(y) - vy * dt, 0.0f), float(ny - 1));
}

__device__ __forceinline__ float lerp1(float a, float b, float t) { return fmaf(t, b - a, a); }

// ---------------------------------------------------------------------------
// Thermal erosion (glsl/thermal_erosion.glsl:28-109) for one cell and layer.
// D[8]: cumulative layer-top differences to the 8 neighbours in the order
// L, R, T, B, LT, RT, LB, RB (OOB neighbours give a very negative D).
// Returns g = S/bk (0 when no direction is marked); out_k = marked_k ? g*D_k : 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool talus_marked(float D, int k, int layer, const KParams& p) {
    return D > (k < 4 ? p.tan_alpha[layer] : p.thr_diag[layer]);
}

// Intermediates of thermal_g kept for the adjoint.
struct ThermalTape {
    float bk, mc, md, H, rmax, sharp, S;
    int kc, kd;            // first index of the largest cross / diagonal D
    bool h_is_layer;       // H = h_l (else H = max(0, mc, md))
    bool sharp_active;     // 1 + atan(rmax) - Kalpha > 1
};

__device__ __forceinline__ float thermal_g_t(const float D[8], float h_layer, int layer,
                                             const KParams& p, ThermalTape& tp) {
    const float ta = p.tan_alpha[layer], td = p.thr_diag[layer];
    float bk = 0.0f, mc = D[0], md = D[4];
    int kc = 0, kd = 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        bk += D[k] > ta ? D[k] : 0.0f;
        if (D[k] > mc) { mc = D[k]; kc = k; }
    }
#pragma unroll
    for (int k = 4; k < 8; ++k) {
        bk += D[k] > td ? D[k] : 0.0f;
        if (D[k] > md) { md = D[k]; kd = k; }
    }
    tp.bk = bk; tp.mc = mc; tp.md = md; tp.kc = kc; tp.kd = kd;
    if (!(bk > 0.0f)) return 0.0f;
    // H = min(h_l, max(0, max_k D_k))                                                  :46-58
    const float M = fmaxf(fmaxf(mc, md), 0.0f);
    tp.h_is_layer = h_layer < M;
    const float H = fminf(h_layer, M);
    // The largest marked D_k/d_k is the largest D_k/d_k overall when any direction
    // is marked (unmarked cross D <= tan a < marked diagonal D/sqrt2 and vice versa).
    const float rmax = fmaxf(mc, md * 0.70710678118654752f);
    const float sharp0 = 1.0f + atanf(rmax) - p.Kalpha[layer];                         // :75-83
    tp.sharp_active = sharp0 > 1.0f;
    float sharp = fmaxf(sharp0, 1.0f);
    tp.sharp = sharp;
    sharp = sharp * (sharp * sharp);                                                   // :92
    const float S = (p.dt * p.Kspeed[layer] * 0.5f) * sharp * H;                       // :95
    tp.H = H; tp.rmax = rmax; tp.S = S;
    return S * rcp(bk);
}

// thermal_g_t without the tape: the largest cross and diagonal differences by
// max3 instead of a tracked compare-and-select.  Both forms give the same g:
// max values are equal (no NaN arises from finite D), and when bk > 0 they
// are > 0, so the sign of a zero maximum cannot reach the result.
__device__ __forceinline__ float thermal_g(const float D[8], float h_layer, int layer,
                                           const KParams& p) {
    const float ta = p.tan_alpha[layer], td = p.thr_diag[layer];
    float bk = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) bk += D[k] > ta ? D[k] : 0.0f;
#pragma unroll
    for (int k = 4; k < 8; ++k) bk += D[k] > td ? D[k] : 0.0f;
    const float mc = fmaxf(fmaxf(D[0], D[1]), fmaxf(D[2], D[3]));
    const float md = fmaxf(fmaxf(D[4], D[5]), fmaxf(D[6], D[7]));
    const float H = fminf(h_layer, fmaxf(fmaxf(mc, md), 0.0f));
    const float rmax = fmaxf(mc, md * 0.70710678118654752f);
    float sharp = fmaxf(1.0f + atanf(rmax) - p.Kalpha[layer], 1.0f);
    sharp = sharp * (sharp * sharp);
    const float S = (p.dt * p.Kspeed[layer] * 0.5f) * sharp * H;
    // Evaluated whatever bk is, with no branch: m = [bk > 0] gives
    // (S*1)*rcp(bk+0) = S*rcp(bk) when bk > 0, and (S*0)*rcp(1) = 0 otherwise
    // (S is finite for finite D).
    const float m = bk > 0.0f ? 1.0f : 0.0f;
    return (S * m) * rcp(bk + (1.0f - m));
}

// Height change of cell p in direction k (thermal_transport.glsl:146-172):
// in_k - out_k, where out_k = g_p*D_k when p marks k and in_k = g_n*(-D_k) when
// the neighbour n = p+k marks the opposite direction (its difference to p is
// exactly -D_k).  At most one of the two is non-zero since the threshold is >= 0.
__device__ __forceinline__ float thermal_term(float Dk, float g_self, float g_nb, float thr) {
    const float G = Dk > thr ? g_self : (-Dk > thr ? g_nb : 0.0f);
    return (-Dk) * G;
}

// ---------------------------------------------------------------------------
// Smoothing, grid branch (glsl/smoothing.glsl:36-74,97-103), interior cell.
// ---------------------------------------------------------------------------
// Extremum tests of the smoothing pass for the rock (er) and surface (eg) channels.
__device__ __forceinline__ void smooth_flags(float r, float d, float lr, float lg, float rr, float rg,
                                             float tr, float tg, float br, float bg, bool& er,
                                             bool& eg) {
    const float dlr = r - lr, dlg = (d - lg) + dlr;
    const float drr = r - rr, drg = (d - rg) + drr;
    const float dtr = r - tr, dtg = (d - tg) + dtr;
    const float dbr = r - br, dbg = (d - bg) + dbr;
    const float gh = fabsf(((dlg + drg) + (dtg + dbg)) * 0.25f);
    const float rh = fabsf(((dlr + drr) + (dtr + dbr)) * 0.25f);
    er = ((-dlr > rh || -drr > rh) && dlr * drr > 0.0f) || ((-dtr > rh || -dbr > rh) && dtr * dbr > 0.0f);
    eg = ((-dlg > gh || -drg > gh) && dlg * drg > 0.0f) || ((-dtg > gh || -dbg > gh) && dtg * dbg > 0.0f);
}

__device__ __forceinline__ void smooth_cell(float& r, float& d, float lr, float lg, float rr,
                                            float rg, float tr, float tg, float br, float bg,
                                            float m) {
    bool er, eg;
    smooth_flags(r, d, lr, lg, rr, rg, tr, tg, br, bg, er, eg);
    const float nr = er ? (r + lr + rr + tr + br) * 0.2f : r;
    const float ng = eg ? (d + lg + rg + tg + bg) * 0.2f : d;
    r = fmaf(m, nr - r, r);
    d = fmaf(m, ng - d, d);
}

}  // namespace hgk
