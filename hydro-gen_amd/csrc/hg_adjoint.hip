// hg_adjoint.hip — reverse-mode gradient of a GRID rollout.
//
// No counterpart in the reference.  For a rollout of nsteps GRID steps from
// the context's state, given dL/dH_final (H = rock + dirt), computes dL/dH0
// (the initial rock height) and dL/d(Kc, Ks, Kd, Ke) (Ks, Kd shared by both
// layers).
//
// Structure:
//  * forward: the ordinary fused steps, saving the 9 state planes every
//    k_ckpt steps;
//  * reverse, one segment of k_ckpt steps at a time: recompute the segment
//    from its checkpoint (fused, saving every step's start state), then for
//    each step from last to first recompute it pass by pass with the per-pass
//    kernels into fresh buffers (so every pass input stays available) and
//    apply the pass adjoints in reverse order.
//  * each pass adjoint is a vector-Jacobian product of that pass: a thread
//    recomputes its cell's forward arithmetic through the same inline code as
//    the forward kernels (hg_cell.hpp, with a tape of the branch decisions), so
//    the reverse sweep differentiates exactly the branches the forward took;
//    clamps, max/min, the talus test and the erosion cases contribute the
//    derivative of the branch taken (a subgradient at switching points).
//    Contributions to neighbouring cells and to the semi-Lagrangian taps are
//    scattered with float atomic adds.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "hg_context.hpp"

using namespace hgctx;

namespace hgk {
namespace {

constexpr int BX = 64, BY = 4;

__device__ __forceinline__ size_t at(int x, int y, int nx) { return size_t(y) * size_t(nx) + size_t(x); }
__device__ __forceinline__ void aadd(float* p, float v) {
    if (v != 0.0f) unsafeAtomicAdd(p, v);
}

// Block reduction of up to 4 doubles, then one atomic add per value.
__device__ void reduce_coef(double* gcoef, double a, double b, double c, double d) {
    __shared__ double sh[4][BX * BY];
    const int t = threadIdx.y * BX + threadIdx.x;
    sh[0][t] = a; sh[1][t] = b; sh[2][t] = c; sh[3][t] = d;
    __syncthreads();
    for (int s = BX * BY / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int j = 0; j < 4; ++j) sh[j][t] += sh[j][t + s];
        __syncthreads();
    }
    if (t == 0)
        for (int j = 0; j < 4; ++j)
            if (sh[j][0] != 0.0) unsafeAtomicAdd(gcoef + j, sh[j][0]);
}

// ---- hydro_flux ----------------------------------------------------------
// x: flux input state; lo: adjoint of the flux outputs (water, flux, velocity);
// li: adjoint of the inputs (fresh zeroed water/flux/velocity planes, rock,
// dirt and sediment accumulated in place); lw: adjoint of the total height.
__global__ __launch_bounds__(BX * BY) void k_adj_flux(Planes x, Planes lo, Planes li, float* lw, int nx,
                                                      int ny, KParams k) {
    const int px = blockIdx.x * BX + threadIdx.x, py = blockIdx.y * BY + threadIdx.y;
    if (px >= nx || py >= ny) return;
    const size_t i = at(px, py, nx);
    const float* R = x.p[HG_P_ROCK];
    const float* D = x.p[HG_P_DIRT];
    const float* B = x.p[HG_P_WATER];
    auto total = [&](int xx, int yy) -> float {
        if (xx < 0 || yy < 0 || xx >= nx || yy >= ny) return HG_OOB_HEIGHT;
        const size_t j = at(xx, yy, nx);
        return R[j] + D[j] + B[j];
    };
    float f[4] = {x.p[HG_P_FLUX_L][i], x.p[HG_P_FLUX_R][i], x.p[HG_P_FLUX_T][i], x.p[HG_P_FLUX_B][i]};
    const float w = R[i] + D[i] + B[i];
    const float wn[4] = {total(px - 1, py), total(px + 1, py), total(px, py + 1), total(px, py - 1)};
    const float fin[4] = {px > 0 ? x.p[HG_P_FLUX_R][i - 1] : 0.0f, px < nx - 1 ? x.p[HG_P_FLUX_L][i + 1] : 0.0f,
                          py < ny - 1 ? x.p[HG_P_FLUX_B][i + nx] : 0.0f, py > 0 ? x.p[HG_P_FLUX_T][i - nx] : 0.0f};
    const unsigned wall = (px == 0 ? 1u : 0u) | (px == nx - 1 ? 2u : 0u) | (py == 0 ? 4u : 0u) |
                          (py == ny - 1 ? 8u : 0u);
    FluxTape tp;
    float vx, vy, vz;
    (void)flux_cell_t(f, B[i], w, wn, fin, wall, k, vx, vy, vz, tp);

    const float lf[4] = {lo.p[HG_P_FLUX_L][i], lo.p[HG_P_FLUX_R][i], lo.p[HG_P_FLUX_T][i], lo.p[HG_P_FLUX_B][i]};
    float ld2 = lo.p[HG_P_WATER][i];
    const float lvx = lo.p[HG_P_VEL_U][i], lvy = lo.p[HG_P_VEL_V][i];
    float lvz = lo.p[HG_P_VEL_DEPTH][i];
    float lfo[4] = {0, 0, 0, 0}, lfin[4] = {0, 0, 0, 0}, lwn[4] = {0, 0, 0, 0};
    float ld1 = 0.0f, lwp = 0.0f;
    if (vz > 0.0f) {                              // v = n * inv, inv = 1 / vz
        const float lnx = lvx * tp.inv, lny = lvy * tp.inv;
        lvz -= (lvx * vx + lvy * vy) * tp.inv;
        lfin[0] += lnx; lfo[0] -= lnx; lfo[1] += lnx; lfin[1] -= lnx;
        lfin[3] += lny; lfo[3] -= lny; lfo[2] += lny; lfin[2] -= lny;
    }
    ld1 += lvz;                                   // vz = d1 + d2
    ld2 += lvz;
    float lsum_out = 0.0f;
    if (tp.arg > 0.0f) {                          // d2 = max(0, d1 + dt (sum_in - sum_out))
        ld1 += ld2;
        const float ls = k.dt * ld2;
        for (int j = 0; j < 4; ++j) lfin[j] += ls;
        lsum_out = -ls;
    }
    const float K = tp.q < 1.0f ? tp.q : 1.0f;
    float lK = lsum_out * tp.S;                   // f = K f1, sum_out = K S
    float lf1[4];
    for (int j = 0; j < 4; ++j) {
        lK += lf[j] * tp.f1[j];
        lf1[j] = (lf[j] + lsum_out) * K;
    }
    if (tp.q < 1.0f) {                            // K = q = d1 / (S dt)
        ld1 += lK * rcp(tp.S * k.dt);
        const float lS = -lK * tp.q * rcp(tp.S);
        for (int j = 0; j < 4; ++j) lf1[j] += lS;
    }
    const float c = k.dt * k.G;                   // f1 = max(0, E f + c (w - wn))
    for (int j = 0; j < 4; ++j)
        if ((tp.act >> j) & 1u) {
            lfo[j] += k.E * lf1[j];
            lwp += c * lf1[j];
            lwn[j] -= c * lf1[j];
        }
    aadd(li.p[HG_P_FLUX_L] + i, lfo[0]);
    aadd(li.p[HG_P_FLUX_R] + i, lfo[1]);
    aadd(li.p[HG_P_FLUX_T] + i, lfo[2]);
    aadd(li.p[HG_P_FLUX_B] + i, lfo[3]);
    if (px > 0) aadd(li.p[HG_P_FLUX_R] + i - 1, lfin[0]);
    if (px < nx - 1) aadd(li.p[HG_P_FLUX_L] + i + 1, lfin[1]);
    if (py < ny - 1) aadd(li.p[HG_P_FLUX_B] + i + nx, lfin[2]);
    if (py > 0) aadd(li.p[HG_P_FLUX_T] + i - nx, lfin[3]);
    aadd(lw + i, lwp);
    if (px > 0) aadd(lw + i - 1, lwn[0]);
    if (px < nx - 1) aadd(lw + i + 1, lwn[1]);
    if (py < ny - 1) aadd(lw + i + nx, lwn[2]);
    if (py > 0) aadd(lw + i - nx, lwn[3]);
    aadd(li.p[HG_P_WATER] + i, ld1);
}

// w = rock + dirt + water: hand the total-height adjoint to the three planes.
__global__ __launch_bounds__(256) void k_adj_total(const float* lw, float* lr, float* ld, float* lb, size_t n) {
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const float v = lw[i];
    if (v != 0.0f) {
        aadd(lr + i, v);
        aadd(ld + i, v);
        aadd(lb + i, v);
    }
}

// ---- hydro_erosion -------------------------------------------------------
__global__ __launch_bounds__(BX * BY) void k_adj_erosion(Planes x, Planes lo, Planes li, int nx, int ny,
                                                         KParams k, double* gcoef) {
    const int px = blockIdx.x * BX + threadIdx.x, py = blockIdx.y * BY + threadIdx.y;
    double gKc = 0, gKs = 0, gKd = 0;
    if (px < nx && py < ny) {
        const size_t i = at(px, py, nx);
        const float* R = x.p[HG_P_ROCK];
        const float* D = x.p[HG_P_DIRT];
        auto rd = [&](int xx, int yy) -> float {
            if (xx < 0 || yy < 0 || xx >= nx || yy >= ny) return 0.0f;
            const size_t j = at(xx, yy, nx);
            return R[j] + D[j];
        };
        const float r_in = R[i], d_in = D[i], s0_in = x.p[HG_P_SED_ROCK][i], s1_in = x.p[HG_P_SED_DIRT][i];
        const float vx = x.p[HG_P_VEL_U][i], vy = x.p[HG_P_VEL_V][i], vz = x.p[HG_P_VEL_DEPTH][i];
        float r = r_in, d = d_in, s0 = s0_in, s1 = s1_in;
        ErosionTape tp;
        erosion_cell_t(r, d, s0, s1, vx, vy, vz, rd(px - 1, py), rd(px + 1, py), rd(px, py + 1), rd(px, py - 1),
                       k, tp);
        const float lr = lo.p[HG_P_ROCK][i], ld = lo.p[HG_P_DIRT][i];
        const float ls0 = lo.p[HG_P_SED_ROCK][i], ls1 = lo.p[HG_P_SED_DIRT][i];
        // conversion s1 += s0 kc, s0 -= s0 kc
        const float kc = k.Kconv * k.dt;
        const float ls0c = ls0 * (1.0f - kc) + ls1 * kc, ls1c = ls1;
        // rock layer: (r, s0, c0) -> (r', s0c)
        float lr_in = lr, ls0_in = ls0c, lc0 = 0.0f, lcap0 = 0.0f, lcap = 0.0f;
        if (tp.case0 == 0 || tp.case0 == 2) {
            const float a = k.dt * (tp.case0 == 0 ? k.Ks[0] : k.Kd[0]);
            ls0_in = lr * a + ls0c * (1.0f - a);
            lc0 = (ls0c - lr) * a;
            if (tp.case0 == 0) gKs += double(k.dt) * double(tp.c0 - s0_in) * double(ls0c - lr);
            else gKd += double(k.dt) * double(s0_in - tp.c0) * double(lr - ls0c);
        } else if (tp.case0 == 1) {               // rock exhausted: r' = 0, s0c = s0 + r
            lr_in = ls0c;
            ls0_in = ls0c;
        }
        if (tp.case0 != 3 && tp.c0 > 0.0f) {      // c0 = max(0, cap0 - cap)
            lcap0 += lc0;
            lcap -= lc0;
        }
        // dirt layer: (d, s1, c1) -> (d', s1c); cap = d when exhausted
        float ld_in = ld, ls1_in = ls1c, lc1 = 0.0f;
        if (tp.case1 == 0 || tp.case1 == 2) {
            const float a = k.dt * (tp.case1 == 0 ? k.Ks[1] : k.Kd[1]);
            ls1_in = ld * a + ls1c * (1.0f - a);
            lc1 = (ls1c - ld) * a;
            if (tp.case1 == 0) gKs += double(k.dt) * double(tp.c1 - s1_in) * double(ls1c - ld);
            else gKd += double(k.dt) * double(s1_in - tp.c1) * double(ld - ls1c);
        } else {                                  // exhausted: d' = 0, s1c = s1 + d, cap = d
            ld_in = ls1c + lcap;
            ls1_in = ls1c;
        }
        if (tp.c1 > 0.0f) lcap0 += lc1;           // c1 = max(0, cap0)
        // cap0 = Kc * max(0.02, sin_a) * ero
        const float m = fmaxf(0.02f, tp.sin_a);
        const float ero = tp.slow ? tp.ero0 * (1.0f - tp.sm) : tp.ero0;
        gKc += double(lcap0) * double(m) * double(ero);
        const float lm = lcap0 * k.Kc * ero, lero = lcap0 * k.Kc * m;
        float lhl = 0.0f, lhr = 0.0f, lht = 0.0f, lhb = 0.0f;
        if (tp.sin_a > 0.02f && tp.sin_a > 0.0f) {  // sin_a = sqrt(g2 / (g2 + 4))
            const float q = tp.g2 + 4.0f;
            const float lg2 = lm * 2.0f * rcp(tp.sin_a * q * q);
            const float ldx = 2.0f * tp.dx * lg2, ldz = 2.0f * tp.dz * lg2;
            lhr = ldx; lhl = -ldx; lht = ldz; lhb = -ldz;
        }
        float lero0 = lero, lvz = 0.0f;
        if (tp.slow) {
            lero0 = lero * (1.0f - tp.sm);
            if (tp.t_inner) {                     // sm = t^2 (3 - 2t), t = (vz - 1e-3) / (5e-4 - 1e-3)
                const float lsm = -lero * tp.ero0;
                lvz = lsm * 6.0f * tp.tt * (1.0f - tp.tt) * (1.0f / (5e-4f - 1e-3f));
            }
        }
        float lvx = 0.0f, lvy = 0.0f;
        if (tp.ero0 > 0.0f) {
            const float inv = rcp(tp.ero0);
            lvx = lero0 * vx * inv;
            lvy = lero0 * vy * inv;
        }
        aadd(li.p[HG_P_ROCK] + i, lr_in);
        aadd(li.p[HG_P_DIRT] + i, ld_in);
        li.p[HG_P_SED_ROCK][i] = ls0_in;
        li.p[HG_P_SED_DIRT][i] = ls1_in;
        li.p[HG_P_VEL_U][i] += lvx;
        li.p[HG_P_VEL_V][i] += lvy;
        li.p[HG_P_VEL_DEPTH][i] += lvz;
        auto nb = [&](int xx, int yy, float v) {
            if (xx < 0 || yy < 0 || xx >= nx || yy >= ny || v == 0.0f) return;
            const size_t j = at(xx, yy, nx);
            aadd(li.p[HG_P_ROCK] + j, v);
            aadd(li.p[HG_P_DIRT] + j, v);
        };
        nb(px - 1, py, lhl);
        nb(px + 1, py, lhr);
        nb(px, py + 1, lht);
        nb(px, py - 1, lhb);
    }
    reduce_coef(gcoef, gKc, gKs, gKd, 0.0);
}

// ---- sediment_transport --------------------------------------------------
__global__ __launch_bounds__(BX * BY) void k_adj_transport(Planes x, Planes lo, Planes li, int nx, int ny,
                                                           KParams k, double* gcoef) {
    const int px = blockIdx.x * BX + threadIdx.x, py = blockIdx.y * BY + threadIdx.y;
    double gKe = 0;
    if (px < nx && py < ny) {
        const size_t i = at(px, py, nx);
        const float bxr = float(px) - x.p[HG_P_VEL_U][i] * k.dt;
        const float byr = float(py) - x.p[HG_P_VEL_V][i] * k.dt;
        float bx, by;
        back_trace(px, py, x.p[HG_P_VEL_U][i], x.p[HG_P_VEL_V][i], k.dt, nx, ny, bx, by);
        const int ix = int(bx), iy = int(by);
        const float fx = bx - float(ix), fy = by - float(iy);
        const int ix1 = min(ix + 1, nx - 1), iy1 = min(iy + 1, ny - 1);
        const size_t j00 = at(ix, iy, nx), j10 = at(ix1, iy, nx), j01 = at(ix, iy1, nx), j11 = at(ix1, iy1, nx);
        const float l0 = lo.p[HG_P_SED_ROCK][i], l1 = lo.p[HG_P_SED_DIRT][i];
        const float w00 = (1.0f - fx) * (1.0f - fy), w10 = fx * (1.0f - fy), w01 = (1.0f - fx) * fy, w11 = fx * fy;
        aadd(li.p[HG_P_SED_ROCK] + j00, l0 * w00);
        aadd(li.p[HG_P_SED_ROCK] + j10, l0 * w10);
        aadd(li.p[HG_P_SED_ROCK] + j01, l0 * w01);
        aadd(li.p[HG_P_SED_ROCK] + j11, l0 * w11);
        aadd(li.p[HG_P_SED_DIRT] + j00, l1 * w00);
        aadd(li.p[HG_P_SED_DIRT] + j10, l1 * w10);
        aadd(li.p[HG_P_SED_DIRT] + j01, l1 * w01);
        aadd(li.p[HG_P_SED_DIRT] + j11, l1 * w11);
        float lbx = 0.0f, lby = 0.0f;
        for (int c = 0; c < 2; ++c) {
            const float* S = x.p[c == 0 ? HG_P_SED_ROCK : HG_P_SED_DIRT];
            const float a00 = S[j00], a10 = S[j10], a01 = S[j01], a11 = S[j11];
            const float l = c == 0 ? l0 : l1;
            lbx += l * ((1.0f - fy) * (a10 - a00) + fy * (a11 - a01));
            lby += l * (lerp1(a01, a11, fx) - lerp1(a00, a10, fx));
        }
        // first writer of the velocity adjoint in a step (velocity is an
        // intermediate of the step: nothing after this step reads it)
        li.p[HG_P_VEL_U][i] = (bxr > 0.0f && bxr < float(nx - 1)) ? -k.dt * lbx : 0.0f;
        li.p[HG_P_VEL_V][i] = (byr > 0.0f && byr < float(ny - 1)) ? -k.dt * lby : 0.0f;
        li.p[HG_P_VEL_DEPTH][i] = 0.0f;
        const float lb = lo.p[HG_P_WATER][i];
        li.p[HG_P_WATER][i] = lb * k.evap;                       // b' = b (1 - Ke dt)
        gKe = double(lb) * double(x.p[HG_P_WATER][i]) * double(-k.dt);
    }
    reduce_coef(gcoef, 0.0, 0.0, 0.0, gKe);
}

__device__ const int TDX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
__device__ const int TDY[8] = {0, 0, 1, -1, 1, 1, -1, -1};

// ---- thermal_erosion (outflow), layer `layer` ----------------------------
__global__ __launch_bounds__(BX * BY) void k_adj_thermal_flux(Planes x, Planes lo, Planes li, int nx, int ny,
                                                              int layer, KParams k) {
    const int px = blockIdx.x * BX + threadIdx.x, py = blockIdx.y * BY + threadIdx.y;
    if (px >= nx || py >= ny) return;
    const size_t i = at(px, py, nx);
    const float* R = x.p[HG_P_ROCK];
    const float* Dt = x.p[HG_P_DIRT];
    const float r = R[i], d = Dt[i];
    float Dk[8];
    bool ok[8];
    size_t jn[8];
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        const int xx = px + TDX[kk], yy = py + TDY[kk];
        ok[kk] = xx >= 0 && yy >= 0 && xx < nx && yy < ny;
        jn[kk] = ok[kk] ? at(xx, yy, nx) : 0;
        const float rn = ok[kk] ? R[jn[kk]] : HG_OOB_HEIGHT;
        float v = r - rn;
        if (layer == 1) v = v + (d - (ok[kk] ? Dt[jn[kk]] : HG_OOB_HEIGHT));
        Dk[kk] = v;
    }
    ThermalTape tp;
    const float g = thermal_g_t(Dk, layer == 0 ? r : d, layer, k, tp);
    if (!(tp.bk > 0.0f)) return;                  // no outflow: outputs are constant 0
    float lD[8], lg = 0.0f;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        const bool m = talus_marked(Dk[kk], kk, layer, k);
        const float lout = lo.p[HG_P_THERM_L + kk][i];
        lD[kk] = m ? lout * g : 0.0f;
        if (m) lg += lout * Dk[kk];
    }
    const float rb = rcp(tp.bk);                  // g = S / bk
    const float lS = lg * rb, lbk = -lg * g * rb;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk)
        if (talus_marked(Dk[kk], kk, layer, k)) lD[kk] += lbk;
    const float A = k.dt * k.Kspeed[layer] * 0.5f;  // S = A sharp^3 H
    const float sh3 = tp.sharp * (tp.sharp * tp.sharp);
    const float lsh3 = lS * A * tp.H, lH = lS * A * sh3;
    if (tp.sharp_active) {                        // sharp = 1 + atan(rmax) - Kalpha
        const float lrmax = lsh3 * 3.0f * tp.sharp * tp.sharp * rcp(1.0f + tp.rmax * tp.rmax);
        if (tp.mc >= tp.md * 0.70710678118654752f) lD[tp.kc] += lrmax;
        else lD[tp.kd] += lrmax * 0.70710678118654752f;
    }
    float lh_layer = 0.0f;
    if (tp.h_is_layer) lh_layer = lH;             // H = min(h_l, max(0, mc, md))
    else if (fmaxf(tp.mc, tp.md) > 0.0f) lD[tp.mc >= tp.md ? tp.kc : tp.kd] += lH;
    float lself = 0.0f;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        lself += lD[kk];
        if (ok[kk] && lD[kk] != 0.0f) {
            aadd(li.p[HG_P_ROCK] + jn[kk], -lD[kk]);
            if (layer == 1) aadd(li.p[HG_P_DIRT] + jn[kk], -lD[kk]);
        }
    }
    aadd(li.p[HG_P_ROCK] + i, lself + (layer == 0 ? lh_layer : 0.0f));
    if (layer == 1) aadd(li.p[HG_P_DIRT] + i, lself + lh_layer);
}

// ---- thermal_transport, layer `layer` (gather) ---------------------------
// h'(p) = h(p) + sum_k (T_opp(k)(p+k) - T_k(p)): T_j(q) enters q's own change
// with -1 and that of q + dir_j with +1.
__global__ __launch_bounds__(BX * BY) void k_adj_thermal_transport(Planes lo, Planes li, int nx, int ny,
                                                                   int layer) {
    const int px = blockIdx.x * BX + threadIdx.x, py = blockIdx.y * BY + threadIdx.y;
    if (px >= nx || py >= ny) return;
    const size_t i = at(px, py, nx);
    const int hp = layer == 0 ? HG_P_ROCK : HG_P_DIRT;
    const float lh = lo.p[hp][i];
    li.p[hp][i] = lh;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        const int xx = px + TDX[kk], yy = py + TDY[kk];
        const bool ok = xx >= 0 && yy >= 0 && xx < nx && yy < ny;
        li.p[HG_P_THERM_L + kk][i] = (ok ? lo.p[hp][at(xx, yy, nx)] : 0.0f) - lh;   // only reader of T_kk
    }
}

// ---- smoothing -----------------------------------------------------------
__global__ __launch_bounds__(BX * BY) void k_adj_smooth(Planes x, Planes lo, Planes li, int nx, int ny,
                                                        KParams k) {
    const int px = blockIdx.x * BX + threadIdx.x, py = blockIdx.y * BY + threadIdx.y;
    if (px >= nx || py >= ny) return;
    const size_t i = at(px, py, nx);
    const float lr = lo.p[HG_P_ROCK][i], ld = lo.p[HG_P_DIRT][i];
    if (!(px > 0 && py > 0 && px < nx - 1 && py < ny - 1)) {
        aadd(li.p[HG_P_ROCK] + i, lr);
        aadd(li.p[HG_P_DIRT] + i, ld);
        return;
    }
    const float* R = x.p[HG_P_ROCK];
    const float* D = x.p[HG_P_DIRT];
    bool er, eg;
    smooth_flags(R[i], D[i], R[i - 1], D[i - 1], R[i + 1], D[i + 1], R[i + nx], D[i + nx], R[i - nx], D[i - nx],
                 er, eg);
    const float m = k.smooth_m;                   // y = x + m (avg - x) when an extremum
    const size_t nbs[4] = {i - 1, i + 1, i + nx, i - nx};
    if (er) {
        aadd(li.p[HG_P_ROCK] + i, lr * (1.0f - 0.8f * m));
        for (int j = 0; j < 4; ++j) aadd(li.p[HG_P_ROCK] + nbs[j], lr * 0.2f * m);
    } else {
        aadd(li.p[HG_P_ROCK] + i, lr);
    }
    if (eg) {
        aadd(li.p[HG_P_DIRT] + i, ld * (1.0f - 0.8f * m));
        for (int j = 0; j < 4; ++j) aadd(li.p[HG_P_DIRT] + nbs[j], ld * 0.2f * m);
    } else {
        aadd(li.p[HG_P_DIRT] + i, ld);
    }
}

// Zeroes `count` planes of n floats (n a multiple of 4: n = side^2, side % 8 == 0).
struct PlaneList {
    float* p[48];
};
__global__ __launch_bounds__(256) void k_zero_planes(PlaneList l, int count, size_t n4) {
    const size_t total = size_t(count) * n4;
    for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += size_t(gridDim.x) * 256) {
        const size_t pl = e / n4, j = e - pl * n4;
        reinterpret_cast<float4*>(l.p[pl])[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

}  // namespace
}  // namespace hgk

namespace {

// Planes each reference pass produces (the adjoint of their input value is
// zero; their input adjoint starts from a fresh zeroed buffer).
std::vector<int> produced(int pass) {
    switch (pass) {
        case HG_PASS_RAIN: return {HG_P_WATER};
        case HG_PASS_FLUX: return {HG_P_WATER, HG_P_FLUX_L, HG_P_FLUX_R, HG_P_FLUX_T, HG_P_FLUX_B,
                                   HG_P_VEL_U, HG_P_VEL_V, HG_P_VEL_DEPTH};
        case HG_PASS_EROSION: return {HG_P_ROCK, HG_P_DIRT, HG_P_SED_ROCK, HG_P_SED_DIRT};
        case HG_PASS_TRANSPORT: return {HG_P_SED_ROCK, HG_P_SED_DIRT, HG_P_WATER};
        case HG_PASS_THERMAL_FLUX_0:
        case HG_PASS_THERMAL_FLUX_1:
            return {HG_P_THERM_L, HG_P_THERM_R, HG_P_THERM_T, HG_P_THERM_B,
                    HG_P_THERM_LT, HG_P_THERM_RT, HG_P_THERM_LB, HG_P_THERM_RB};
        case HG_PASS_THERMAL_TRANSPORT_0: return {HG_P_ROCK};
        case HG_PASS_THERMAL_TRANSPORT_1: return {HG_P_DIRT};
        case HG_PASS_SMOOTH: return {HG_P_ROCK, HG_P_DIRT};
        default: return {};
    }
}

// Whether the adjoint kernel of `pass` accumulates (with atomics) into the
// fresh buffer of produced plane `p`, which must then start at zero; the other
// fresh buffers are written with plain stores or never read.
bool needs_zero(int pass, int p) {
    switch (pass) {
        case HG_PASS_FLUX: return p == HG_P_WATER || (p >= HG_P_FLUX_L && p <= HG_P_FLUX_B);
        case HG_PASS_EROSION: return p == HG_P_ROCK || p == HG_P_DIRT;
        case HG_PASS_TRANSPORT: return p == HG_P_SED_ROCK || p == HG_P_SED_DIRT;
        case HG_PASS_SMOOTH: return true;
        default: return false;
    }
}

// Device buffers of the reverse sweep.
struct AdjWork {
    size_t pn = 0, pitch = 0;
    void* block = nullptr;
    std::vector<float*> ckpt;      // (#checkpoints) x 9 planes
    std::vector<float*> seg;       // (k_ckpt + 1) x 9 planes: each step's start state in a segment
    std::vector<float*> tape;      // k_ckpt x NTAPE planes: each step's intermediates
    std::vector<float*> pool;      // adjoint planes (free list)
    float* zero = nullptr;
    float* lw = nullptr;
    double* gcoef = nullptr;
};

// Intermediates of one step recorded by the fused kernels.
enum { T_B1, T_B2, T_S0, T_S1, T_VX, T_VY, T_VZ, T_R3, T_D3, T_R5, T_D6, NTAPE };

float* take(std::vector<float*>& pool) {
    float* p = pool.back();
    pool.pop_back();
    return p;
}

// One fused GRID step from the 9 planes `in` into `out`, recording the step's
// intermediates in `tp`: hydro writes rock/dirt after erosion to tp[T_R3/T_D3]
// (the thermal kernel then produces out's rock/dirt from them).
int step_into(hg_ctx* c, float* const* in, float* const* out, float* const* tp, uint32_t step, bool rain,
              bool thermal, bool smooth) {
    const bool th = thermal || smooth;
    hgk::HydroIO io;
    io.r = in[HG_P_ROCK]; io.d = in[HG_P_DIRT]; io.b = in[HG_P_WATER];
    io.fl = in[HG_P_FLUX_L]; io.fr = in[HG_P_FLUX_R]; io.ft = in[HG_P_FLUX_T]; io.fb = in[HG_P_FLUX_B];
    io.s0 = in[HG_P_SED_ROCK]; io.s1 = in[HG_P_SED_DIRT];
    io.ro = th ? tp[T_R3] : out[HG_P_ROCK];
    io.dout = th ? tp[T_D3] : out[HG_P_DIRT];
    io.bo = out[HG_P_WATER];
    io.flo = out[HG_P_FLUX_L]; io.fro = out[HG_P_FLUX_R]; io.fto = out[HG_P_FLUX_T]; io.fbo = out[HG_P_FLUX_B];
    io.s0o = out[HG_P_SED_ROCK]; io.s1o = out[HG_P_SED_DIRT];
    io.vx = tp[T_VX]; io.vy = tp[T_VY]; io.vz = tp[T_VZ];
    io.tb1 = tp[T_B1]; io.tb2 = tp[T_B2]; io.ts0 = tp[T_S0]; io.ts1 = tp[T_S1];
    HG_HIP(c, hgk::launch_fused_hydro(c->hydro_variant, io, c->n, c->n, step, 0u, rain, c->xcd_order, c->kp,
                                      c->stream));
    if (!th) return HG_OK;
    hgk::ThermIO tio;
    tio.r = tp[T_R3]; tio.d = tp[T_D3];
    tio.ro = out[HG_P_ROCK]; tio.dout = out[HG_P_DIRT];
    tio.tr5 = tp[T_R5]; tio.td6 = tp[T_D6];
    HG_HIP(c, hgk::launch_fused_thermal(c->thermal_variant, tio, c->n, c->n, thermal, smooth, c->xcd_order, c->kp,
                                        c->stream));
    return HG_OK;
}

int zero_planes(hg_ctx* c, const std::vector<float*>& planes, size_t pn) {
    for (size_t b = 0; b < planes.size(); b += 48) {
        hgk::PlaneList l{};
        const int cnt = int(std::min<size_t>(48, planes.size() - b));
        for (int j = 0; j < cnt; ++j) l.p[j] = planes[b + size_t(j)];
        const size_t n4 = pn / 4;
        const unsigned grid = unsigned(std::min<size_t>((size_t(cnt) * n4 + 255) / 256, 8192));
        hgk::k_zero_planes<<<grid, 256, 0, c->stream>>>(l, cnt, n4);
        HG_HIP(c, hipGetLastError());
    }
    return HG_OK;
}

}  // namespace

extern "C" int hg_adjoint_rollout(hg_ctx* c, uint32_t nsteps, uint32_t k_ckpt, uint32_t flags,
                                  const float* dL_dH_final, float* dL_dH0, double* dL_dcoef) {
    using namespace hgk;
    if (!c) return HG_ERR_INVALID;
    if (!dL_dH_final || !dL_dH0 || !dL_dcoef) return fail(c, HG_ERR_INVALID, "hg_adjoint_rollout: NULL buffer");
    if (nsteps == 0 || k_ckpt == 0 || k_ckpt > 1024)
        return fail(c, HG_ERR_INVALID, "hg_adjoint_rollout: need nsteps > 0 and 0 < k_ckpt <= 1024");
    if (flags & ~(HG_STEP_RAIN | HG_STEP_THERMAL | HG_STEP_SMOOTH))
        return fail(c, HG_ERR_INVALID, "hg_adjoint_rollout: only rain/thermal/smooth flags");
    DeviceGuard guard(c->device);
    const bool rain = (flags & HG_STEP_RAIN) && c->rain_mode != HG_RAIN_OFF;
    const bool thermal = (flags & HG_STEP_THERMAL) != 0, smooth = (flags & HG_STEP_SMOOTH) != 0;
    const int n = c->n;
    const size_t pn = plane_n(c);
    const uint32_t nck = (nsteps + k_ckpt - 1) / k_ckpt;
    AdjWork w;
    w.pn = pn;
    w.pitch = (pn * sizeof(float) + 255) / 256 * 256;
    const size_t n_pool = 64;
    const size_t nplanes = size_t(nck) * HG_NUM_STATE_PLANES + size_t(k_ckpt + 1) * HG_NUM_STATE_PLANES +
                           size_t(k_ckpt) * NTAPE + n_pool + 2;
    HG_HIP(c, hipMalloc(&w.block, nplanes * w.pitch + 256));
    struct Free {
        void* p;
        ~Free() { if (p) (void)hipFree(p); }
    } free_block{w.block};
    char* base = static_cast<char*>(w.block);
    size_t used = 0;
    auto plane = [&]() { float* p = reinterpret_cast<float*>(base + used * w.pitch); ++used; return p; };
    for (size_t i = 0; i < size_t(nck) * HG_NUM_STATE_PLANES; ++i) w.ckpt.push_back(plane());
    for (size_t i = 0; i < size_t(k_ckpt + 1) * HG_NUM_STATE_PLANES; ++i) w.seg.push_back(plane());
    for (size_t i = 0; i < size_t(k_ckpt) * NTAPE; ++i) w.tape.push_back(plane());
    for (size_t i = 0; i < n_pool; ++i) w.pool.push_back(plane());
    w.zero = plane();
    w.lw = plane();
    w.gcoef = reinterpret_cast<double*>(base + used * w.pitch);
    HG_HIP(c, hipMemsetAsync(w.zero, 0, pn * sizeof(float), c->stream));
    HG_HIP(c, hipMemsetAsync(w.gcoef, 0, 4 * sizeof(double), c->stream));
    const uint64_t step0 = c->step;

    // ---- forward with checkpoints every k_ckpt steps
    for (uint32_t s = 0; s < nsteps; ++s) {
        if (s % k_ckpt == 0)
            for (int p = 0; p < HG_NUM_STATE_PLANES; ++p)
                HG_HIP(c, hipMemcpyAsync(w.ckpt[(s / k_ckpt) * HG_NUM_STATE_PLANES + p], c->cur[p], pn * sizeof(float),
                                         hipMemcpyDeviceToDevice, c->stream));
        const int rc = fused_step(c, rain, thermal, smooth, false);
        if (rc != HG_OK) return rc;
        c->step += 1;
    }
    const uint64_t step_end = c->step;

    // ---- adjoint of the final state: dL/dH on rock and dirt
    float* lam[HG_NUM_PLANES];
    {
        std::vector<float*> z;
        for (int p = 0; p < HG_NUM_PLANES; ++p) z.push_back(lam[p] = take(w.pool));
        const int rz = zero_planes(c, z, pn);
        if (rz != HG_OK) return rz;
    }
    HG_HIP(c, hipMemcpyAsync(lam[HG_P_ROCK], dL_dH_final, pn * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HG_HIP(c, hipMemcpyAsync(lam[HG_P_DIRT], dL_dH_final, pn * sizeof(float), hipMemcpyHostToDevice, c->stream));

    std::vector<int> order;
    order.push_back(HG_PASS_FLUX);
    order.push_back(HG_PASS_EROSION);
    order.push_back(HG_PASS_TRANSPORT);
    if (thermal)
        for (int p = HG_PASS_THERMAL_FLUX_0; p <= HG_PASS_THERMAL_TRANSPORT_1; ++p) order.push_back(p);
    if (smooth) order.push_back(HG_PASS_SMOOTH);

    const dim3 blk(BX, BY), grd((n + BX - 1) / BX, (n + BY - 1) / BY);
    const unsigned g1 = unsigned((pn + 255) / 256);
    for (int seg = int(nck) - 1; seg >= 0; --seg) {
        const uint32_t s_begin = uint32_t(seg) * k_ckpt;
        const uint32_t s_end = std::min(nsteps, s_begin + k_ckpt);
        // recompute the segment from its checkpoint, recording every step's
        // start state (slot 0 is the checkpoint itself) and intermediates
        std::vector<float*>& slots = w.seg;
        for (int p = 0; p < HG_NUM_STATE_PLANES; ++p) slots[size_t(p)] = w.ckpt[size_t(seg) * HG_NUM_STATE_PLANES + p];
        for (uint32_t s = s_begin; s < s_end; ++s) {
            const size_t j = s - s_begin;
            const int rc = step_into(c, &slots[j * HG_NUM_STATE_PLANES], &slots[(j + 1) * HG_NUM_STATE_PLANES],
                                     &w.tape[j * NTAPE], uint32_t(step0 + s), rain, thermal, smooth);
            if (rc != HG_OK) return rc;
        }
        // reverse the steps of the segment
        for (int s = int(s_end) - 1; s >= int(s_begin); --s) {
            float* const* in = &slots[size_t(s - int(s_begin)) * HG_NUM_STATE_PLANES];
            float* const* tp = &w.tape[size_t(s - int(s_begin)) * NTAPE];
            // pass inputs, from the step's start state and its tape
            auto x_of = [&](int pass) {
                Planes X;
                for (int p = 0; p < HG_NUM_PLANES; ++p) X.p[p] = w.zero;
                switch (pass) {
                    case HG_PASS_FLUX:
                        X.p[HG_P_ROCK] = in[HG_P_ROCK]; X.p[HG_P_DIRT] = in[HG_P_DIRT]; X.p[HG_P_WATER] = tp[T_B1];
                        for (int p = HG_P_FLUX_L; p <= HG_P_FLUX_B; ++p) X.p[p] = in[p];
                        break;
                    case HG_PASS_EROSION:
                        X.p[HG_P_ROCK] = in[HG_P_ROCK]; X.p[HG_P_DIRT] = in[HG_P_DIRT];
                        X.p[HG_P_SED_ROCK] = in[HG_P_SED_ROCK]; X.p[HG_P_SED_DIRT] = in[HG_P_SED_DIRT];
                        X.p[HG_P_VEL_U] = tp[T_VX]; X.p[HG_P_VEL_V] = tp[T_VY]; X.p[HG_P_VEL_DEPTH] = tp[T_VZ];
                        break;
                    case HG_PASS_TRANSPORT:
                        X.p[HG_P_SED_ROCK] = tp[T_S0]; X.p[HG_P_SED_DIRT] = tp[T_S1]; X.p[HG_P_WATER] = tp[T_B2];
                        X.p[HG_P_VEL_U] = tp[T_VX]; X.p[HG_P_VEL_V] = tp[T_VY]; X.p[HG_P_VEL_DEPTH] = tp[T_VZ];
                        break;
                    case HG_PASS_THERMAL_FLUX_0:
                        X.p[HG_P_ROCK] = tp[T_R3]; X.p[HG_P_DIRT] = tp[T_D3];
                        break;
                    case HG_PASS_THERMAL_FLUX_1:
                        X.p[HG_P_ROCK] = tp[T_R5]; X.p[HG_P_DIRT] = tp[T_D3];
                        break;
                    case HG_PASS_SMOOTH:
                        X.p[HG_P_ROCK] = tp[T_R5]; X.p[HG_P_DIRT] = tp[T_D6];
                        break;
                    default: break;
                }
                return X;
            };
            // fresh adjoint buffers of every pass output, taken up front; the
            // ones that accumulate are zeroed by one launch
            std::vector<std::vector<float*>> fresh_of(order.size());
            std::vector<float*> to_zero;
            for (size_t oi = 0; oi < order.size(); ++oi) {
                for (int p : produced(order[oi])) {
                    float* f = take(w.pool);
                    fresh_of[oi].push_back(f);
                    if (needs_zero(order[oi], p)) to_zero.push_back(f);
                }
                if (order[oi] == HG_PASS_FLUX) to_zero.push_back(w.lw);
            }
            int rc = zero_planes(c, to_zero, pn);
            if (rc != HG_OK) return rc;
            // pass adjoints in reverse order (rain, water += rain, is the identity)
            for (int oi = int(order.size()) - 1; oi >= 0; --oi) {
                const int pass = order[oi];
                const Planes X = x_of(pass);
                const Planes lo = planes_of(lam);
                float* fresh[HG_NUM_PLANES];
                std::memcpy(fresh, lam, sizeof(fresh));
                const std::vector<int> prod = produced(pass);
                for (size_t q = 0; q < prod.size(); ++q) fresh[prod[q]] = fresh_of[oi][q];
                const Planes li = planes_of(fresh);
                switch (pass) {
                    case HG_PASS_FLUX:
                        k_adj_flux<<<grd, blk, 0, c->stream>>>(X, lo, li, w.lw, n, n, c->kp);
                        HG_HIP(c, hipGetLastError());
                        k_adj_total<<<g1, 256, 0, c->stream>>>(w.lw, fresh[HG_P_ROCK], fresh[HG_P_DIRT],
                                                                fresh[HG_P_WATER], pn);
                        break;
                    case HG_PASS_EROSION:
                        k_adj_erosion<<<grd, blk, 0, c->stream>>>(X, lo, li, n, n, c->kp, w.gcoef);
                        break;
                    case HG_PASS_TRANSPORT:
                        k_adj_transport<<<grd, blk, 0, c->stream>>>(X, lo, li, n, n, c->kp, w.gcoef);
                        break;
                    case HG_PASS_THERMAL_FLUX_0:
                    case HG_PASS_THERMAL_FLUX_1:
                        k_adj_thermal_flux<<<grd, blk, 0, c->stream>>>(X, lo, li, n, n,
                                                                        pass == HG_PASS_THERMAL_FLUX_0 ? 0 : 1, c->kp);
                        break;
                    case HG_PASS_THERMAL_TRANSPORT_0:
                    case HG_PASS_THERMAL_TRANSPORT_1:
                        k_adj_thermal_transport<<<grd, blk, 0, c->stream>>>(lo, li, n, n,
                                                                             pass == HG_PASS_THERMAL_TRANSPORT_0 ? 0 : 1);
                        break;
                    case HG_PASS_SMOOTH:
                        k_adj_smooth<<<grd, blk, 0, c->stream>>>(X, lo, li, n, n, c->kp);
                        break;
                    default: break;
                }
                HG_HIP(c, hipGetLastError());
                for (int p : prod) {
                    w.pool.push_back(lam[p]);
                    lam[p] = fresh[p];
                }
            }
        }
    }
    c->step = step_end;
    HG_HIP(c, hipMemcpyAsync(dL_dH0, lam[HG_P_ROCK], pn * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HG_HIP(c, hipMemcpyAsync(dL_dcoef, w.gcoef, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HG_HIP(c, hipStreamSynchronize(c->stream));
    return HG_OK;
}
