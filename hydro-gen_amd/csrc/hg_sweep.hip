// hg_sweep.hip — the whole GRID step in one kernel.
//
// A workgroup owns a strip of W columns and a segment of R rows of the grid
// and sweeps it from the bottom row to the top, H rows per iteration.  Each
// iteration runs every pass of the step on a block of H rows, each pass a
// fixed number of rows behind the one it depends on:
//
//   phase  pass (reference shader)                          rows behind  columns
//     0    load the 9 state planes (+ rain)                      0      -8 .. W+8
//     1    hydro_flux + hydro_erosion                             1      -5 .. W+5
//     2    thermal_erosion layer 0 (outflow scale g)              2      -4 .. W+4
//          sediment_transport (back-trace + bilinear gather)    1+HA     0 .. W
//     3    thermal_transport layer 0 (rock)                       3      -3 .. W+3
//     4    thermal_erosion layer 1                                4      -2 .. W+2
//     5    thermal_transport layer 1 (dirt)                       5      -1 .. W+1
//     6    smoothing, write rock and dirt                         6       0 .. W
//
// Every pass reads its inputs one cell around, so a pass one row behind its
// producer (after a barrier) finds them computed.  Each intermediate lives in
// an LDS ring of the rows its consumers still read; the state planes are read
// once (72 B/cell with the writes) apart from the halo columns shared with the
// neighbouring strips and the 12 rows shared with the neighbouring segments.
// The loads of the next iteration's rows are issued before the current
// iteration's passes run and stored to LDS at its start.
//
// The per-cell arithmetic is the same inline code as the per-pass and
// two-kernel paths (hg_cell.hpp), evaluated in the same order on the same
// inputs, so all three agree bit for bit.  Back-traces whose bilinear taps
// leave the LDS window recompute the eroded sediment of the taps from the
// step's input planes (eroded_sediment_global), which is exact.
#include "hg_kernels.hpp"

namespace hgk {

namespace {

__device__ __forceinline__ size_t gat(int x, int y, int n) { return size_t(y) * size_t(n) + size_t(x); }

template <bool NTS>
__device__ __forceinline__ void st_out(float* p, float v) {
    if constexpr (NTS)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <int W_, int H_, int NT_, int HA_>
struct Sweep {
    static constexpr int W = W_, H = H_, NT = NT_, HA = HA_;
    static constexpr int XO = 8;                // LDS column of strip column 0
    static constexpr int WL = W + 2 * XO;       // LDS row width
    static constexpr int QW = WL / 4;           // quads per loaded row
    // columns each side evaluated by a pass
    static constexpr int SF = 5, SG0 = 4, SU0 = 3, SG1 = 2, SU1 = 1;
    // rows behind the load front
    static constexpr int LF = 1, LG0 = 2, LU0 = 3, LG1 = 4, LU1 = 5, LSM = 6, LTR = LF + HA;
    // ring sizes (rows): producer lag p, deepest consumer lag c reading one row
    // below its block: H + c + 1 - p
    static constexpr int RIN = H + 2;           // inputs, read by flux/erosion
    static constexpr int RR1 = H + 3;           // eroded rock, read by the two layer-0 passes
    static constexpr int RD1 = H + 5;           // eroded dirt, read by the two layer-1 passes
    static constexpr int RSE = H + 2 * HA;      // eroded sediment, read by the transport window
    static constexpr int RV = H + HA;           // velocity, read by the transport (own cell)
    static constexpr int RG = H + 2;            // g of a layer, read by its transport
    static constexpr int RR2 = H + 4;           // rock after layer 0, read by layer 1 and smoothing
    static constexpr int RDD = H + 2;           // dirt after layer 1, read by smoothing
    // input planes in LDS
    enum { I_R, I_D, I_RD, I_B, I_FL, I_FR, I_FT, I_FB, I_S0, I_S1, NIN };
    static constexpr int O_IN = 0;
    static constexpr int O_R1 = O_IN + NIN * RIN * WL;
    static constexpr int O_D1 = O_R1 + RR1 * WL;
    static constexpr int O_E0 = O_D1 + RD1 * WL;
    static constexpr int O_E1 = O_E0 + RSE * WL;
    static constexpr int O_G0 = O_E1 + RSE * WL;
    static constexpr int O_R2 = O_G0 + RG * WL;
    static constexpr int O_G1 = O_R2 + RR2 * WL;
    static constexpr int O_DD = O_G1 + RG * WL;
    static constexpr int O_VX = O_DD + RDD * WL;
    static constexpr int O_VY = O_VX + RV * W;
    static constexpr int LDS_FLOATS = O_VY + RV * W;
    // load items: (plane group, row, quad); groups: rock+dirt, water, flux L/R,
    // flux T/B, sediment
    static constexpr int NGRP = 5, NLOAD = NGRP * H * QW, LOAD_IT = (NLOAD + NT - 1) / NT;
    static_assert(H * (W + 2 * SF) <= NT, "one flux/erosion cell per thread");
    static_assert(W % 4 == 0, "strip width must be a multiple of 4");
};

// ring slot of row y (y >= -64)
template <int RS>
__device__ __forceinline__ int slot(int y) { return (y + 64 * RS) % RS; }

struct SweepArgs {
    HydroIO io;
    int n, seg_rows, nstrips, nsegs;
    uint32_t step, terrain;
    int rain, do_thermal, do_smooth;
};

template <int W, int H, int NT, int HA, bool NTS>
__global__ __launch_bounds__(NT) void k_step_sweep(SweepArgs a, KParams k) {
    using S = Sweep<W, H, NT, HA>;
    __shared__ float lds[S::LDS_FLOATS];
    const HydroIO& io = a.io;
    const int n = a.n;
    int tid = threadIdx.x;
    // strip and segment of this workgroup: with 8 | strips, the 8 XCDs (which
    // the hardware deals workgroups to round robin) each own a run of adjacent
    // strips, so the halo columns two strips share are read through one L2
    int strip, seg;
    {
        const int b = blockIdx.x;
        if ((a.nstrips & 7) == 0) {
            const int per = a.nstrips >> 3, x = b & 7, j = b >> 3;
            strip = x * per + j % per;
            seg = j / per;
        } else {
            strip = b % a.nstrips;
            seg = b / a.nstrips;
        }
    }
    const int x0 = strip * W, y0 = seg * a.seg_rows;
    const int R = min(a.seg_rows, n - y0);
    const int lx0 = x0 - S::XO;                       // grid x of LDS column 0
    float* const in = lds + S::O_IN;
    auto inp = [&](int plane, int y, int lc) -> float& { return in[(plane * S::RIN + slot<S::RIN>(y)) * S::WL + lc]; };
    auto ring = [&](int off, int rs, int y, int lc) -> float& { return lds[off + ((y + 64 * rs) % rs) * S::WL + lc]; };
    auto in_grid = [&](int gx, int gy) { return gx >= 0 && gy >= 0 && gx < n && gy < n; };

    // ---- prefetch registers of the load items --------------------------------
    float4 pa[S::LOAD_IT], pb[S::LOAD_IT];
    auto issue_loads = [&](int front) {
#pragma unroll
        for (int it = 0; it < S::LOAD_IT; ++it) {
            const int q = tid + it * NT;
            pa[it] = make_float4(0.f, 0.f, 0.f, 0.f);
            pb[it] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (S::NLOAD % NT != 0 && q >= S::NLOAD) continue;
            const int grp = q / (H * S::QW), rem = q - grp * (H * S::QW);
            const int rr = rem / S::QW, qx = rem - rr * S::QW;
            const int gy = front + rr, gx = lx0 + 4 * qx;
            if (gy < y0 - 6 || gy >= y0 + R + 6) continue;
            if (!(gy >= 0 && gy < n && gx >= 0 && gx + 3 < n)) continue;
            const size_t g = gat(gx, gy, n);
            switch (grp) {
                case 0: pa[it] = *reinterpret_cast<const float4*>(io.r + g);
                        pb[it] = *reinterpret_cast<const float4*>(io.d + g); break;
                case 1: pa[it] = *reinterpret_cast<const float4*>(io.b + g); break;
                case 2: pa[it] = *reinterpret_cast<const float4*>(io.fl + g);
                        pb[it] = *reinterpret_cast<const float4*>(io.fr + g); break;
                case 3: pa[it] = *reinterpret_cast<const float4*>(io.ft + g);
                        pb[it] = *reinterpret_cast<const float4*>(io.fb + g); break;
                default: pa[it] = *reinterpret_cast<const float4*>(io.s0 + g);
                         pb[it] = *reinterpret_cast<const float4*>(io.s1 + g); break;
            }
        }
    };
    auto store_loads = [&](int front) {
#pragma unroll
        for (int it = 0; it < S::LOAD_IT; ++it) {
            const int q = tid + it * NT;
            if (S::NLOAD % NT != 0 && q >= S::NLOAD) continue;
            const int grp = q / (H * S::QW), rem = q - grp * (H * S::QW);
            const int rr = rem / S::QW, qx = rem - rr * S::QW;
            const int gy = front + rr, gx = lx0 + 4 * qx, lc = 4 * qx;
            if (gy < y0 - 6 || gy >= y0 + R + 6) continue;
            // a quad lies wholly inside or wholly outside the grid (n is a
            // multiple of 8, quads start at multiples of 4)
            const bool inside = gy >= 0 && gy < n && gx >= 0 && gx + 3 < n;
            const float a4[4] = {pa[it].x, pa[it].y, pa[it].z, pa[it].w};
            const float b4[4] = {pb[it].x, pb[it].y, pb[it].z, pb[it].w};
            switch (grp) {
                case 0:
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        inp(S::I_R, gy, lc + j) = a4[j];
                        inp(S::I_D, gy, lc + j) = b4[j];
                        inp(S::I_RD, gy, lc + j) = a4[j] + b4[j];   // 0 outside: erosion reads 0 there
                    }
                    break;
                case 1: {
                    float inc[4] = {0.f, 0.f, 0.f, 0.f};
                    if (inside && a.rain) {
                        if (k.rain_mode == HG_RAIN_UNIFORM) {
                            uint32_t o[4];
                            philox4x32_10(uint32_t(gx) >> 2, uint32_t(gy), a.step, a.terrain, k.seed_lo, k.seed_hi, o);
#pragma unroll
                            for (int j = 0; j < 4; ++j) inc[j] = k.rain_amount * u24(o[j]);
                        } else if (k.rain_mode == HG_RAIN_CONSTANT) {
#pragma unroll
                            for (int j = 0; j < 4; ++j) inc[j] = k.rain_amount;
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j)   // outside: total height HG_OOB_HEIGHT for the flux
                        inp(S::I_B, gy, lc + j) = inside ? (a.rain ? a4[j] + inc[j] : a4[j]) : HG_OOB_HEIGHT;
                    break;
                }
                case 2:
#pragma unroll
                    for (int j = 0; j < 4; ++j) { inp(S::I_FL, gy, lc + j) = a4[j]; inp(S::I_FR, gy, lc + j) = b4[j]; }
                    break;
                case 3:
#pragma unroll
                    for (int j = 0; j < 4; ++j) { inp(S::I_FT, gy, lc + j) = a4[j]; inp(S::I_FB, gy, lc + j) = b4[j]; }
                    break;
                default:
#pragma unroll
                    for (int j = 0; j < 4; ++j) { inp(S::I_S0, gy, lc + j) = a4[j]; inp(S::I_S1, gy, lc + j) = b4[j]; }
                    break;
            }
        }
    };

    constexpr int DXK[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
    constexpr int DYK[8] = {0, 0, 1, -1, 1, 1, -1, -1};
    const int niter = (R + 12 + H - 1) / H;
    int front = y0 - 6;
    issue_loads(front);
    for (int iter = 0; iter < niter; ++iter, front += H) {
        // ---- phase 0: loads of this block to LDS; next block's loads issued
        store_loads(front);
        issue_loads(front + H);
        __syncthreads();

        // ---- phase 1: flux + erosion, rows front-1 .., columns -SF .. W+SF
        {
            const int i = tid;
            if (i < H * (W + 2 * S::SF)) {
                const int rr = i / (W + 2 * S::SF), c = i - rr * (W + 2 * S::SF) - S::SF;
                const int gy = front - S::LF + rr, gx = x0 + c, lc = c + S::XO;
                if (gy >= y0 - 5 && gy < y0 + R + 5) {
                    if (!in_grid(gx, gy)) {
                        ring(S::O_R1, S::RR1, gy, lc) = HG_OOB_HEIGHT;
                        ring(S::O_D1, S::RD1, gy, lc) = HG_OOB_HEIGHT;
                        ring(S::O_E0, S::RSE, gy, lc) = 0.0f;
                        ring(S::O_E1, S::RSE, gy, lc) = 0.0f;
                    } else {
                        const int so = slot<S::RIN>(gy) * S::WL, su = slot<S::RIN>(gy + 1) * S::WL,
                                  sd = slot<S::RIN>(gy - 1) * S::WL;
                        auto P = [&](int plane, int s, int col) { return in[plane * S::RIN * S::WL + s + col]; };
                        float f[4] = {P(S::I_FL, so, lc), P(S::I_FR, so, lc), P(S::I_FT, so, lc), P(S::I_FB, so, lc)};
                        const float b = P(S::I_B, so, lc);
                        const float w = P(S::I_RD, so, lc) + b;
                        const float wn[4] = {P(S::I_RD, so, lc - 1) + P(S::I_B, so, lc - 1),
                                             P(S::I_RD, so, lc + 1) + P(S::I_B, so, lc + 1),
                                             P(S::I_RD, su, lc) + P(S::I_B, su, lc),
                                             P(S::I_RD, sd, lc) + P(S::I_B, sd, lc)};
                        const float fin[4] = {P(S::I_FR, so, lc - 1), P(S::I_FL, so, lc + 1), P(S::I_FB, su, lc),
                                              P(S::I_FT, sd, lc)};
                        const unsigned wall = (gx == 0 ? 1u : 0u) | (gx == n - 1 ? 2u : 0u) |
                                              (gy == 0 ? 4u : 0u) | (gy == n - 1 ? 8u : 0u);
                        float vx, vy, vz;
                        const float d2 = flux_cell(f, b, w, wn, fin, wall, k, vx, vy, vz);
                        float r = P(S::I_R, so, lc), d = P(S::I_D, so, lc);
                        float s0 = P(S::I_S0, so, lc), s1 = P(S::I_S1, so, lc);
                        erosion_cell(r, d, s0, s1, vx, vy, vz, P(S::I_RD, so, lc - 1), P(S::I_RD, so, lc + 1),
                                     P(S::I_RD, su, lc), P(S::I_RD, sd, lc), k);
                        ring(S::O_R1, S::RR1, gy, lc) = r;
                        ring(S::O_D1, S::RD1, gy, lc) = d;
                        ring(S::O_E0, S::RSE, gy, lc) = s0;
                        ring(S::O_E1, S::RSE, gy, lc) = s1;
                        if (c >= 0 && c < W) {
                            lds[S::O_VX + slot<S::RV>(gy) * W + c] = vx;
                            lds[S::O_VY + slot<S::RV>(gy) * W + c] = vy;
                            if (gy >= y0 && gy < y0 + R) {
                                const size_t g = gat(gx, gy, n);
                                st_out<NTS>(&io.bo[g], d2 * k.evap);     // sediment_transport.glsl:75
                                st_out<NTS>(&io.flo[g], f[0]);
                                st_out<NTS>(&io.fro[g], f[1]);
                                st_out<NTS>(&io.fto[g], f[2]);
                                st_out<NTS>(&io.fbo[g], f[3]);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();

        // ---- phase 2: layer-0 outflow scale; sediment transport
        if (a.do_thermal) {
            const int i = tid;
            if (i < H * (W + 2 * S::SG0)) {
                const int rr = i / (W + 2 * S::SG0), c = i - rr * (W + 2 * S::SG0) - S::SG0;
                const int gy = front - S::LG0 + rr, gx = x0 + c, lc = c + S::XO;
                if (gy >= y0 - 4 && gy < y0 + R + 4) {
                    const float r = ring(S::O_R1, S::RR1, gy, lc);
                    float D[8];
#pragma unroll
                    for (int kk = 0; kk < 8; ++kk) D[kk] = r - ring(S::O_R1, S::RR1, gy + DYK[kk], lc + DXK[kk]);
                    const float g = thermal_g(D, r, 0, k);
                    ring(S::O_G0, S::RG, gy, lc) = in_grid(gx, gy) ? g : 0.0f;
                }
            }
        }
        {
            const int i = tid;
            if (i < H * W) {
                const int rr = i / W, c = i - rr * W;
                const int gy = front - S::LTR + rr, gx = x0 + c;
                if (gy >= y0 && gy < y0 + R && gx < n) {
                    const float vx = lds[S::O_VX + slot<S::RV>(gy) * W + c];
                    const float vy = lds[S::O_VY + slot<S::RV>(gy) * W + c];
                    float bx, by;
                    back_trace(gx, gy, vx, vy, k.dt, n, n, bx, by);
                    const int ix = int(bx), iy = int(by);
                    const float fx = bx - float(ix), fy = by - float(iy);
                    const int ix1 = min(ix + 1, n - 1), iy1 = min(iy + 1, n - 1);
                    float o0, o1;
                    if (ix >= x0 - S::SF && ix1 < x0 + W + S::SF && iy >= gy - HA && iy1 <= gy + HA) {
                        const int c0 = ix - lx0, c1 = ix1 - lx0;
                        auto E = [&](int off, int yy, int cc) { return ring(off, S::RSE, yy, cc); };
                        o0 = lerp1(lerp1(E(S::O_E0, iy, c0), E(S::O_E0, iy, c1), fx),
                                   lerp1(E(S::O_E0, iy1, c0), E(S::O_E0, iy1, c1), fx), fy);
                        o1 = lerp1(lerp1(E(S::O_E1, iy, c0), E(S::O_E1, iy, c1), fx),
                                   lerp1(E(S::O_E1, iy1, c0), E(S::O_E1, iy1, c1), fx), fy);
                    } else {
                        // taps outside the window: recompute the eroded sediment
                        // of the taps (00, 10, 01, 11) with the same lerp order
                        float p0 = 0.0f, p1 = 0.0f, t0 = 0.0f, t1 = 0.0f, b0 = 0.0f, b1 = 0.0f;
#pragma unroll 1
                        for (int tap = 0; tap < 4; ++tap) {
                            const int qx = (tap & 1) ? ix1 : ix, qy = (tap & 2) ? iy1 : iy;
                            float v0, v1;
                            eroded_sediment_global(io, qx, qy, n, n, a.step, a.terrain, a.rain != 0, k, v0, v1);
                            if ((tap & 1) == 0) {
                                p0 = v0; p1 = v1;
                            } else if (tap == 1) {
                                t0 = lerp1(p0, v0, fx); t1 = lerp1(p1, v1, fx);
                            } else {
                                b0 = lerp1(p0, v0, fx); b1 = lerp1(p1, v1, fx);
                            }
                        }
                        o0 = lerp1(t0, b0, fy);
                        o1 = lerp1(t1, b1, fy);
                    }
                    const size_t g = gat(gx, gy, n);
                    st_out<NTS>(&io.s0o[g], o0);
                    st_out<NTS>(&io.s1o[g], o1);
                }
            }
        }
        __syncthreads();

        // ---- phase 3: layer-0 transport (rock)
        {
            const int i = tid;
            if (i < H * (W + 2 * S::SU0)) {
                const int rr = i / (W + 2 * S::SU0), c = i - rr * (W + 2 * S::SU0) - S::SU0;
                const int gy = front - S::LU0 + rr, gx = x0 + c, lc = c + S::XO;
                if (gy >= y0 - 3 && gy < y0 + R + 3) {
                    const float r = ring(S::O_R1, S::RR1, gy, lc);
                    float nr = r;
                    if (a.do_thermal && in_grid(gx, gy)) {
                        const float g = ring(S::O_G0, S::RG, gy, lc);
                        float acc = 0.0f;
#pragma unroll
                        for (int kk = 0; kk < 8; ++kk) {
                            const float Dk = r - ring(S::O_R1, S::RR1, gy + DYK[kk], lc + DXK[kk]);
                            acc = acc + thermal_term(Dk, g, ring(S::O_G0, S::RG, gy + DYK[kk], lc + DXK[kk]),
                                                     kk < 4 ? k.tan_alpha[0] : k.thr_diag[0]);
                        }
                        nr = r + acc;
                    }
                    ring(S::O_R2, S::RR2, gy, lc) = nr;
                }
            }
        }
        __syncthreads();

        // ---- phase 4: layer-1 outflow scale (rock after layer 0)
        if (a.do_thermal) {
            const int i = tid;
            if (i < H * (W + 2 * S::SG1)) {
                const int rr = i / (W + 2 * S::SG1), c = i - rr * (W + 2 * S::SG1) - S::SG1;
                const int gy = front - S::LG1 + rr, gx = x0 + c, lc = c + S::XO;
                if (gy >= y0 - 2 && gy < y0 + R + 2) {
                    const float r = ring(S::O_R2, S::RR2, gy, lc), d = ring(S::O_D1, S::RD1, gy, lc);
                    float D[8];
#pragma unroll
                    for (int kk = 0; kk < 8; ++kk)
                        D[kk] = (r - ring(S::O_R2, S::RR2, gy + DYK[kk], lc + DXK[kk])) +
                                (d - ring(S::O_D1, S::RD1, gy + DYK[kk], lc + DXK[kk]));
                    const float g = thermal_g(D, d, 1, k);
                    ring(S::O_G1, S::RG, gy, lc) = in_grid(gx, gy) ? g : 0.0f;
                }
            }
        }
        __syncthreads();

        // ---- phase 5: layer-1 transport (dirt)
        {
            const int i = tid;
            if (i < H * (W + 2 * S::SU1)) {
                const int rr = i / (W + 2 * S::SU1), c = i - rr * (W + 2 * S::SU1) - S::SU1;
                const int gy = front - S::LU1 + rr, gx = x0 + c, lc = c + S::XO;
                if (gy >= y0 - 1 && gy < y0 + R + 1) {
                    const float d = ring(S::O_D1, S::RD1, gy, lc);
                    float nd = d;
                    if (a.do_thermal && in_grid(gx, gy)) {
                        const float r = ring(S::O_R2, S::RR2, gy, lc), g = ring(S::O_G1, S::RG, gy, lc);
                        float acc = 0.0f;
#pragma unroll
                        for (int kk = 0; kk < 8; ++kk) {
                            const float Dk = (r - ring(S::O_R2, S::RR2, gy + DYK[kk], lc + DXK[kk])) +
                                             (d - ring(S::O_D1, S::RD1, gy + DYK[kk], lc + DXK[kk]));
                            acc = acc + thermal_term(Dk, g, ring(S::O_G1, S::RG, gy + DYK[kk], lc + DXK[kk]),
                                                     kk < 4 ? k.tan_alpha[1] : k.thr_diag[1]);
                        }
                        nd = d + acc;
                    }
                    ring(S::O_DD, S::RDD, gy, lc) = nd;
                }
            }
        }
        __syncthreads();

        // ---- phase 6: smoothing, write rock and dirt
        {
            const int i = tid;
            if (i < H * W) {
                const int rr = i / W, c = i - rr * W;
                const int gy = front - S::LSM + rr, gx = x0 + c, lc = c + S::XO;
                if (gy >= y0 && gy < y0 + R && gx < n) {
                    float r = ring(S::O_R2, S::RR2, gy, lc), d = ring(S::O_DD, S::RDD, gy, lc);
                    if (a.do_smooth && gx > 0 && gy > 0 && gx < n - 1 && gy < n - 1)
                        smooth_cell(r, d, ring(S::O_R2, S::RR2, gy, lc - 1), ring(S::O_DD, S::RDD, gy, lc - 1),
                                    ring(S::O_R2, S::RR2, gy, lc + 1), ring(S::O_DD, S::RDD, gy, lc + 1),
                                    ring(S::O_R2, S::RR2, gy + 1, lc), ring(S::O_DD, S::RDD, gy + 1, lc),
                                    ring(S::O_R2, S::RR2, gy - 1, lc), ring(S::O_DD, S::RDD, gy - 1, lc),
                                    k.smooth_m);
                    const size_t g = gat(gx, gy, n);
                    st_out<NTS>(&io.ro[g], r);
                    st_out<NTS>(&io.dout[g], d);
                }
            }
        }
        // no barrier: the next phase 0 writes only the input ring, which phase
        // 1 of this iteration has finished reading (barrier after phase 1)
    }
}

//                      W   H   NT  HA  rows per segment
#define HG_SWEEP_VARIANTS(X)   \
    X(0, 128, 4, 576, 2, 256)  \
    X(1, 128, 4, 576, 2, 128)  \
    X(2, 128, 4, 576, 2, 512)

}  // namespace

int sweep_variant_count() {
#define X(i, w, h, nt, ha, rs) +1
    return 0 HG_SWEEP_VARIANTS(X);
#undef X
}

bool sweep_variant_info(int v, int* strip_w, int* rows, int* threads, int* seg_rows) {
    switch (v) {
#define X(i, W_, H_, NT_, HA_, RS_) \
    case i: *strip_w = W_; *rows = H_; *threads = NT_; *seg_rows = RS_; return true;
        HG_SWEEP_VARIANTS(X)
#undef X
        default: return false;
    }
}

hipError_t launch_step_sweep(int variant, const HydroIO& io, int n, uint32_t step, uint32_t terrain, bool rain,
                             bool thermal, bool smooth, const KParams& k, hipStream_t s) {
    switch (variant) {
#define X(i, W_, H_, NT_, HA_, RS_)                                                               \
    case i: {                                                                                    \
        SweepArgs a;                                                                             \
        a.io = io; a.n = n; a.seg_rows = RS_;                                                    \
        a.nstrips = (n + W_ - 1) / W_; a.nsegs = (n + RS_ - 1) / RS_;                            \
        a.step = step; a.terrain = terrain;                                                      \
        a.rain = rain ? 1 : 0; a.do_thermal = thermal ? 1 : 0; a.do_smooth = smooth ? 1 : 0;     \
        const unsigned grid = unsigned(a.nstrips) * unsigned(a.nsegs);                           \
        if (k.store_nt)                                                                          \
            k_step_sweep<W_, H_, NT_, HA_, true><<<grid, NT_, 0, s>>>(a, k);                     \
        else                                                                                     \
            k_step_sweep<W_, H_, NT_, HA_, false><<<grid, NT_, 0, s>>>(a, k);                    \
        break;                                                                                   \
    }
        HG_SWEEP_VARIANTS(X)
#undef X
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace hgk
