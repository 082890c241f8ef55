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
    static constexpr int WL = W + 2 * XO;       // LDS row width (floats)
    static constexpr int QW = WL / 4;           // quads per loaded row
    static constexpr int SF = 5;                // columns each side of flux/erosion
    static constexpr int CW = W + 2 * SF;       // columns of a thread block row
    // rows behind the load front
    static constexpr int LF = 1, LG0 = 2, LU0 = 3, LG1 = 4, LU1 = 5, LSM = 6, LTR = LF + HA;
    // Every intermediate lives in a ring of RS rows (row y in slot y mod RS),
    // the same for all, so one row index serves every ring a pass touches.
    // A ring must hold the rows of its producer's block and of everything its
    // deepest consumer still reads: H + (consumer lag + 1) - producer lag rows
    // (eroded dirt: H + 5; eroded sediment: H + 2 HA).  The own-cell inputs of
    // flux/erosion (rock, dirt, sediment) need H + 1.
    static constexpr int RS = 8, RSO = 4;
    static_assert(H + 5 <= RS && H + 2 * HA <= RS && H + 2 <= RS && H + 1 <= RSO, "ring too short");
    enum { P_RD, P_B, P_FL, P_FR, P_FT, P_FB, P_R1, P_D1, P_E0, P_E1, P_G0, P_R2, P_G1, P_DD, P_VX, P_VY, NP };
    enum { Q_R, Q_D, Q_S0, Q_S1, NQ };
    static constexpr int PLANE = RS * WL, OWN = RSO * WL;
    static constexpr int O_OWN = NP * PLANE;
    static constexpr int LDS_FLOATS = O_OWN + NQ * OWN;
    // load items: (plane group, row, quad); groups: rock+dirt, water, flux L/R,
    // flux T/B, sediment
    static constexpr int NGRP = 5, NLOAD = NGRP * H * QW, LOAD_IT = (NLOAD + NT - 1) / NT;
    static_assert(H * CW <= NT, "one flux/erosion cell per thread");
    static_assert(W % 4 == 0 && WL % 4 == 0, "strip width must be a multiple of 4");
};

struct SweepArgs {
    HydroIO io;
    int n, seg_rows, nstrips, nsegs;
    uint32_t step, terrain;
    int rain, do_thermal, do_smooth;
};

// LDS index of neighbour k (L, R, T, B, LT, RT, LB, RB) given the indices of
// the cell's row (i0), the row below (im) and the row above (ip)
template <int K>
__device__ __forceinline__ int nbi(int im, int i0, int ip) {
    constexpr int DX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
    constexpr int DY[8] = {0, 0, 1, -1, 1, 1, -1, -1};
    return (DY[K] > 0 ? ip : DY[K] < 0 ? im : i0) + DX[K];
}

template <int W, int H, int NT, int HA, bool NTS>
__global__ __launch_bounds__(NT) void k_step_sweep(SweepArgs a, KParams k) {
    using S = Sweep<W, H, NT, HA>;
    __shared__ __align__(16) float lds[S::LDS_FLOATS];
    const HydroIO& io = a.io;
    const int n = a.n;
    const int tid = threadIdx.x;
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
    auto M = [&](int plane, int i) -> float& { return lds[plane * S::PLANE + i]; };
    auto Q = [&](int plane, int i) -> float& { return lds[S::O_OWN + plane * S::OWN + i]; };

    // this thread's cell of every pass: row rr of the block, column c
    const int rr = tid / S::CW, c = tid - rr * S::CW - S::SF;
    const bool active = tid < H * S::CW;
    const int lc = c + S::XO, gx = x0 + c;
    const bool gx_in = gx >= 0 && gx < n;
    auto cols = [&](int side) { return active && c >= -side && c < W + side; };
    // LDS indices of rows y - 1, y, y + 1 in the rings (y >= -32)
    auto rows = [&](int y, int& im, int& i0, int& ip) {
        const int u = y + 64;
        im = ((u - 1) & (S::RS - 1)) * S::WL + lc;
        i0 = (u & (S::RS - 1)) * S::WL + lc;
        ip = ((u + 1) & (S::RS - 1)) * S::WL + lc;
    };

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
            const int lr = rem / S::QW, qx = rem - lr * S::QW;
            const int gy = front + lr, qgx = lx0 + 4 * qx;
            if (gy < y0 - 6 || gy >= y0 + R + 6) continue;
            if (!(gy >= 0 && gy < n && qgx >= 0 && qgx + 3 < n)) continue;
            const size_t g = gat(qgx, gy, n);
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
            const int lr = rem / S::QW, qx = rem - lr * S::QW;
            const int gy = front + lr, qgx = lx0 + 4 * qx;
            if (gy < y0 - 6 || gy >= y0 + R + 6) continue;
            const int u = gy + 64;
            const int im = (u & (S::RS - 1)) * S::WL + 4 * qx, io_ = (u & (S::RSO - 1)) * S::WL + 4 * qx;
            // a quad lies wholly inside or wholly outside the grid (n is a
            // multiple of 8, quads start at multiples of 4)
            const bool inside = gy >= 0 && gy < n && qgx >= 0 && qgx + 3 < n;
            const float4 A = pa[it], B = pb[it];
            auto put = [&](float* p, float4 v) { *reinterpret_cast<float4*>(p) = v; };
            switch (grp) {
                case 0:   // rock + dirt is 0 outside: erosion reads 0 there
                    put(&Q(S::Q_R, io_), A);
                    put(&Q(S::Q_D, io_), B);
                    put(&M(S::P_RD, im), make_float4(A.x + B.x, A.y + B.y, A.z + B.z, A.w + B.w));
                    break;
                case 1: {
                    // outside: water HG_OOB_HEIGHT, so the flux reads that total there
                    float4 w = inside ? A : make_float4(HG_OOB_HEIGHT, HG_OOB_HEIGHT, HG_OOB_HEIGHT, HG_OOB_HEIGHT);
                    if (inside && a.rain) {
                        float inc[4] = {0.f, 0.f, 0.f, 0.f};
                        if (k.rain_mode == HG_RAIN_UNIFORM) {
                            uint32_t o[4];
                            philox4x32_10(uint32_t(qgx) >> 2, uint32_t(gy), a.step, a.terrain, k.seed_lo, k.seed_hi, o);
#pragma unroll
                            for (int j = 0; j < 4; ++j) inc[j] = k.rain_amount * u24(o[j]);
                        } else if (k.rain_mode == HG_RAIN_CONSTANT) {
#pragma unroll
                            for (int j = 0; j < 4; ++j) inc[j] = k.rain_amount;
                        }
                        w = make_float4(A.x + inc[0], A.y + inc[1], A.z + inc[2], A.w + inc[3]);
                    }
                    put(&M(S::P_B, im), w);
                    break;
                }
                case 2: put(&M(S::P_FL, im), A); put(&M(S::P_FR, im), B); break;
                case 3: put(&M(S::P_FT, im), A); put(&M(S::P_FB, im), B); break;
                default: put(&Q(S::Q_S0, io_), A); put(&Q(S::Q_S1, io_), B); break;
            }
        }
    };

    const int niter = (R + 12 + H - 1) / H;
    int front = y0 - 6;
    issue_loads(front);
    for (int iter = 0; iter < niter; ++iter, front += H) {
        // ---- phase 0: loads of this block to LDS; next block's loads issued
        store_loads(front);
        issue_loads(front + H);
        __syncthreads();

        // ---- phase 1: flux + erosion, rows front-1 .., columns -5 .. W+5
        if (active) {
            const int gy = front - S::LF + rr;
            if (gy >= y0 - 5 && gy < y0 + R + 5) {
                int im, i0, ip;
                rows(gy, im, i0, ip);
                if (!(gx_in && gy >= 0 && gy < n)) {
                    M(S::P_R1, i0) = HG_OOB_HEIGHT;
                    M(S::P_D1, i0) = HG_OOB_HEIGHT;
                    M(S::P_E0, i0) = 0.0f;
                    M(S::P_E1, i0) = 0.0f;
                } else {
                    const int iq = ((gy + 64) & (S::RSO - 1)) * S::WL + lc;
                    float f[4] = {M(S::P_FL, i0), M(S::P_FR, i0), M(S::P_FT, i0), M(S::P_FB, i0)};
                    const float b = M(S::P_B, i0);
                    const float w = M(S::P_RD, i0) + b;
                    const float hl = M(S::P_RD, i0 - 1), hr = M(S::P_RD, i0 + 1), ht = M(S::P_RD, ip),
                                hb = M(S::P_RD, im);
                    const float wn[4] = {hl + M(S::P_B, i0 - 1), hr + M(S::P_B, i0 + 1), ht + M(S::P_B, ip),
                                         hb + M(S::P_B, im)};
                    const float fin[4] = {M(S::P_FR, i0 - 1), M(S::P_FL, i0 + 1), M(S::P_FB, ip), M(S::P_FT, im)};
                    const unsigned wall = (gx == 0 ? 1u : 0u) | (gx == n - 1 ? 2u : 0u) |
                                          (gy == 0 ? 4u : 0u) | (gy == n - 1 ? 8u : 0u);
                    float vx, vy, vz;
                    const float d2 = flux_cell(f, b, w, wn, fin, wall, k, vx, vy, vz);
                    float r = Q(S::Q_R, iq), d = Q(S::Q_D, iq), s0 = Q(S::Q_S0, iq), s1 = Q(S::Q_S1, iq);
                    erosion_cell(r, d, s0, s1, vx, vy, vz, hl, hr, ht, hb, k);
                    M(S::P_R1, i0) = r;
                    M(S::P_D1, i0) = d;
                    M(S::P_E0, i0) = s0;
                    M(S::P_E1, i0) = s1;
                    if (c >= 0 && c < W) {
                        M(S::P_VX, i0) = vx;
                        M(S::P_VY, i0) = vy;
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
        __syncthreads();

        // ---- phase 2: layer-0 outflow scale; sediment transport
        if (a.do_thermal && cols(4)) {
            const int gy = front - S::LG0 + rr;
            if (gy >= y0 - 4 && gy < y0 + R + 4) {
                int im, i0, ip;
                rows(gy, im, i0, ip);
                const float r = M(S::P_R1, i0);
                const float D[8] = {r - M(S::P_R1, nbi<0>(im, i0, ip)), r - M(S::P_R1, nbi<1>(im, i0, ip)),
                                    r - M(S::P_R1, nbi<2>(im, i0, ip)), r - M(S::P_R1, nbi<3>(im, i0, ip)),
                                    r - M(S::P_R1, nbi<4>(im, i0, ip)), r - M(S::P_R1, nbi<5>(im, i0, ip)),
                                    r - M(S::P_R1, nbi<6>(im, i0, ip)), r - M(S::P_R1, nbi<7>(im, i0, ip))};
                const float g = thermal_g(D, r, 0, k);
                M(S::P_G0, i0) = (gx_in && gy >= 0 && gy < n) ? g : 0.0f;
            }
        }
        if (cols(0)) {
            const int gy = front - S::LTR + rr;
            if (gy >= y0 && gy < y0 + R && gx < n) {
                const int i0 = ((gy + 64) & (S::RS - 1)) * S::WL + lc;
                const float vx = M(S::P_VX, i0), vy = M(S::P_VY, i0);
                float bx, by;
                back_trace(gx, gy, vx, vy, k.dt, n, n, bx, by);
                const int ix = int(bx), iy = int(by);
                const float fx = bx - float(ix), fy = by - float(iy);
                const int ix1 = min(ix + 1, n - 1), iy1 = min(iy + 1, n - 1);
                float o0, o1;
                if (ix >= x0 - S::SF && ix1 < x0 + W + S::SF && iy >= gy - HA && iy1 <= gy + HA) {
                    const int e00 = ((iy + 64) & (S::RS - 1)) * S::WL + (ix - lx0);
                    const int e01 = ((iy1 + 64) & (S::RS - 1)) * S::WL + (ix - lx0);
                    const int dx1 = ix1 - ix;
                    o0 = lerp1(lerp1(M(S::P_E0, e00), M(S::P_E0, e00 + dx1), fx),
                               lerp1(M(S::P_E0, e01), M(S::P_E0, e01 + dx1), fx), fy);
                    o1 = lerp1(lerp1(M(S::P_E1, e00), M(S::P_E1, e00 + dx1), fx),
                               lerp1(M(S::P_E1, e01), M(S::P_E1, e01 + dx1), fx), fy);
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
        __syncthreads();

        // ---- phase 3: layer-0 transport (rock)
        if (cols(3)) {
            const int gy = front - S::LU0 + rr;
            if (gy >= y0 - 3 && gy < y0 + R + 3) {
                int im, i0, ip;
                rows(gy, im, i0, ip);
                const float r = M(S::P_R1, i0);
                float nr = r;
                if (a.do_thermal && gx_in && gy >= 0 && gy < n) {
                    const float g = M(S::P_G0, i0);
                    const float ta = k.tan_alpha[0], td = k.thr_diag[0];
                    float acc = 0.0f;
                    acc = acc + thermal_term(r - M(S::P_R1, nbi<0>(im, i0, ip)), g, M(S::P_G0, nbi<0>(im, i0, ip)), ta);
                    acc = acc + thermal_term(r - M(S::P_R1, nbi<1>(im, i0, ip)), g, M(S::P_G0, nbi<1>(im, i0, ip)), ta);
                    acc = acc + thermal_term(r - M(S::P_R1, nbi<2>(im, i0, ip)), g, M(S::P_G0, nbi<2>(im, i0, ip)), ta);
                    acc = acc + thermal_term(r - M(S::P_R1, nbi<3>(im, i0, ip)), g, M(S::P_G0, nbi<3>(im, i0, ip)), ta);
                    acc = acc + thermal_term(r - M(S::P_R1, nbi<4>(im, i0, ip)), g, M(S::P_G0, nbi<4>(im, i0, ip)), td);
                    acc = acc + thermal_term(r - M(S::P_R1, nbi<5>(im, i0, ip)), g, M(S::P_G0, nbi<5>(im, i0, ip)), td);
                    acc = acc + thermal_term(r - M(S::P_R1, nbi<6>(im, i0, ip)), g, M(S::P_G0, nbi<6>(im, i0, ip)), td);
                    acc = acc + thermal_term(r - M(S::P_R1, nbi<7>(im, i0, ip)), g, M(S::P_G0, nbi<7>(im, i0, ip)), td);
                    nr = r + acc;
                }
                M(S::P_R2, i0) = nr;
            }
        }
        __syncthreads();

        // ---- phase 4: layer-1 outflow scale (rock after layer 0)
        if (a.do_thermal && cols(2)) {
            const int gy = front - S::LG1 + rr;
            if (gy >= y0 - 2 && gy < y0 + R + 2) {
                int im, i0, ip;
                rows(gy, im, i0, ip);
                const float r = M(S::P_R2, i0), d = M(S::P_D1, i0);
#define HG_D1(K) ((r - M(S::P_R2, nbi<K>(im, i0, ip))) + (d - M(S::P_D1, nbi<K>(im, i0, ip))))
                const float D[8] = {HG_D1(0), HG_D1(1), HG_D1(2), HG_D1(3), HG_D1(4), HG_D1(5), HG_D1(6), HG_D1(7)};
                const float g = thermal_g(D, d, 1, k);
                M(S::P_G1, i0) = (gx_in && gy >= 0 && gy < n) ? g : 0.0f;
            }
        }
        __syncthreads();

        // ---- phase 5: layer-1 transport (dirt)
        if (cols(1)) {
            const int gy = front - S::LU1 + rr;
            if (gy >= y0 - 1 && gy < y0 + R + 1) {
                int im, i0, ip;
                rows(gy, im, i0, ip);
                const float d = M(S::P_D1, i0);
                float nd = d;
                if (a.do_thermal && gx_in && gy >= 0 && gy < n) {
                    const float r = M(S::P_R2, i0), g = M(S::P_G1, i0);
                    const float ta = k.tan_alpha[1], td = k.thr_diag[1];
                    float acc = 0.0f;
                    acc = acc + thermal_term(HG_D1(0), g, M(S::P_G1, nbi<0>(im, i0, ip)), ta);
                    acc = acc + thermal_term(HG_D1(1), g, M(S::P_G1, nbi<1>(im, i0, ip)), ta);
                    acc = acc + thermal_term(HG_D1(2), g, M(S::P_G1, nbi<2>(im, i0, ip)), ta);
                    acc = acc + thermal_term(HG_D1(3), g, M(S::P_G1, nbi<3>(im, i0, ip)), ta);
                    acc = acc + thermal_term(HG_D1(4), g, M(S::P_G1, nbi<4>(im, i0, ip)), td);
                    acc = acc + thermal_term(HG_D1(5), g, M(S::P_G1, nbi<5>(im, i0, ip)), td);
                    acc = acc + thermal_term(HG_D1(6), g, M(S::P_G1, nbi<6>(im, i0, ip)), td);
                    acc = acc + thermal_term(HG_D1(7), g, M(S::P_G1, nbi<7>(im, i0, ip)), td);
                    nd = d + acc;
                }
#undef HG_D1
                M(S::P_DD, i0) = nd;
            }
        }
        __syncthreads();

        // ---- phase 6: smoothing, write rock and dirt
        if (cols(0)) {
            const int gy = front - S::LSM + rr;
            if (gy >= y0 && gy < y0 + R && gx < n) {
                int im, i0, ip;
                rows(gy, im, i0, ip);
                float r = M(S::P_R2, i0), d = M(S::P_DD, i0);
                if (a.do_smooth && gx > 0 && gy > 0 && gx < n - 1 && gy < n - 1)
                    smooth_cell(r, d, M(S::P_R2, i0 - 1), M(S::P_DD, i0 - 1), M(S::P_R2, i0 + 1),
                                M(S::P_DD, i0 + 1), M(S::P_R2, ip), M(S::P_DD, ip), M(S::P_R2, im),
                                M(S::P_DD, im), k.smooth_m);
                const size_t g = gat(gx, gy, n);
                st_out<NTS>(&io.ro[g], r);
                st_out<NTS>(&io.dout[g], d);
            }
        }
        // no barrier: the next phase 0 writes only the input rings, which
        // phase 1 of this iteration has finished reading (barrier after it)
    }
}

// ---------------------------------------------------------------------------
// The same sweep with one wave per strip and no LDS stencils.  Lane l of a
// wave holds column xs - 6 + l of a strip of 52 output columns (lanes 6..57),
// and the wave walks its segment one row per iteration.  Every pass keeps the
// rows it and its consumers need in registers (a window that moves up one
// row per iteration); a cell's left and right neighbours come from the lanes
// beside it through DPP wave shifts.  Passes evaluate lanes 1..62 (flux and
// erosion) down to 6..57 (smoothing), one lane fewer each side per pass.
// The eroded sediment that the transport's bilinear taps read sits in a ring
// of 8 rows x 64 lanes of LDS private to the wave (taps within 2 rows and
// inside the wave's columns; the rest recompute exactly).  No pass waits for
// another wave, so the kernel has no barrier.
// ---------------------------------------------------------------------------
constexpr int WV_OUT = 52, WV_XL = 6, WV_WAVES = 4;

// value of the lane to the left (lane - 1) / right (lane + 1) of this one
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

template <int N>
__device__ __forceinline__ void push(float (&w)[N], float v) {
#pragma unroll
    for (int i = N - 1; i > 0; --i) w[i] = w[i - 1];
    w[0] = v;
}

template <int RSEG, int OCC, bool NTS>
__global__ __launch_bounds__(64 * WV_WAVES) __attribute__((amdgpu_waves_per_eu(OCC))) void k_step_wave(
    SweepArgs a, KParams k) {
    __shared__ float esh[WV_WAVES][2][8][64];
    const HydroIO& io = a.io;
    const int n = a.n;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int gw = blockIdx.x * WV_WAVES + wv;
    const int strip = gw % a.nstrips, seg = gw / a.nstrips;
    if (seg >= a.nsegs) return;                       // the whole wave
    const int xs = strip * WV_OUT, xl0 = xs - WV_XL;  // grid x of lane 0
    const int x = xl0 + lane;
    const bool x_in = x >= 0 && x < n;
    const int y0 = seg * RSEG, R = min(RSEG, n - y0);
    const bool out_lane = lane >= WV_XL && lane < WV_XL + WV_OUT && x < n;
    float* e0ring = &esh[wv][0][0][0];
    float* e1ring = &esh[wv][1][0][0];

    // windows: index 0 is the newest row
    float rd[3] = {}, bw[3] = {}, wfl[2] = {}, wfr[2] = {}, wft[3] = {}, wfb[2] = {};
    float wr[2] = {}, wd[2] = {}, ws0[2] = {}, ws1[2] = {};
    float r1[4] = {}, d1[6] = {}, vx3[3] = {}, vy3[3] = {}, g0[3] = {}, r2[5] = {}, g1[3] = {}, dd[3] = {};
    // prefetched row
    float pr = 0.f, pd = 0.f, pb = 0.f, pfl = 0.f, pfr = 0.f, pft = 0.f, pfb = 0.f, ps0 = 0.f, ps1 = 0.f;
    auto load_row = [&](int y) {
        pr = pd = pb = pfl = pfr = pft = pfb = ps0 = ps1 = 0.f;
        if (x_in && y >= 0 && y < n) {
            const size_t g = gat(x, y, n);
            pr = io.r[g]; pd = io.d[g]; pb = io.b[g];
            pfl = io.fl[g]; pfr = io.fr[g]; pft = io.ft[g]; pfb = io.fb[g];
            ps0 = io.s0[g]; ps1 = io.s1[g];
        }
    };
    const int niter = R + 12;
    int f = y0 - 6;
    load_row(f);
    for (int t = 0; t < niter; ++t, ++f) {
        // ---- row f enters the input windows (out of the grid: rock + dirt 0,
        // total height HG_OOB_HEIGHT through the water, no flux)
        {
            const bool inside = x_in && f >= 0 && f < n;
            float b = inside ? pb : HG_OOB_HEIGHT;
            if (inside && a.rain) b = b + rain_incr(x, f, a.step, a.terrain, k);
            push(rd, pr + pd);
            push(bw, b);
            push(wfl, pfl); push(wfr, pfr); push(wft, pft); push(wfb, pfb);
            push(wr, pr); push(wd, pd); push(ws0, ps0); push(ws1, ps1);
        }
        if (t + 1 < niter) load_row(f + 1);

        // ---- flux + erosion, row f - 1 (inputs: [0] row f, [1] row f - 1, [2] row f - 2)
        {
            const int gy = f - 1;
            const float w0 = rd[1] + bw[1];
            const float wl = from_left(w0), wrt = from_right(w0);
            const float hl = from_left(rd[1]), hr = from_right(rd[1]);
            const float fin_l = from_left(wfr[1]), fin_r = from_right(wfl[1]);
            float nr1 = HG_OOB_HEIGHT, nd1 = HG_OOB_HEIGHT, ne0 = 0.0f, ne1 = 0.0f, nvx = 0.0f, nvy = 0.0f;
            if (lane >= 1 && lane < 63 && gy >= y0 - 5 && gy < y0 + R + 5 && x_in && gy >= 0 && gy < n) {
                float f4[4] = {wfl[1], wfr[1], wft[1], wfb[1]};
                const float wn[4] = {wl, wrt, rd[0] + bw[0], rd[2] + bw[2]};
                const float fin[4] = {fin_l, fin_r, wfb[0], wft[2]};
                const unsigned wall = (x == 0 ? 1u : 0u) | (x == n - 1 ? 2u : 0u) | (gy == 0 ? 4u : 0u) |
                                      (gy == n - 1 ? 8u : 0u);
                float vz;
                const float d2 = flux_cell(f4, bw[1], w0, wn, fin, wall, k, nvx, nvy, vz);
                float r = wr[1], d = wd[1], s0 = ws0[1], s1 = ws1[1];
                erosion_cell(r, d, s0, s1, nvx, nvy, vz, hl, hr, rd[0], rd[2], k);
                nr1 = r; nd1 = d; ne0 = s0; ne1 = s1;
                if (out_lane && gy >= y0 && gy < y0 + R) {
                    const size_t g = gat(x, gy, n);
                    st_out<NTS>(&io.bo[g], d2 * k.evap);     // sediment_transport.glsl:75
                    st_out<NTS>(&io.flo[g], f4[0]);
                    st_out<NTS>(&io.fro[g], f4[1]);
                    st_out<NTS>(&io.fto[g], f4[2]);
                    st_out<NTS>(&io.fbo[g], f4[3]);
                }
            }
            push(r1, nr1); push(d1, nd1); push(vx3, nvx); push(vy3, nvy);
            const int sl = ((gy + 64) & 7) * 64 + lane;
            e0ring[sl] = ne0;
            e1ring[sl] = ne1;
        }

        // ---- thermal layer 0 outflow scale, row f - 2 (r1: [0] f-1 .. [3] f-4)
        {
            const int gy = f - 2;
            const float c = r1[1];
            const float D[8] = {c - from_left(r1[1]), c - from_right(r1[1]), c - r1[0], c - r1[2],
                                c - from_left(r1[0]), c - from_right(r1[0]), c - from_left(r1[2]),
                                c - from_right(r1[2])};
            float ng = 0.0f;
            if (a.do_thermal && lane >= 2 && lane < 62 && gy >= y0 - 4 && gy < y0 + R + 4) {
                const float g = thermal_g(D, c, 0, k);
                ng = (x_in && gy >= 0 && gy < n) ? g : 0.0f;
            }
            push(g0, ng);
        }

        // ---- sediment transport, row f - 3 (velocity [2]; eroded sediment rows
        // f-8 .. f-1 in the ring)
        {
            const int gy = f - 3;
            if (out_lane && gy >= y0 && gy < y0 + R) {
                float bx, by;
                back_trace(x, gy, vx3[2], vy3[2], k.dt, n, n, bx, by);
                const int ix = int(bx), iy = int(by);
                const float fx = bx - float(ix), fy = by - float(iy);
                const int ix1 = min(ix + 1, n - 1), iy1 = min(iy + 1, n - 1);
                const int l0 = ix - xl0, l1 = ix1 - xl0;
                float o0, o1;
                if (l0 >= 1 && l1 <= 62 && iy >= gy - 2 && iy1 <= gy + 2) {
                    const int a0 = ((iy + 64) & 7) * 64, a1 = ((iy1 + 64) & 7) * 64;
                    o0 = lerp1(lerp1(e0ring[a0 + l0], e0ring[a0 + l1], fx), lerp1(e0ring[a1 + l0], e0ring[a1 + l1], fx), fy);
                    o1 = lerp1(lerp1(e1ring[a0 + l0], e1ring[a0 + l1], fx), lerp1(e1ring[a1 + l0], e1ring[a1 + l1], fx), fy);
                } else {
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
                const size_t g = gat(x, gy, n);
                st_out<NTS>(&io.s0o[g], o0);
                st_out<NTS>(&io.s1o[g], o1);
            }
        }

        // ---- thermal layer 0 transport, row f - 3 (r1 [1..3], g0 [0..2])
        {
            const int gy = f - 3;
            const float r = r1[2], g = g0[1];
            const float nbr[8] = {from_left(r1[2]), from_right(r1[2]), r1[1], r1[3],
                                  from_left(r1[1]), from_right(r1[1]), from_left(r1[3]), from_right(r1[3])};
            const float nbg[8] = {from_left(g0[1]), from_right(g0[1]), g0[0], g0[2],
                                  from_left(g0[0]), from_right(g0[0]), from_left(g0[2]), from_right(g0[2])};
            float nr = r;
            if (a.do_thermal && lane >= 3 && lane < 61 && gy >= y0 - 3 && gy < y0 + R + 3 && x_in && gy >= 0 &&
                gy < n) {
                float acc = 0.0f;
#pragma unroll
                for (int kk = 0; kk < 8; ++kk)
                    acc = acc + thermal_term(r - nbr[kk], g, nbg[kk], kk < 4 ? k.tan_alpha[0] : k.thr_diag[0]);
                nr = r + acc;
            }
            push(r2, nr);
        }

        // ---- thermal layer 1 outflow scale, row f - 4 (r2 [0..2], d1 [2..4])
        {
            const int gy = f - 4;
            const float r = r2[1], d = d1[3];
            const float D[8] = {
                (r - from_left(r2[1])) + (d - from_left(d1[3])), (r - from_right(r2[1])) + (d - from_right(d1[3])),
                (r - r2[0]) + (d - d1[2]), (r - r2[2]) + (d - d1[4]),
                (r - from_left(r2[0])) + (d - from_left(d1[2])), (r - from_right(r2[0])) + (d - from_right(d1[2])),
                (r - from_left(r2[2])) + (d - from_left(d1[4])), (r - from_right(r2[2])) + (d - from_right(d1[4]))};
            float ng = 0.0f;
            if (a.do_thermal && lane >= 4 && lane < 60 && gy >= y0 - 2 && gy < y0 + R + 2) {
                const float g = thermal_g(D, d, 1, k);
                ng = (x_in && gy >= 0 && gy < n) ? g : 0.0f;
            }
            push(g1, ng);
        }

        // ---- thermal layer 1 transport, row f - 5 (r2 [1..3], d1 [3..5], g1 [0..2])
        {
            const int gy = f - 5;
            const float r = r2[2], d = d1[4], g = g1[1];
            const float nbr[8] = {from_left(r2[2]), from_right(r2[2]), r2[1], r2[3],
                                  from_left(r2[1]), from_right(r2[1]), from_left(r2[3]), from_right(r2[3])};
            const float nbd[8] = {from_left(d1[4]), from_right(d1[4]), d1[3], d1[5],
                                  from_left(d1[3]), from_right(d1[3]), from_left(d1[5]), from_right(d1[5])};
            const float nbg[8] = {from_left(g1[1]), from_right(g1[1]), g1[0], g1[2],
                                  from_left(g1[0]), from_right(g1[0]), from_left(g1[2]), from_right(g1[2])};
            float nd = d;
            if (a.do_thermal && lane >= 5 && lane < 59 && gy >= y0 - 1 && gy < y0 + R + 1 && x_in && gy >= 0 &&
                gy < n) {
                float acc = 0.0f;
#pragma unroll
                for (int kk = 0; kk < 8; ++kk)
                    acc = acc + thermal_term((r - nbr[kk]) + (d - nbd[kk]), g, nbg[kk],
                                             kk < 4 ? k.tan_alpha[1] : k.thr_diag[1]);
                nd = d + acc;
            }
            push(dd, nd);
        }

        // ---- smoothing, row f - 6 (r2 [2..4], dd [0..2]); write rock and dirt
        {
            const int gy = f - 6;
            float r = r2[3], d = dd[1];
            const float lr = from_left(r2[3]), lg = from_left(dd[1]), rrt = from_right(r2[3]), rg = from_right(dd[1]);
            if (out_lane && gy >= y0 && gy < y0 + R) {
                if (a.do_smooth && x > 0 && gy > 0 && x < n - 1 && gy < n - 1)
                    smooth_cell(r, d, lr, lg, rrt, rg, r2[2], dd[0], r2[4], dd[2], k.smooth_m);
                const size_t g = gat(x, gy, n);
                st_out<NTS>(&io.ro[g], r);
                st_out<NTS>(&io.dout[g], d);
            }
        }
    }
}

// rows per segment, waves per SIMD the registers are sized for
#define HG_WAVE_VARIANTS(X) \
    X(0, 64, 4)             \
    X(1, 96, 4)             \
    X(2, 128, 4)            \
    X(3, 64, 3)             \
    X(4, 128, 3)

//                      W   H   NT  HA  rows per segment
#define HG_SWEEP_VARIANTS(X)   \
    X(0, 112, 3, 384, 2, 256)  \
    X(1, 120, 3, 448, 2, 256)  \
    X(2, 112, 3, 384, 2, 128)  \
    X(3, 240, 3, 768, 2, 256)  \
    X(4,  48, 3, 192, 2, 128)

}  // namespace

// Variants 0 .. (wave variants - 1) are the one-wave-per-strip kernel, the
// rest the workgroup kernel with LDS rings.
int sweep_variant_count() {
#define X(i, ...) +1
    return 0 HG_WAVE_VARIANTS(X) HG_SWEEP_VARIANTS(X);
#undef X
}

namespace {
constexpr int kWaveVariants = 0
#define X(i, ...) +1
    HG_WAVE_VARIANTS(X)
#undef X
    ;
}  // namespace

bool sweep_variant_info(int v, int* strip_w, int* rows, int* threads, int* seg_rows) {
    if (v < kWaveVariants) {
        switch (v) {
#define X(i, RS_, OCC_) \
    case i: *strip_w = WV_OUT; *rows = 1; *threads = 64 * WV_WAVES; *seg_rows = RS_; return true;
            HG_WAVE_VARIANTS(X)
#undef X
            default: return false;
        }
    }
    switch (v - kWaveVariants) {
#define X(i, W_, H_, NT_, HA_, RS_) \
    case i: *strip_w = W_; *rows = H_; *threads = NT_; *seg_rows = RS_; return true;
        HG_SWEEP_VARIANTS(X)
#undef X
        default: return false;
    }
}

hipError_t launch_step_sweep(int variant, const HydroIO& io, int n, uint32_t step, uint32_t terrain, bool rain,
                             bool thermal, bool smooth, const KParams& k, hipStream_t s) {
    SweepArgs a;
    a.io = io; a.n = n;
    a.step = step; a.terrain = terrain;
    a.rain = rain ? 1 : 0; a.do_thermal = thermal ? 1 : 0; a.do_smooth = smooth ? 1 : 0;
    if (variant < kWaveVariants) {
        switch (variant) {
#define X(i, RS_, OCC_)                                                                            \
    case i: {                                                                                     \
        a.seg_rows = RS_;                                                                         \
        a.nstrips = (n + WV_OUT - 1) / WV_OUT; a.nsegs = (n + RS_ - 1) / RS_;                     \
        const unsigned waves = unsigned(a.nstrips) * unsigned(a.nsegs);                           \
        const unsigned grid = (waves + WV_WAVES - 1) / WV_WAVES;                                  \
        if (k.store_nt)                                                                           \
            k_step_wave<RS_, OCC_, true><<<grid, 64 * WV_WAVES, 0, s>>>(a, k);                    \
        else                                                                                      \
            k_step_wave<RS_, OCC_, false><<<grid, 64 * WV_WAVES, 0, s>>>(a, k);                   \
        break;                                                                                    \
    }
            HG_WAVE_VARIANTS(X)
#undef X
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (variant - kWaveVariants) {
#define X(i, W_, H_, NT_, HA_, RS_)                                                               \
    case i: {                                                                                    \
        a.seg_rows = RS_;                                                                        \
        a.nstrips = (n + W_ - 1) / W_; a.nsegs = (n + RS_ - 1) / RS_;                            \
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
