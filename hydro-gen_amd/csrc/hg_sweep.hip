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

//                      W   H   NT  HA  rows per segment
#define HG_SWEEP_VARIANTS(X)   \
    X(0, 112, 3, 384, 2, 256)  \
    X(1, 120, 3, 448, 2, 256)  \
    X(2, 112, 3, 384, 2, 128)  \
    X(3, 240, 3, 768, 2, 256)  \
    X(4,  48, 3, 192, 2, 128)

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
