// hg_fused.hip — the GRID step in two fused kernels.
//
//  hydro   : rain -> hydro_flux -> hydro_erosion -> sediment_transport
//            (glsl/rain.glsl, hydro_flux.glsl, hydro_erosion.glsl,
//            sediment_transport.glsl).  Reads the 9 state planes once and
//            writes them once (72 B/cell).
//  thermal : thermal_erosion + thermal_transport for layer 0 then layer 1,
//            then smoothing (glsl/thermal_erosion.glsl, thermal_transport.glsl,
//            smoothing.glsl).  Reads rock+dirt once, writes them once (16 B/cell).
//
// Both kernels work on a TX x TY output tile per workgroup.  The tile plus a
// halo of its dependency radius is staged in LDS, every intermediate stage is
// evaluated on a region that shrinks by one cell per stage, and only the tile
// is written back.  The per-cell arithmetic is the same inline code as the
// per-pass kernels (hg_cell.hpp), so the fused and per-pass paths agree bit
// for bit.
//
// The semi-Lagrangian gather of sediment_transport has no bounded radius
// (|v|*dt is unbounded).  The hydro kernel evaluates eroded sediment on the
// tile plus an HA-cell halo; a back-trace whose bilinear taps fall outside it
// recomputes the eroded sediment of those taps from the step's input planes
// (still intact: every pass is ping-pong), which is exact.
#include "hg_kernels.hpp"

namespace hgk {

namespace {

__device__ __forceinline__ size_t at(int x, int y, int nx) { return size_t(y) * size_t(nx) + size_t(x); }
// Output and tape planes are written once per launch and not read back by it.
// When the step's state outgrows the last-level cache (KParams::store_nt, a
// compile-time NTS in the kernels: a run-time choice between two stores of the
// same value is folded into one plain store) they are written with
// non-temporal stores: hydro kernel 270 -> 262 us at 4096^2, same results.
// Below that size ordinary stores are faster (the next launch reads the planes
// from the cache).  Non-temporal loads of the inputs read once per workgroup
// measured slower and are not used.
template <bool NTS>
__device__ __forceinline__ void st_out(float* p, float v) {
    if constexpr (NTS)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// ---------------------------------------------------------------------------
// Cells of the region of radius RAD around a TX x TY tile.  When the tile
// width is a multiple of 64, the region is split into its core columns (the
// tile's own: one wave per 64-wide row chunk, so lanes touch 64 consecutive
// LDS words, free of bank conflicts, and global rows start 256-B aligned) and
// the two side strips of RAD columns, spread over the threads.  Otherwise the
// region is walked linearly.
// ---------------------------------------------------------------------------
template <int TX, int TY, int NT, int RAD>
struct Region {
    static constexpr int W = TX + 2 * RAD, H = TY + 2 * RAD, N = W * H;
    static constexpr bool ROWS = TX % 64 == 0 && NT % 64 == 0;
    static constexpr int NW = NT / 64, CPR = TX / 64;       // waves; 64-wide chunks per row
    static constexpr int NCORE = H * CPR, NSIDE = 2 * RAD * H;
    static constexpr int ITC = ROWS ? (NCORE + NW - 1) / NW : 0;
    static constexpr int ITS = ROWS ? (NSIDE + NT - 1) / NT : 0;
    static constexpr int ITER = ROWS ? ITC + ITS : (N + NT - 1) / NT;
};

// Calls fn(it, rx, ry) for the cells of the region, in region coordinates
// rx in [0, W), ry in [0, H); `it` < Region::ITER numbers a thread's calls.
template <int TX, int TY, int NT, int RAD, class F>
__device__ __forceinline__ void tile_region_for(int tid, F&& fn) {
    using R = Region<TX, TY, NT, RAD>;
    if constexpr (R::ROWS) {
        const int lane = tid & 63;
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
        for (int it = 0; it < R::ITC; ++it) {
            // The scalar guard stays even where every wave passes it: the
            // branch bounds how far the scheduler hoists LDS reads across
            // cells, which otherwise spills registers.
            const int q = wave + it * R::NW;              // wave-uniform
            if (q < R::NCORE) {
                const int row = q / R::CPR, c = q - row * R::CPR;
                fn(it, RAD + c * 64 + lane, row);
            }
        }
        if constexpr (RAD > 0) {
#pragma unroll
            for (int it = 0; it < R::ITS; ++it) {
                const int j = tid + it * NT;
                if (R::NSIDE % NT == 0 || j < R::NSIDE) {
                    const int row = j / (2 * RAD), c = j - row * (2 * RAD);
                    fn(R::ITC + it, c < RAD ? c : TX + c, row);
                }
            }
        }
    } else {
#pragma unroll
        for (int it = 0; it < R::ITER; ++it) {
            const int i = tid + it * NT;
            if (R::N % NT == 0 || i < R::N) {
                const int ry = i / R::W, rx = i - ry * R::W;
                fn(it, rx, ry);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// hydro kernel
// ---------------------------------------------------------------------------

template <int TX, int TY, int HA, int NT>
struct HydroTile {
    // input region: tile +- (HA+1) cells, loaded in aligned quads, so the LDS
    // region spans x in [x0 - QH*4, x0 + TX + QH*4)
    static constexpr int QH = (HA + 1 + 3) / 4;           // halo quads per side
    static constexpr int WI = TX + 8 * QH;                 // LDS row width (cells)
    static constexpr int OX = 4 * QH;                      // LDS x of tile column 0
    static constexpr int HI = TY + 2 * (HA + 1);           // LDS rows
    static constexpr int OY = HA + 1;                      // LDS y of tile row 0
    static constexpr int WE = TX + 2 * HA, HE = TY + 2 * HA;  // eroded-sediment region
    static constexpr int NE = WE * HE;
    static constexpr int NTILE = TX * TY;
    static constexpr int ITER_T = (NTILE + NT - 1) / NT;
    static constexpr int NLDS = WI * HI;
    static_assert(2 * NE + 2 * NTILE <= 6 * NLDS, "aliased outputs must fit the input arrays");
    static_assert(TX % 4 == 0, "tile width must be a multiple of 4");
};

// One output tile (bx, by) of the hydro kernel; `lds` holds 6*NLDS floats.
// Ends with LDS still in use: callers looping over tiles must barrier first.
// EXTRA: also write the optional outputs (velocity, adjoint tape) that are
// not null; without it their pointers are never read, which keeps them out
// of the scalar registers of the common case.
template <int TX, int TY, int HA, int NT, bool EXTRA = false, bool NTS = false>
__device__ __forceinline__ void hydro_tile(float* lds, const HydroIO& io, int nx, int ny, int bx,
                                           int by, uint32_t step, uint32_t terrain, int rain,
                                           const KParams& k) {
    using T = HydroTile<TX, TY, HA, NT>;
    // LDS planes (pitch WI): rd = rock+dirt, b = water after rain, fl..fb = old flux.
    float* s_rd = lds;
    float* s_b = lds + T::NLDS;
    float* s_fl = lds + 2 * T::NLDS;
    float* s_fr = lds + 3 * T::NLDS;
    float* s_ft = lds + 4 * T::NLDS;
    float* s_fb = lds + 5 * T::NLDS;

    // Opaque to the optimiser: values derived from the thread index are then
    // recomputed per tile instead of being hoisted out of the batched tile
    // loop and held in registers across tiles.
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int x0 = bx * TX, y0 = by * TY;
    const int lx0 = x0 - T::OX, ly0 = y0 - T::OY;   // grid coords of LDS (0,0)

    // ---- stage 0: stage inputs in aligned quads; rain is added here -------
    {
        constexpr int QW = T::WI / 4;
        constexpr int NQ = QW * T::HI;
        for (int q = tid; q < NQ; q += NT) {
            const int qy = q / QW, qxq = q - qy * QW;
            const int gx = lx0 + 4 * qxq, gy = ly0 + qy;
            const int li = qy * T::WI + 4 * qxq;
            float rd[4], b[4], fl[4], fr[4], ft[4], fb[4];
            if (gy >= 0 && gy < ny && gx >= 0 && gx + 3 < nx) {
                const size_t g = at(gx, gy, nx);
                const float4 r4 = *reinterpret_cast<const float4*>(io.r + g);
                const float4 d4 = *reinterpret_cast<const float4*>(io.d + g);
                const float4 b4 = *reinterpret_cast<const float4*>(io.b + g);
                const float4 l4 = *reinterpret_cast<const float4*>(io.fl + g);
                const float4 rr4 = *reinterpret_cast<const float4*>(io.fr + g);
                const float4 t4 = *reinterpret_cast<const float4*>(io.ft + g);
                const float4 bb4 = *reinterpret_cast<const float4*>(io.fb + g);
                rd[0] = r4.x + d4.x; rd[1] = r4.y + d4.y; rd[2] = r4.z + d4.z; rd[3] = r4.w + d4.w;
                b[0] = b4.x; b[1] = b4.y; b[2] = b4.z; b[3] = b4.w;
                fl[0] = l4.x; fl[1] = l4.y; fl[2] = l4.z; fl[3] = l4.w;
                fr[0] = rr4.x; fr[1] = rr4.y; fr[2] = rr4.z; fr[3] = rr4.w;
                ft[0] = t4.x; ft[1] = t4.y; ft[2] = t4.z; ft[3] = t4.w;
                fb[0] = bb4.x; fb[1] = bb4.y; fb[2] = bb4.z; fb[3] = bb4.w;
                if (rain) {
                    float inc[4];
                    if (k.rain_mode == HG_RAIN_UNIFORM) {
                        uint32_t o[4];
                        philox4x32_10(uint32_t(gx) >> 2, uint32_t(gy), step, terrain, k.seed_lo, k.seed_hi, o);
#pragma unroll
                        for (int j = 0; j < 4; ++j) inc[j] = k.rain_amount * u24(o[j]);
                    } else {
                        const float a = k.rain_mode == HG_RAIN_CONSTANT ? k.rain_amount : 0.0f;
#pragma unroll
                        for (int j = 0; j < 4; ++j) inc[j] = a;
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) b[j] = b[j] + inc[j];
                }
            } else {
                // whole quad outside the grid (n is a multiple of 8, so quads
                // never straddle the border): flux reads a total of
                // HG_OOB_HEIGHT and zero flux, erosion reads rock+dirt 0.
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    rd[j] = 0.0f; b[j] = HG_OOB_HEIGHT;
                    fl[j] = 0.0f; fr[j] = 0.0f; ft[j] = 0.0f; fb[j] = 0.0f;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                s_rd[li + j] = rd[j]; s_b[li + j] = b[j];
                s_fl[li + j] = fl[j]; s_fr[li + j] = fr[j]; s_ft[li + j] = ft[j]; s_fb[li + j] = fb[j];
            }
        }
    }
    __syncthreads();

    // ---- stage 1: flux + erosion on the tile +- HA region -----------------
    // (region coordinates ex in [0, TX + 2HA), ey in [0, TY + 2HA))
    constexpr int IE = Region<TX, TY, NT, HA>::ITER;
    auto walk = [&](auto&& fn) { tile_region_for<TX, TY, NT, HA>(tid, fn); };
    float es0[IE], es1[IE], evx[IE], evy[IE];
#pragma unroll
    for (int it = 0; it < IE; ++it) { es0[it] = 0.0f; es1[it] = 0.0f; evx[it] = 0.0f; evy[it] = 0.0f; }
    // rock, dirt and sediment of the region, at clamped addresses, all
    // issued before the first cell is evaluated
    float gr[IE], gd[IE], g0[IE], g1[IE];
    walk([&](int it, int ex, int ey) {
        const int gx = min(max(x0 - HA + ex, 0), nx - 1), gy = min(max(y0 - HA + ey, 0), ny - 1);
        const size_t g = at(gx, gy, nx);
        gr[it] = io.r[g];
        gd[it] = io.d[g];
        g0[it] = io.s0[g];
        g1[it] = io.s1[g];
    });
    walk([&](int it, int ex, int ey) {
        const int gx = x0 - HA + ex, gy = y0 - HA + ey;
        if (gx < 0 || gy < 0 || gx >= nx || gy >= ny) return;
        const int li = (ey + 1) * T::WI + (ex + T::OX - HA);   // LDS index of this cell
        float f[4] = {s_fl[li], s_fr[li], s_ft[li], s_fb[li]};
        const float b = s_b[li];
        const float w = s_rd[li] + b;
        const float wn[4] = {s_rd[li - 1] + s_b[li - 1], s_rd[li + 1] + s_b[li + 1],
                             s_rd[li + T::WI] + s_b[li + T::WI], s_rd[li - T::WI] + s_b[li - T::WI]};
        const float fin[4] = {s_fr[li - 1], s_fl[li + 1], s_fb[li + T::WI], s_ft[li - T::WI]};
        const unsigned wall = (gx == 0 ? 1u : 0u) | (gx == nx - 1 ? 2u : 0u) | (gy == 0 ? 4u : 0u) |
                              (gy == ny - 1 ? 8u : 0u);
        float vx, vy, vz;
        const float d2 = flux_cell(f, b, w, wn, fin, wall, k, vx, vy, vz);
        const size_t g = at(gx, gy, nx);
        float r = gr[it], d = gd[it];
        float s0 = g0[it], s1 = g1[it];
        erosion_cell(r, d, s0, s1, vx, vy, vz, s_rd[li - 1], s_rd[li + 1], s_rd[li + T::WI],
                     s_rd[li - T::WI], k);
        es0[it] = s0; es1[it] = s1; evx[it] = vx; evy[it] = vy;
        const bool tile = ex >= HA && ex < HA + TX && ey >= HA && ey < HA + TY;
        if (tile) {
            st_out<NTS>(&io.ro[g], r);
            st_out<NTS>(&io.dout[g], d);
            st_out<NTS>(&io.bo[g], d2 * k.evap);     // sediment_transport.glsl:75
            st_out<NTS>(&io.flo[g], f[0]);
            st_out<NTS>(&io.fro[g], f[1]);
            st_out<NTS>(&io.fto[g], f[2]);
            st_out<NTS>(&io.fbo[g], f[3]);
            if (EXTRA && io.vx) {
                st_out<NTS>(&io.vx[g], vx);
                st_out<NTS>(&io.vy[g], vy);
                st_out<NTS>(&io.vz[g], vz);
            }
            if (EXTRA && io.tb1) {
                st_out<NTS>(&io.tb1[g], b);
                st_out<NTS>(&io.tb2[g], d2);
                st_out<NTS>(&io.ts0[g], s0);
                st_out<NTS>(&io.ts1[g], s1);
            }
        }
    });
    __syncthreads();   // all reads of the staged inputs are done

    // eroded sediment of the region and velocity of the tile, aliased over LDS
    float* s_s0 = lds;
    float* s_s1 = lds + T::NE;
    float* s_vx = lds + 2 * T::NE;
    float* s_vy = lds + 2 * T::NE + T::NTILE;
    walk([&](int it, int ex, int ey) {
        const int e = ey * T::WE + ex;
        s_s0[e] = es0[it];
        s_s1[e] = es1[it];
        if (ex >= HA && ex < HA + TX && ey >= HA && ey < HA + TY) {
            const int t = (ey - HA) * TX + (ex - HA);
            s_vx[t] = evx[it];
            s_vy[t] = evy[it];
        }
    });
    __syncthreads();

    // ---- stage 2: back-trace + bilinear gather on the tile -----------------
    for (int t = tid; t < T::NTILE; t += NT) {
        const int ty = t / TX, tx = t - ty * TX;
        const int gx = x0 + tx, gy = y0 + ty;
        if (gx >= nx || gy >= ny) continue;
        float bx, by;
        back_trace(gx, gy, s_vx[t], s_vy[t], k.dt, nx, ny, bx, by);
        const int ix = int(bx), iy = int(by);
        const float fx = bx - float(ix), fy = by - float(iy);
        const int ix1 = min(ix + 1, nx - 1), iy1 = min(iy + 1, ny - 1);
        // region-local coordinates of the taps
        const int ex0 = ix - (x0 - HA), ey0 = iy - (y0 - HA);
        const int ex1 = ix1 - (x0 - HA), ey1 = iy1 - (y0 - HA);
        float o0, o1;
        if (ex0 >= 0 && ey0 >= 0 && ex1 < T::WE && ey1 < T::HE) {
            const int e00 = ey0 * T::WE + ex0, e10 = ey0 * T::WE + ex1;
            const int e01 = ey1 * T::WE + ex0, e11 = ey1 * T::WE + ex1;
            o0 = lerp1(lerp1(s_s0[e00], s_s0[e10], fx), lerp1(s_s0[e01], s_s0[e11], fx), fy);
            o1 = lerp1(lerp1(s_s1[e00], s_s1[e10], fx), lerp1(s_s1[e01], s_s1[e11], fx), fy);
        } else {
            // taps outside the halo (rare): one copy of the recompute, looped
            // over the taps (00, 10, 01, 11) with the same lerp order
            float p0 = 0.0f, p1 = 0.0f, t0 = 0.0f, t1 = 0.0f, b0 = 0.0f, b1 = 0.0f;
#pragma unroll 1
            for (int tap = 0; tap < 4; ++tap) {
                const int qx = (tap & 1) ? ix1 : ix, qy = (tap & 2) ? iy1 : iy;
                float v0, v1;
                eroded_sediment_global(io, qx, qy, nx, ny, step, terrain, rain != 0, k, v0, v1);
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
        const size_t g = at(gx, gy, nx);
        st_out<NTS>(&io.s0o[g], o0);
        st_out<NTS>(&io.s1o[g], o1);
    }
}

// Tile of workgroup b in a 1-D grid of nb.  Workgroups are dealt to the 8 XCDs
// round robin (b -> XCD b mod 8); with `xcd` set, XCD x gets the contiguous run
// of tiles starting at x*q + min(x, r) (q = nb/8, r = nb mod 8), so a tile and
// its neighbours, whose halos overlap, are read through the same L2.
__device__ __forceinline__ void tile_of(int b, int nb, int gx, int xcd, int& bx, int& by) {
    if (xcd == 2 && (gx & 7) == 0 && (nb & 7) == 0) {
        // XCD x owns the column strip of tiles [x*sw, x*sw + sw) and walks it
        // row by row: its L2 sees both the horizontal and the vertical
        // neighbours of a tile, and all XCDs advance through the same rows
        const int sw = gx >> 3, x = b & 7, j = b >> 3;
        by = j / sw;
        bx = x * sw + (j - by * sw);
        return;
    }
    int t = b;
    if (xcd == 1) {
        const int q = nb >> 3, r = nb & 7, x = b & 7, j = b >> 3;
        t = x * q + min(x, r) + j;
    }
    by = t / gx;
    bx = t - by * gx;
}

template <int TX, int TY, int HA, int NT, bool EXTRA, bool NTS>
__global__ __launch_bounds__(NT) void k_fused_hydro(HydroIO io, int nx, int ny, uint32_t step,
                                                    uint32_t terrain, int rain, int xcd, KParams k) {
    __shared__ float lds[6 * HydroTile<TX, TY, HA, NT>::NLDS];
    int bx, by;
    tile_of(blockIdx.x, gridDim.x, (nx + TX - 1) / TX, xcd, bx, by);
    hydro_tile<TX, TY, HA, NT, EXTRA, NTS>(lds, io, nx, ny, bx, by, step, terrain, rain, k);
}

// ---------------------------------------------------------------------------
// thermal kernel: thermal layer 0, thermal layer 1, smoothing
//
// LDS holds rock, dirt and the outflow scale g on the tile +- 5 cells.  Stages
// (radius = cells around the tile they are evaluated on):
//   A r4: g of layer 0      B r3: rock += thermal change (written in place)
//   C r2: g of layer 1      D r1: dirt += thermal change (written in place)
//   E r0: smoothing, write back.
// In-place updates go through registers between two barriers.
// ---------------------------------------------------------------------------
constexpr int TH = 5;   // dependency radius of thermal L0 + L1 + smoothing

// Calls fn(it, lx, ly) for the cells of the region of radius RAD around the
// thermal tile, in LDS coordinates.
template <int TX, int TY, int NT, int RAD, class F>
__device__ __forceinline__ void region_for(int tid, F&& fn) {
    tile_region_for<TX, TY, NT, RAD>(tid, [&](int it, int rx, int ry) { fn(it, rx + (TH - RAD), ry + (TH - RAD)); });
}

template <int TX, int TY>
struct ThermalTile {
    static constexpr int W0 = TX + 2 * TH, N0 = W0 * (TY + 2 * TH);
};

// One output tile (bx, by) of the thermal kernel; `lds` holds 3*N0 floats.
template <int TX, int TY, int NT, bool EXTRA = false, bool NTS = false>
__device__ __forceinline__ void thermal_tile(float* lds, const ThermIO& io, int nx, int ny, int bx,
                                             int by, int do_thermal, int do_smooth, const KParams& k) {
    constexpr int W0 = ThermalTile<TX, TY>::W0, N0 = ThermalTile<TX, TY>::N0;
    float* s_r = lds;
    float* s_d = lds + N0;
    float* s_g = lds + 2 * N0;
    // Opaque to the optimiser: values derived from the thread index are then
    // recomputed per tile instead of being hoisted out of the batched tile
    // loop and held in registers across tiles.
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int x0 = bx * TX - TH, y0 = by * TY - TH;   // grid coords of LDS (0,0)
    constexpr int DX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
    constexpr int DY[8] = {0, 0, 1, -1, 1, 1, -1, -1};
    constexpr int OFF[8] = {-1, 1, W0, -W0, W0 - 1, W0 + 1, -W0 - 1, -W0 + 1};
    (void)DX; (void)DY;
    auto in_grid = [&](int lx, int ly) {
        const int gx = x0 + lx, gy = y0 + ly;
        return gx >= 0 && gy >= 0 && gx < nx && gy < ny;
    };

    // load rock, dirt; out of grid reads HG_OOB_HEIGHT (thermal_erosion.glsl:20-26).
    // Loads use clamped addresses and are all issued before the first LDS write.
    {
        constexpr int IL = Region<TX, TY, NT, TH>::ITER;
        float lr[IL], ld[IL];
        region_for<TX, TY, NT, TH>(tid, [&](int it, int lx, int ly) {
            const int gx = x0 + lx, gy = y0 + ly;
            const size_t g = at(min(max(gx, 0), nx - 1), min(max(gy, 0), ny - 1), nx);
            const bool in = gx >= 0 && gy >= 0 && gx < nx && gy < ny;
            lr[it] = in ? io.r[g] : HG_OOB_HEIGHT;
            ld[it] = in ? io.d[g] : HG_OOB_HEIGHT;
        });
        region_for<TX, TY, NT, TH>(tid, [&](int it, int lx, int ly) {
            s_r[ly * W0 + lx] = lr[it];
            s_d[ly * W0 + lx] = ld[it];
        });
    }
    __syncthreads();

    if (do_thermal) {
        // ---- A: layer-0 outflow scale
        // (evaluated out of the grid too, on the out-of-grid height, and dropped)
        region_for<TX, TY, NT, 4>(tid, [&](int, int lx, int ly) {
            const int li = ly * W0 + lx;
            const float r = s_r[li];
            float D[8];
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) D[kk] = r - s_r[li + OFF[kk]];
            const float g = thermal_g(D, r, 0, k);
            s_g[li] = in_grid(lx, ly) ? g : 0.0f;
        });
        __syncthreads();
        // ---- B: layer-0 transport, rock in place
        {
            float nr[Region<TX, TY, NT, 3>::ITER];
            region_for<TX, TY, NT, 3>(tid, [&](int it, int lx, int ly) {
                const int li = ly * W0 + lx;
                const float r = s_r[li], g = s_g[li];
                float acc = 0.0f;
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    const float Dk = r - s_r[li + OFF[kk]];
                    acc = acc + thermal_term(Dk, g, s_g[li + OFF[kk]],
                                             kk < 4 ? k.tan_alpha[0] : k.thr_diag[0]);
                }
                nr[it] = r + acc;
            });
            __syncthreads();
            region_for<TX, TY, NT, 3>(tid, [&](int it, int lx, int ly) {
                if (in_grid(lx, ly)) s_r[ly * W0 + lx] = nr[it];
            });
            __syncthreads();
        }
        // ---- C: layer-1 outflow scale (rock after layer 0)
        region_for<TX, TY, NT, 2>(tid, [&](int, int lx, int ly) {
            const int li = ly * W0 + lx;
            const float r = s_r[li], d = s_d[li];
            float D[8];
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) D[kk] = (r - s_r[li + OFF[kk]]) + (d - s_d[li + OFF[kk]]);
            const float g = thermal_g(D, d, 1, k);
            s_g[li] = in_grid(lx, ly) ? g : 0.0f;
        });
        __syncthreads();
        // ---- D: layer-1 transport, dirt in place
        {
            float nd[Region<TX, TY, NT, 1>::ITER];
            region_for<TX, TY, NT, 1>(tid, [&](int it, int lx, int ly) {
                const int li = ly * W0 + lx;
                const float r = s_r[li], d = s_d[li], g = s_g[li];
                float acc = 0.0f;
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    const float Dk = (r - s_r[li + OFF[kk]]) + (d - s_d[li + OFF[kk]]);
                    acc = acc + thermal_term(Dk, g, s_g[li + OFF[kk]],
                                             kk < 4 ? k.tan_alpha[1] : k.thr_diag[1]);
                }
                nd[it] = d + acc;
            });
            __syncthreads();
            region_for<TX, TY, NT, 1>(tid, [&](int it, int lx, int ly) {
                if (in_grid(lx, ly)) s_d[ly * W0 + lx] = nd[it];
            });
            __syncthreads();
        }
    }

    // ---- E: smoothing on the tile, write back
    region_for<TX, TY, NT, 0>(tid, [&](int, int lx, int ly) {
        const int gx = x0 + lx, gy = y0 + ly;
        const int li = ly * W0 + lx;
        const float r0 = s_r[li], d0 = s_d[li];
        float r = r0, d = d0;
        if (do_smooth && gx > 0 && gy > 0 && gx < nx - 1 && gy < ny - 1)
            smooth_cell(r, d, s_r[li - 1], s_d[li - 1], s_r[li + 1], s_d[li + 1], s_r[li + W0],
                        s_d[li + W0], s_r[li - W0], s_d[li - W0], k.smooth_m);
        if (gx < nx && gy < ny) {
            const size_t g = at(gx, gy, nx);
            if (EXTRA && io.tr5) {
                st_out<NTS>(&io.tr5[g], r0);
                st_out<NTS>(&io.td6[g], d0);
            }
            st_out<NTS>(&io.ro[g], r);
            st_out<NTS>(&io.dout[g], d);
        }
    });
}

template <int TX, int TY, int NT, bool EXTRA, bool NTS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_fused_thermal(
    ThermIO io, int nx, int ny, int do_thermal, int do_smooth, int xcd, KParams k) {
    __shared__ float lds[3 * ThermalTile<TX, TY>::N0];
    int bx, by;
    tile_of(blockIdx.x, gridDim.x, (nx + TX - 1) / TX, xcd, bx, by);
    thermal_tile<TX, TY, NT, EXTRA, NTS>(lds, io, nx, ny, bx, by, do_thermal, do_smooth, k);
}

// ---------------------------------------------------------------------------
// Batched persistent kernels: one launch per pass covers the tiles of every
// terrain of a batch.  Workgroup b processes work-list entries b, b+G, ...
// (G = grid size); an entry packs (terrain << 24) | (tile_y << 12) | tile_x.
// Each terrain's descriptor holds both buffers of its 9 state planes; bit p of
// `cur` selects which buffer of plane p is the current (input) one.
// ---------------------------------------------------------------------------
__device__ __forceinline__ HydroIO batch_hydro_io(const BatchTerrain& T, uint32_t cur) {
    HydroIO io;
    const float** in[9] = {&io.r, &io.d, &io.b, &io.fl, &io.fr, &io.ft, &io.fb, &io.s0, &io.s1};
    float** out[9] = {&io.ro, &io.dout, &io.bo, &io.flo, &io.fro, &io.fto, &io.fbo, &io.s0o, &io.s1o};
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const int c = (cur >> p) & 1;
        *in[p] = T.plane[p][c];
        *out[p] = T.plane[p][c ^ 1];
    }
    io.vx = io.vy = io.vz = nullptr;
    return io;
}

template <int TX, int TY, int HA, int NT, bool NTS>
__global__ __launch_bounds__(NT) void k_batch_hydro(const BatchTerrain* __restrict__ terr,
                                                    const uint32_t* __restrict__ tiles,
                                                    int ntiles, uint32_t cur, const uint32_t* step_ptr,
                                                    uint32_t step_off, int rain) {
    __shared__ float lds[6 * HydroTile<TX, TY, HA, NT>::NLDS];
    const uint32_t step = *step_ptr + step_off;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t e = tiles[t];
        if (e == 0xffffffffu) continue;
        const int ti = int(e >> 24), by = int((e >> 12) & 0xfffu), bx = int(e & 0xfffu);
        // copy the descriptor into registers: a reference into global memory
        // would be re-read after every global store of the tile
        const BatchTerrain& T = terr[ti];
        const HydroIO io = batch_hydro_io(T, cur);
        const KParams k = T.k;
        const int n = T.n;
        hydro_tile<TX, TY, HA, NT, false, NTS>(lds, io, n, n, bx, by, step, uint32_t(ti), rain, k);
        __syncthreads();
    }
}

template <int TX, int TY, int NT, bool NTS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_thermal(
    const BatchTerrain* __restrict__ terr, const uint32_t* __restrict__ tiles, int ntiles,
    uint32_t cur, int do_thermal,
    int do_smooth) {
    __shared__ float lds[3 * ThermalTile<TX, TY>::N0];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t e = tiles[t];
        if (e == 0xffffffffu) continue;
        const int ti = int(e >> 24), by = int((e >> 12) & 0xfffu), bx = int(e & 0xfffu);
        const BatchTerrain& T = terr[ti];
        ThermIO io;
        const int cr = cur & 1, cd = (cur >> 1) & 1;
        io.r = T.plane[HG_P_ROCK][cr];
        io.ro = T.plane[HG_P_ROCK][cr ^ 1];
        io.d = T.plane[HG_P_DIRT][cd];
        io.dout = T.plane[HG_P_DIRT][cd ^ 1];
        const KParams k = T.k;
        const int n = T.n;
        thermal_tile<TX, TY, NT, false, NTS>(lds, io, n, n, bx, by, do_thermal, do_smooth, k);
        __syncthreads();
    }
}

// Compiled tile variants (the autotuner picks among them per grid size).
//                      TX   TY  HA   NT
#define HG_HYDRO_VARIANTS(X) \
    X(0,  64, 16, 2, 256)    \
    X(1,  64, 32, 2, 512)    \
    X(2, 128, 16, 2, 512)    \
    X(3,  64,  8, 2, 256)    \
    X(4,  64, 16, 3, 256)    \
    X(5, 128,  8, 2, 256)    \
    X(6,  64, 16, 2, 512)    \
    X(7, 128, 16, 2, 1024)
//                        TX   TY   NT
#define HG_THERMAL_VARIANTS(X) \
    X(0,  64, 16, 256)         \
    X(1,  64, 32, 256)         \
    X(2,  64, 32, 512)         \
    X(3,  64, 64, 512)         \
    X(4, 128, 32, 512)         \
    X(5,  32, 32, 256)         \
    X(6, 128, 16, 512)         \
    X(7,  64, 64, 1024)        \
    X(8, 128, 32, 1024)

}  // namespace

int hydro_variant_count() {
#define X(i, tx, ty, ha, nt) +1
    return 0 HG_HYDRO_VARIANTS(X);
#undef X
}

int thermal_variant_count() {
#define X(i, tx, ty, nt) +1
    return 0 HG_THERMAL_VARIANTS(X);
#undef X
}

bool hydro_variant_info(int v, int* tx, int* ty, int* halo, int* nt) {
    switch (v) {
#define X(i, TX_, TY_, HA_, NT_) \
    case i: *tx = TX_; *ty = TY_; *halo = HA_; *nt = NT_; return true;
        HG_HYDRO_VARIANTS(X)
#undef X
        default: return false;
    }
}

bool thermal_variant_info(int v, int* tx, int* ty, int* halo, int* nt) {
    switch (v) {
#define X(i, TX_, TY_, NT_) \
    case i: *tx = TX_; *ty = TY_; *halo = TH; *nt = NT_; return true;
        HG_THERMAL_VARIANTS(X)
#undef X
        default: return false;
    }
}

hipError_t launch_fused_hydro(int variant, const HydroIO& io, int nx, int ny, uint32_t step,
                              uint32_t terrain, bool rain, int xcd, const KParams& k, hipStream_t s) {
    switch (variant) {
#define X(i, TX_, TY_, HA_, NT_)                                                              \
    case i: {                                                                                 \
        const unsigned grid = unsigned((nx + TX_ - 1) / TX_) * unsigned((ny + TY_ - 1) / TY_); \
        const bool ex = io.vx || io.tb1;                                                      \
        const int r = rain ? 1 : 0;                                                           \
        if (ex && k.store_nt)                                                                 \
            k_fused_hydro<TX_, TY_, HA_, NT_, true, true><<<grid, NT_, 0, s>>>(io, nx, ny, step, terrain, r, xcd, k); \
        else if (ex)                                                                          \
            k_fused_hydro<TX_, TY_, HA_, NT_, true, false><<<grid, NT_, 0, s>>>(io, nx, ny, step, terrain, r, xcd, k); \
        else if (k.store_nt)                                                                  \
            k_fused_hydro<TX_, TY_, HA_, NT_, false, true><<<grid, NT_, 0, s>>>(io, nx, ny, step, terrain, r, xcd, k); \
        else                                                                                  \
            k_fused_hydro<TX_, TY_, HA_, NT_, false, false><<<grid, NT_, 0, s>>>(io, nx, ny, step, terrain, r, xcd, k); \
        break;                                                                                \
    }
        HG_HYDRO_VARIANTS(X)
#undef X
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_hydro(int variant, const BatchTerrain* terr, const uint32_t* tiles, int ntiles,
                              int grid, uint32_t cur, const uint32_t* step_ptr, uint32_t step_off,
                              bool rain, bool nts, hipStream_t s) {
    switch (variant) {
#define X(i, TX_, TY_, HA_, NT_)                                                                   \
    case i:                                                                                        \
        if (nts)                                                                                   \
            k_batch_hydro<TX_, TY_, HA_, NT_, true><<<grid, NT_, 0, s>>>(terr, tiles, ntiles, cur, step_ptr, \
                                                                      step_off, rain ? 1 : 0);    \
        else                                                                                       \
            k_batch_hydro<TX_, TY_, HA_, NT_, false><<<grid, NT_, 0, s>>>(terr, tiles, ntiles, cur, step_ptr, \
                                                                       step_off, rain ? 1 : 0);   \
        break;
        HG_HYDRO_VARIANTS(X)
#undef X
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_thermal(int variant, const BatchTerrain* terr, const uint32_t* tiles, int ntiles,
                                int grid, uint32_t cur, bool thermal, bool smooth, bool nts, hipStream_t s) {
    switch (variant) {
#define X(i, TX_, TY_, NT_)                                                                   \
    case i:                                                                                   \
        if (nts)                                                                              \
            k_batch_thermal<TX_, TY_, NT_, true><<<grid, NT_, 0, s>>>(terr, tiles, ntiles, cur, \
                                                                   thermal ? 1 : 0, smooth ? 1 : 0); \
        else                                                                                  \
            k_batch_thermal<TX_, TY_, NT_, false><<<grid, NT_, 0, s>>>(terr, tiles, ntiles, cur, \
                                                                    thermal ? 1 : 0, smooth ? 1 : 0); \
        break;
        HG_THERMAL_VARIANTS(X)
#undef X
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__global__ void k_step_advance(uint32_t* step_ptr, uint32_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step_ptr += n;
}

hipError_t launch_step_advance(uint32_t* step_ptr, uint32_t n, hipStream_t s) {
    k_step_advance<<<1, 64, 0, s>>>(step_ptr, n);
    return hipGetLastError();
}

int batch_hydro_blocks_per_cu(int variant) {
    int nb = 0;
    switch (variant) {
#define X(i, TX_, TY_, HA_, NT_)                                                                  \
    case i:                                                                                       \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_hydro<TX_, TY_, HA_, NT_, true>, \
                                                         NT_, 0) != hipSuccess)                  \
            nb = 1;                                                                               \
        break;
        HG_HYDRO_VARIANTS(X)
#undef X
        default: nb = 1;
    }
    return nb > 0 ? nb : 1;
}

int batch_thermal_blocks_per_cu(int variant) {
    int nb = 0;
    switch (variant) {
#define X(i, TX_, TY_, NT_)                                                                   \
    case i:                                                                                   \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_thermal<TX_, TY_, NT_, true>, \
                                                         NT_, 0) != hipSuccess)              \
            nb = 1;                                                                           \
        break;
        HG_THERMAL_VARIANTS(X)
#undef X
        default: nb = 1;
    }
    return nb > 0 ? nb : 1;
}

hipError_t launch_fused_thermal(int variant, const ThermIO& io, int nx, int ny, bool thermal,
                                bool smooth, int xcd, const KParams& k, hipStream_t s) {
    switch (variant) {
#define X(i, TX_, TY_, NT_)                                                                  \
    case i: {                                                                                \
        const unsigned grid = unsigned((nx + TX_ - 1) / TX_) * unsigned((ny + TY_ - 1) / TY_); \
        const int th = thermal ? 1 : 0, sm = smooth ? 1 : 0;                                 \
        if (io.tr5 && k.store_nt)                                                            \
            k_fused_thermal<TX_, TY_, NT_, true, true><<<grid, NT_, 0, s>>>(io, nx, ny, th, sm, xcd, k); \
        else if (io.tr5)                                                                     \
            k_fused_thermal<TX_, TY_, NT_, true, false><<<grid, NT_, 0, s>>>(io, nx, ny, th, sm, xcd, k); \
        else if (k.store_nt)                                                                 \
            k_fused_thermal<TX_, TY_, NT_, false, true><<<grid, NT_, 0, s>>>(io, nx, ny, th, sm, xcd, k); \
        else                                                                                 \
            k_fused_thermal<TX_, TY_, NT_, false, false><<<grid, NT_, 0, s>>>(io, nx, ny, th, sm, xcd, k); \
        break;                                                                               \
    }
        HG_THERMAL_VARIANTS(X)
#undef X
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace hgk
