// hg_adjoint.hip — reverse-mode gradient of a GRID rollout.
//
// No counterpart in the reference.  For a rollout of nsteps GRID steps from
// the context's state, given dL/dH_final (H = rock + dirt), computes dL/dH0
// (the initial rock height) and dL/d(Kc, Ks, Kd, Ke) (Ks, Kd shared by both
// layers).
//
// Structure:
//  * forward: the ordinary fused steps, saving the 9 state planes every
//    k_ckpt steps; the last two segments run through the taped kernels (next
//    item) at once, so they are not recomputed;
//  * reverse, one segment of k_ckpt steps at a time: recompute the segment
//    from its checkpoint with the fused kernels, which also record each
//    step's start state and its intermediates (water after rain and before
//    evaporation, velocity, eroded sediment, rock/dirt after erosion and after
//    each thermal layer); then, for each step from last to first, apply the
//    pass adjoints in reverse order to those recorded inputs.
//  * each pass adjoint is a vector-Jacobian product of that pass: a thread
//    recomputes its cell's forward arithmetic through the same inline code as
//    the forward kernels (hg_cell.hpp, with a tape of the branch decisions), so
//    the reverse sweep differentiates exactly the branches the forward took;
//    clamps, max/min, the talus test and the erosion cases contribute the
//    derivative of the branch taken (a subgradient at switching points).
//  * contributions a cell sends to its neighbours are gathered, not
//    scattered: each tile evaluates its cells and a one-cell ring, passes
//    what they send through LDS and writes every output once.  Only
//    semi-Lagrangian back-traces longer than one cell scatter, with float
//    atomic adds, after the gather.
//  * by default two pairs of pass adjoints run as one kernel each (erosion +
//    flux, smoothing + thermal layer 1), with the same arithmetic as the
//    kernels of the single passes.
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

// The tile a workgroup of a gather kernel works on: tile (x, y) of the
// GTX x GTY tiling, and its slot of the coefficient partial sums (the tile's
// row-major index, whatever order the tiles run in).
struct TileId {
    int x, y, slot;
};
// Tile of the b-th of nb workgroups over a tiling gx tiles wide.  Workgroups
// go to the 8 XCDs in turn; when gx and nb are multiples of 8, XCD x takes
// the column strip of tiles [x*gx/8, (x+1)*gx/8) row by row, so its L2 holds
// the tiles' left, right, upper and lower neighbours' staged rows.
// Otherwise row-major.
__device__ __forceinline__ TileId tile_at(int b, int nb, int gx) {
    int x, y;
    if ((gx & 7) == 0 && (nb & 7) == 0) {
        const int sw = gx >> 3, j = b >> 3;
        y = j / sw;
        x = (b & 7) * sw + (j - y * sw);
    } else {
        y = b / gx;
        x = b - y * gx;
    }
    return {x, y, y * gx + x};
}
// The workgroup's own tile, one per workgroup of a (tiles in x) x (tiles in
// y) grid.
__device__ __forceinline__ TileId block_tile() {
    return tile_at(int(blockIdx.y * gridDim.x + blockIdx.x), int(gridDim.x * gridDim.y), int(gridDim.x));
}

// Block reduction of up to 4 doubles, added by thread 0 to the block's own 4
// slots of `gpart` (one slot set per workgroup of the launch, so no atomics:
// same-address atomics from every workgroup serialise in one L2 channel).
// Each wave sums its lanes by a fixed butterfly, then thread 0 adds the wave
// sums in wave order, so the result does not depend on scheduling.  The host
// sums the slots in workgroup order at the end of the rollout.
template <int NT>
__device__ void reduce_coef(double* gpart, int slot, double a, double b, double c, double d) {
    static_assert(NT % 64 == 0, "whole waves");
    constexpr int NW = NT / 64;
    __shared__ double sh[NW][4];
    double v[4] = {a, b, c, d};
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] += __shfl_xor(v[j], m, 64);
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    if ((t & 63) == 0)
#pragma unroll
        for (int j = 0; j < 4; ++j) sh[t >> 6][j] = v[j];
    __syncthreads();
    if (t == 0) {
        double* q = gpart + 4 * size_t(slot);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double s = sh[0][j];
            for (int w = 1; w < NW; ++w) s += sh[w][j];
            if (s != 0.0) q[j] += s;
        }
    }
}

// ---------------------------------------------------------------------------
// Tiled gather.  A pass whose adjoint sends contributions to the 4 or 8
// neighbours of a cell is evaluated on a TX x TY tile plus a one-cell ring:
// phase 1 computes every ring and tile cell's contributions into LDS, phase 2
// sums, for each tile cell, its own part and what its neighbours send it, and
// writes each output once.  No atomics and no zeroed buffers.
//
// Loads come first and together: the inputs a ring cell reads at its
// neighbours are staged in LDS over the tile + 2-cell region (stage_walk),
// and a `pre` pass loads each ring cell's own inputs into registers for all
// of a thread's cells before any is evaluated.  Loads use clamped addresses;
// values off the grid are replaced by the pass's fill after the load.
// ---------------------------------------------------------------------------
constexpr int GTX = 64, GTY = 8, GNT = 256;   // 64 x 16 / 512 threads measured slower at 1024^2
constexpr int GRW = GTX + 2, GRH = GTY + 2, GNR = GRW * GRH;
constexpr int GIT = (GNR + GNT - 1) / GNT;          // ring/tile cells per thread
constexpr int SRW = GTX + 4, SRH = GTY + 4, SRN = SRW * SRH;
constexpr int SIT = (SRN + GNT - 1) / GNT;          // staged cells per thread

// Staged-region index of ring cell e.
__device__ __forceinline__ int sidx(int e) {
    const int ey = e / GRW, ex = e - ey * GRW;
    return (ey + 1) * SRW + ex + 1;
}

// fn(it, s, i, in, gx, gy) for the cells of the tile + H region: s its index
// in a (GTX + 2H) x (GTY + 2H) array, i the (clamped) global index.  Threads
// past the end redo the last cell.
template <int H, class F>
__device__ __forceinline__ void region_walk(const TileId& T, int nx, int ny, F&& fn) {
    constexpr int W = GTX + 2 * H, N = W * (GTY + 2 * H), IT = (N + GNT - 1) / GNT;
    const int x0 = T.x * GTX - H, y0 = T.y * GTY - H;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int s = min(int(threadIdx.x) + it * GNT, N - 1);
        const int sy = s / W, sx = s - sy * W;
        const int gx = x0 + sx, gy = y0 + sy;
        const bool in = gx >= 0 && gy >= 0 && gx < nx && gy < ny;
        fn(it, s, at(min(max(gx, 0), nx - 1), min(max(gy, 0), ny - 1), nx), in, gx, gy);
    }
}

// fn(it, s, i, in) for the cells of the tile + 2-cell region: s its index in
// the staged (SRW x SRH) arrays, i the (clamped) global index, in whether it is
// on the grid.
template <class F>
__device__ __forceinline__ void stage_walk(const TileId& T, int nx, int ny, F&& fn) {
    static_assert((SRN + GNT - 1) / GNT == SIT, "one iteration per staged cell of a thread");
    region_walk<2>(T, nx, ny, [&](int it, int s, size_t i, bool in, int, int) { fn(it, s, i, in); });
}

// A thread's cells of the tile + ring, numbered j = threadIdx.x + it * GNT:
// the GTX * GTY tile cells first, row by row (whole waves, GTIT iterations),
// then the ring (bottom row, top row, then the left and right columns).
constexpr int GTIT = GTX * GTY / GNT;
static_assert(GTX * GTY % GNT == 0, "the tile fills whole iterations");
static_assert(GIT == GTIT + 1 && GNR - GTX * GTY <= GNT, "the ring fits one iteration");
__device__ __forceinline__ void ring_cell(int it, int& ex, int& ey) {
    const int j = int(threadIdx.x) + it * GNT;
    if (it < GTIT) {
        ey = j / GTX + 1;
        ex = j - (ey - 1) * GTX + 1;
    } else {
        const int r = min(j - GTX * GTY, GNR - GTX * GTY - 1);   // past the end: the last (loads only)
        if (r < 2 * GRW) {
            ey = r < GRW ? 0 : GRH - 1;
            ex = r < GRW ? r : r - GRW;
        } else {
            const int c = r - 2 * GRW;
            ey = (c >> 1) + 1;
            ex = (c & 1) ? GRW - 1 : 0;
        }
    }
}

// pre(it, e, i): loads of ring cell e (i its clamped global index) for all of
// a thread's cells; cell(it, e, px, py, in, tile): phase 1; gather(it, e, px,
// py): phase 2 for the tile cells on the grid.  Phase 2 visits the tile cells
// with the same (thread, it) as phase 1, so what a cell computes for itself
// can stay in registers indexed by `it` (it < GTIT); only what it sends to
// its neighbours goes through LDS.
template <class Pre, class Cell, class Gather>
__device__ __forceinline__ void tile_gather(const TileId& T, int nx, int ny, Pre&& pre, Cell&& cell,
                                            Gather&& gather) {
    const int x0 = T.x * GTX, y0 = T.y * GTY;
#pragma unroll
    for (int it = 0; it < GIT; ++it) {
        int ex, ey;
        ring_cell(it, ex, ey);
        const int gx = x0 - 1 + ex, gy = y0 - 1 + ey;
        pre(it, ey * GRW + ex, at(min(max(gx, 0), nx - 1), min(max(gy, 0), ny - 1), nx));
    }
    __syncthreads();   // staged LDS regions are complete
#pragma unroll
    for (int it = 0; it < GIT; ++it) {
        if (it == GTIT && int(threadIdx.x) >= GNR - GTX * GTY) break;   // past the end of the ring
        int ex, ey;
        ring_cell(it, ex, ey);
        const int gx = x0 - 1 + ex, gy = y0 - 1 + ey;
        const bool in = gx >= 0 && gy >= 0 && gx < nx && gy < ny;
        cell(it, ey * GRW + ex, gx, gy, in, it < GTIT);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < GTIT; ++it) {
        int ex, ey;
        ring_cell(it, ex, ey);
        const int gx = x0 - 1 + ex, gy = y0 - 1 + ey;
        if (gx < nx && gy < ny) gather(it, ey * GRW + ex, gx, gy);
    }
}

// tile_gather for kernels whose staged inputs and sends share LDS: cell(it,
// e, px, py, in, tile) keeps what a cell sends in registers; after a barrier
// (every read of the staged regions done) send(it, e) writes it to LDS, and
// after another the gather runs.
template <class Pre, class Cell, class Send, class Gather>
__device__ __forceinline__ void tile_gather_late(const TileId& T, int nx, int ny, Pre&& pre, Cell&& cell,
                                                 Send&& send, Gather&& gather) {
    const int x0 = T.x * GTX, y0 = T.y * GTY;
#pragma unroll
    for (int it = 0; it < GIT; ++it) {
        int ex, ey;
        ring_cell(it, ex, ey);
        const int gx = x0 - 1 + ex, gy = y0 - 1 + ey;
        pre(it, ey * GRW + ex, at(min(max(gx, 0), nx - 1), min(max(gy, 0), ny - 1), nx));
    }
    __syncthreads();   // staged LDS regions are complete
#pragma unroll
    for (int it = 0; it < GIT; ++it) {
        if (it == GTIT && int(threadIdx.x) >= GNR - GTX * GTY) break;   // past the end of the ring
        int ex, ey;
        ring_cell(it, ex, ey);
        const int gx = x0 - 1 + ex, gy = y0 - 1 + ey;
        const bool in = gx >= 0 && gy >= 0 && gx < nx && gy < ny;
        cell(it, ey * GRW + ex, gx, gy, in, it < GTIT);
    }
    __syncthreads();   // the staged regions are no longer read
#pragma unroll
    for (int it = 0; it < GIT; ++it) {
        if (it == GTIT && int(threadIdx.x) >= GNR - GTX * GTY) break;
        int ex, ey;
        ring_cell(it, ex, ey);
        send(it, ey * GRW + ex);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < GTIT; ++it) {
        int ex, ey;
        ring_cell(it, ex, ey);
        const int gx = x0 - 1 + ex, gy = y0 - 1 + ey;
        if (gx < nx && gy < ny) gather(it, ey * GRW + ex, gx, gy);
    }
}

// ---- hydro_flux ----------------------------------------------------------
// VJP of the flux pass at an in-grid cell p = (px, py).  f: p's old outflow
// (L, R, T, B); b: its water after rain; w, wn: its own and its L, R, T, B
// neighbours' total height; fin: the neighbours' old outflow towards p; lf,
// ld2, lvx/lvy/lvz: the adjoints of p's new outflow, water and velocity.
// Adds to lfo (p's old outflow), lfin (the neighbours' old outflow towards
// p), lwn (their total height), ld1 (p's water after rain) and lwp (p's own
// total height), all zero on entry.  Nothing is sent off the grid.
__device__ __forceinline__ void flux_vjp(float f[4], float b, float w, const float wn[4], const float fin[4],
                                         int px, int py, int nx, int ny, const float lf[4], float ld2,
                                         float lvx, float lvy, float lvz, const KParams& k, float lfo[4],
                                         float lfin[4], float lwn[4], float& ld1, float& lwp) {
    const unsigned wall = (px == 0 ? 1u : 0u) | (px == nx - 1 ? 2u : 0u) | (py == 0 ? 4u : 0u) |
                          (py == ny - 1 ? 8u : 0u);
    FluxTape tp;
    float vx, vy, vz;
    (void)flux_cell_t(f, b, w, wn, fin, wall, k, vx, vy, vz, tp);
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
    // contributions to neighbours outside the grid are dropped
    if (px == 0) { lfin[0] = 0.0f; lwn[0] = 0.0f; }
    if (px == nx - 1) { lfin[1] = 0.0f; lwn[1] = 0.0f; }
    if (py == ny - 1) { lfin[2] = 0.0f; lwn[2] = 0.0f; }
    if (py == 0) { lfin[3] = 0.0f; lwn[3] = 0.0f; }
}

// x: flux inputs (rock, dirt, water after rain, old flux); lo: adjoint of the
// flux outputs (water, flux, velocity).  Writes li's water and flux (fresh)
// and adds the total-height adjoint to li's rock and dirt.
__global__ __launch_bounds__(GNT) void k_adj_flux(Planes x, Planes lo, Planes li, int nx, int ny, KParams k) {
    const TileId T = block_tile();
    // sent to neighbours, per ring/tile cell: inflow adj (4), total-height adj (4);
    // kept per thread: outflow adj (4), own total-height adj, water adj
    __shared__ float s_fin[4][GNR], s_wn[4][GNR];
    float o_fo[GTIT][4], o_wp[GTIT], o_d1[GTIT];
    __shared__ float t_w[SRN], t_fl[SRN], t_fr[SRN], t_ft[SRN], t_fb[SRN];
    const float* R = x.p[HG_P_ROCK];
    const float* D = x.p[HG_P_DIRT];
    const float* B = x.p[HG_P_WATER];
    {
        // total height (out of grid: HG_OOB_HEIGHT) and old flux (0) of the region
        float w[SIT], fl[SIT], fr[SIT], ft[SIT], fb[SIT];
        stage_walk(T, nx, ny, [&](int it, int, size_t i, bool in) {
            const float t = (R[i] + D[i]) + B[i];
            const float a = x.p[HG_P_FLUX_L][i], b = x.p[HG_P_FLUX_R][i], c = x.p[HG_P_FLUX_T][i],
                        d = x.p[HG_P_FLUX_B][i];
            w[it] = in ? t : HG_OOB_HEIGHT;
            fl[it] = in ? a : 0.0f; fr[it] = in ? b : 0.0f; ft[it] = in ? c : 0.0f; fb[it] = in ? d : 0.0f;
        });
        stage_walk(T, nx, ny, [&](int it, int s, size_t, bool) {
            t_w[s] = w[it]; t_fl[s] = fl[it]; t_fr[s] = fr[it]; t_ft[s] = ft[it]; t_fb[s] = fb[it];
        });
    }
    float pb[GIT], plf[GIT][4], pld2[GIT], plv[GIT][3];
    tile_gather(T, nx, ny,
        [&](int it, int, size_t i) {
            pb[it] = B[i];
            plf[it][0] = lo.p[HG_P_FLUX_L][i]; plf[it][1] = lo.p[HG_P_FLUX_R][i];
            plf[it][2] = lo.p[HG_P_FLUX_T][i]; plf[it][3] = lo.p[HG_P_FLUX_B][i];
            pld2[it] = lo.p[HG_P_WATER][i];
            plv[it][0] = lo.p[HG_P_VEL_U][i]; plv[it][1] = lo.p[HG_P_VEL_V][i]; plv[it][2] = lo.p[HG_P_VEL_DEPTH][i];
        },
        [&](int it, int e, int px, int py, bool in, bool) {
            float lfo[4] = {0, 0, 0, 0}, lfin[4] = {0, 0, 0, 0}, lwn[4] = {0, 0, 0, 0};
            float ld1 = 0.0f, lwp = 0.0f;
            if (in) {
                const int sc = sidx(e);
                float f[4] = {t_fl[sc], t_fr[sc], t_ft[sc], t_fb[sc]};
                const float w = t_w[sc];
                const float wn[4] = {t_w[sc - 1], t_w[sc + 1], t_w[sc + SRW], t_w[sc - SRW]};
                const float fin[4] = {t_fr[sc - 1], t_fl[sc + 1], t_fb[sc + SRW], t_ft[sc - SRW]};
                flux_vjp(f, pb[it], w, wn, fin, px, py, nx, ny, plf[it], pld2[it], plv[it][0], plv[it][1],
                         plv[it][2], k, lfo, lfin, lwn, ld1, lwp);
            }
            for (int j = 0; j < 4; ++j) {
                s_fin[j][e] = lfin[j];
                s_wn[j][e] = lwn[j];
            }
            if (it < GTIT) {                                  // a tile cell keeps its own part
                for (int j = 0; j < 4; ++j) o_fo[it][j] = lfo[j];
                o_wp[it] = lwp;
                o_d1[it] = ld1;
            }
        },
        [&](int it, int e, int px, int py) {
            const size_t i = at(px, py, nx);
            // neighbour q = p - 1 sends its R-inflow adjoint (index 1) to p's L
            // outflow, q = p + 1 its index 0 to p's R, q = p - nx its index 2
            // to p's B, q = p + nx its index 3 to p's T
            li.p[HG_P_FLUX_L][i] = o_fo[it][0] + s_fin[1][e - 1];
            li.p[HG_P_FLUX_R][i] = o_fo[it][1] + s_fin[0][e + 1];
            li.p[HG_P_FLUX_T][i] = o_fo[it][2] + s_fin[3][e + GRW];
            li.p[HG_P_FLUX_B][i] = o_fo[it][3] + s_fin[2][e - GRW];
            const float lw = o_wp[it] + s_wn[0][e + 1] + s_wn[1][e - 1] + s_wn[2][e - GRW] + s_wn[3][e + GRW];
            li.p[HG_P_WATER][i] = o_d1[it] + lw;               // w = rock + dirt + water
            li.p[HG_P_ROCK][i] += lw;
            li.p[HG_P_DIRT][i] += lw;
        });
}

// ---- hydro_erosion -------------------------------------------------------
// VJP of the erosion pass at an in-grid cell.  r_in .. vz: the cell's rock,
// dirt, sediment and velocity before the pass; hl, hr, ht, hb: the total
// heights (rock + dirt) of its x-1, x+1, y+1 and y-1 neighbours; lr .. ls1: the
// adjoints of its rock, dirt and sediment after the pass.
struct ErosionVjp {
    float lr, ld, ls0, ls1;     // adjoints of the cell's rock, dirt, sediment before the pass
    float lvx, lvy, lvz;        // and of its velocity
    float ldx, ldz;             // slope adjoints: +ldx to x+1, -ldx to x-1, +ldz to y+1, -ldz to y-1
    double cKc, cKs, cKd;       // coefficient terms
};
__device__ __forceinline__ ErosionVjp erosion_vjp(float r_in, float d_in, float s0_in, float s1_in, float vx,
                                                  float vy, float vz, float hl, float hr, float ht, float hb,
                                                  float lr, float ld, float ls0, float ls1, const KParams& k) {
    ErosionVjp o;
    float r = r_in, d = d_in, s0 = s0_in, s1 = s1_in;
    ErosionTape tp;
    erosion_cell_t(r, d, s0, s1, vx, vy, vz, hl, hr, ht, hb, k, tp);
    double cKc = 0, cKs = 0, cKd = 0;
    // conversion s1 += s0 kc, s0 -= s0 kc
    const float kc = k.Kconv * k.dt;
    const float ls0c = ls0 * (1.0f - kc) + ls1 * kc, ls1c = ls1;
    // rock layer: (r, s0, c0) -> (r', s0c)
    float lc0 = 0.0f, lcap0 = 0.0f, lcap = 0.0f;
    float ls0_in = ls0c;
    float lr_in = lr;
    if (tp.case0 == 0 || tp.case0 == 2) {
        const float a = k.dt * (tp.case0 == 0 ? k.Ks[0] : k.Kd[0]);
        ls0_in = lr * a + ls0c * (1.0f - a);
        lc0 = (ls0c - lr) * a;
        if (tp.case0 == 0) cKs += double(k.dt) * double(tp.c0 - s0_in) * double(ls0c - lr);
        else cKd += double(k.dt) * double(s0_in - tp.c0) * double(lr - ls0c);
    } else if (tp.case0 == 1) {               // rock exhausted: r' = 0, s0c = s0 + r
        lr_in = ls0c;
        ls0_in = ls0c;
    }
    if (tp.case0 != 3 && tp.c0 > 0.0f) {      // c0 = max(0, cap0 - cap)
        lcap0 += lc0;
        lcap -= lc0;
    }
    // dirt layer: (d, s1, c1) -> (d', s1c); cap = d when exhausted
    float lc1 = 0.0f;
    float ls1_in = ls1c;
    float ld_in = ld;
    if (tp.case1 == 0 || tp.case1 == 2) {
        const float a = k.dt * (tp.case1 == 0 ? k.Ks[1] : k.Kd[1]);
        ls1_in = ld * a + ls1c * (1.0f - a);
        lc1 = (ls1c - ld) * a;
        if (tp.case1 == 0) cKs += double(k.dt) * double(tp.c1 - s1_in) * double(ls1c - ld);
        else cKd += double(k.dt) * double(s1_in - tp.c1) * double(ld - ls1c);
    } else {                                  // exhausted: d' = 0, s1c = s1 + d, cap = d
        ld_in = ls1c + lcap;
        ls1_in = ls1c;
    }
    if (tp.c1 > 0.0f) lcap0 += lc1;           // c1 = max(0, cap0)
    // cap0 = Kc * max(0.02, sin_a) * ero
    const float m = fmaxf(0.02f, tp.sin_a);
    const float ero = tp.slow ? tp.ero0 * (1.0f - tp.sm) : tp.ero0;
    cKc += double(lcap0) * double(m) * double(ero);
    const float lm = lcap0 * k.Kc * ero, lero = lcap0 * k.Kc * m;
    float ldx = 0.0f, ldz = 0.0f, lvx = 0.0f, lvy = 0.0f, lvz = 0.0f;
    if (tp.sin_a > 0.02f && tp.sin_a > 0.0f) {  // sin_a = sqrt(g2 / (g2 + 4))
        const float q = tp.g2 + 4.0f;
        const float lg2 = lm * 2.0f * rcp(tp.sin_a * q * q);
        ldx = 2.0f * tp.dx * lg2;
        ldz = 2.0f * tp.dz * lg2;
    }
    float lero0 = lero;
    if (tp.slow) {
        lero0 = lero * (1.0f - tp.sm);
        if (tp.t_inner) {                     // sm = t^2 (3 - 2t), t = (vz - 1e-3) / (5e-4 - 1e-3)
            const float lsm = -lero * tp.ero0;
            lvz = lsm * 6.0f * tp.tt * (1.0f - tp.tt) * (1.0f / (5e-4f - 1e-3f));
        }
    }
    if (tp.ero0 > 0.0f) {
        const float inv = rcp(tp.ero0);
        lvx = lero0 * vx * inv;
        lvy = lero0 * vy * inv;
    }
    o.lr = lr_in; o.ld = ld_in; o.ls0 = ls0_in; o.ls1 = ls1_in;
    o.lvx = lvx; o.lvy = lvy; o.lvz = lvz;
    o.ldx = ldx; o.ldz = ldz;
    o.cKc = cKc; o.cKs = cKs; o.cKd = cKd;
    return o;
}

// Writes li's rock, dirt and sediment (fresh) and adds to li's velocity.
__global__ __launch_bounds__(GNT) void k_adj_erosion(Planes x, Planes lo, Planes li, int nx, int ny, KParams k,
                                                     double* gpart) {
    const TileId T = block_tile();
    // sent to neighbours: the slope adjoints; kept per thread: the cell's own
    // rock, dirt, sediment and velocity adjoints
    __shared__ float s_dx[GNR], s_dz[GNR];
    __shared__ float t_r[SRN], t_d[SRN];
    float o_r[GTIT], o_d[GTIT], o_s0[GTIT], o_s1[GTIT], o_vx[GTIT], o_vy[GTIT], o_vz[GTIT];
    const float* R = x.p[HG_P_ROCK];
    const float* D = x.p[HG_P_DIRT];
    {
        // rock and dirt of the region (0 out of the grid)
        float a[SIT], b[SIT];
        stage_walk(T, nx, ny, [&](int it, int, size_t i, bool in) {
            const float r = R[i], d = D[i];
            a[it] = in ? r : 0.0f;
            b[it] = in ? d : 0.0f;
        });
        stage_walk(T, nx, ny, [&](int it, int s, size_t, bool) { t_r[s] = a[it]; t_d[s] = b[it]; });
    }
    auto rd = [&](int sc) -> float { return t_r[sc] + t_d[sc]; };
    float p0[GIT], p1[GIT], pvx[GIT], pvy[GIT], pvz[GIT], plr[GIT], pld[GIT], pl0[GIT], pl1[GIT];
    double gKc = 0, gKs = 0, gKd = 0;
    tile_gather(T, nx, ny,
        [&](int it, int, size_t i) {
            p0[it] = x.p[HG_P_SED_ROCK][i]; p1[it] = x.p[HG_P_SED_DIRT][i];
            pvx[it] = x.p[HG_P_VEL_U][i]; pvy[it] = x.p[HG_P_VEL_V][i]; pvz[it] = x.p[HG_P_VEL_DEPTH][i];
            plr[it] = lo.p[HG_P_ROCK][i]; pld[it] = lo.p[HG_P_DIRT][i];
            pl0[it] = lo.p[HG_P_SED_ROCK][i]; pl1[it] = lo.p[HG_P_SED_DIRT][i];
        },
        [&](int it, int e, int px, int py, bool in, bool tile) {
            ErosionVjp ev{};
            if (in) {
                const int sc = sidx(e);
                ev = erosion_vjp(t_r[sc], t_d[sc], p0[it], p1[it], pvx[it], pvy[it], pvz[it], rd(sc - 1),
                                 rd(sc + 1), rd(sc + SRW), rd(sc - SRW), plr[it], pld[it], pl0[it], pl1[it], k);
                if (tile) {                               // coefficient terms counted once
                    gKc += ev.cKc; gKs += ev.cKs; gKd += ev.cKd;
                }
            }
            s_dx[e] = ev.ldx;
            s_dz[e] = ev.ldz;
            if (it < GTIT) {                                  // a tile cell keeps its own part
                o_r[it] = ev.lr;
                o_d[it] = ev.ld;
                o_s0[it] = ev.ls0;
                o_s1[it] = ev.ls1;
                o_vx[it] = ev.lvx;
                o_vy[it] = ev.lvy;
                o_vz[it] = ev.lvz;
            }
        },
        [&](int it, int e, int px, int py) {
            const size_t i = at(px, py, nx);
            const float nb = (s_dx[e - 1] - s_dx[e + 1]) + (s_dz[e - GRW] - s_dz[e + GRW]);
            li.p[HG_P_ROCK][i] = o_r[it] + nb;
            li.p[HG_P_DIRT][i] = o_d[it] + nb;
            li.p[HG_P_SED_ROCK][i] = o_s0[it];
            li.p[HG_P_SED_DIRT][i] = o_s1[it];
            li.p[HG_P_VEL_U][i] += o_vx[it];
            li.p[HG_P_VEL_V][i] += o_vy[it];
            li.p[HG_P_VEL_DEPTH][i] += o_vz[it];
        });
    reduce_coef<GNT>(gpart, T.slot, gKc, gKs, gKd, 0.0);
}

// ---- hydro_erosion and hydro_flux in one launch ----------------------------
// k_adj_erosion followed by k_adj_flux, with the same arithmetic in the same
// order.  Both passes read the step's start rock and dirt, and the erosion's
// velocity adjoint is the cell's own, so each ring and tile cell evaluates the
// erosion VJP, completes its velocity adjoint in registers (the transport's
// part plus the erosion's) and evaluates the flux VJP; the gather then sums
// what both passes send.  x: rock, dirt, water after rain, old flux, sediment
// and velocity of the step; lo: the adjoints after erosion.  Writes li's rock,
// dirt, sediment, water and flux (fresh).  The velocity adjoint ends here: it
// is not written, and the transport VJP of the step before writes it anew.
__global__ __launch_bounds__(GNT) void k_adj_erosion_flux(Planes x, Planes lo, Planes li, int nx, int ny,
                                                          KParams k, double* gpart) {
    const TileId T = block_tile();
    // One LDS array (26.4 KB), first for the inputs staged on the tile + 2-cell
    // region: rock and dirt (0 off the grid), total height (HG_OOB_HEIGHT off
    // the grid) and old flux (0); then, once every cell is evaluated, for what
    // the cells send their neighbours: erosion's slope adjoints, flux's inflow
    // (4) and total-height (4) adjoints.
    constexpr int NU = 7 * SRN > 10 * GNR ? 7 * SRN : 10 * GNR;
    __shared__ float u[NU];
    float* t_r = u;
    float* t_d = u + SRN;
    float* t_w = u + 2 * SRN;
    float* t_fl = u + 3 * SRN;
    float* t_fr = u + 4 * SRN;
    float* t_ft = u + 5 * SRN;
    float* t_fb = u + 6 * SRN;
    float* s_dx = u;
    float* s_dz = u + GNR;
    float (*s_fin)[GNR] = reinterpret_cast<float (*)[GNR]>(u + 2 * GNR);
    float (*s_wn)[GNR] = reinterpret_cast<float (*)[GNR]>(u + 6 * GNR);
    float o_r[GTIT], o_d[GTIT], o_s0[GTIT], o_s1[GTIT], o_fo[GTIT][4], o_wp[GTIT], o_d1[GTIT];
    float snd[GIT][10];                          // what each of a thread's cells sends
    const float* R = x.p[HG_P_ROCK];
    const float* D = x.p[HG_P_DIRT];
    const float* B = x.p[HG_P_WATER];
    {
        float a[SIT], b[SIT], w[SIT], fl[SIT], fr[SIT], ft[SIT], fb[SIT];
        stage_walk(T, nx, ny, [&](int it, int, size_t i, bool in) {
            const float r = R[i], d = D[i];
            const float t = (r + d) + B[i];
            const float f0 = x.p[HG_P_FLUX_L][i], f1 = x.p[HG_P_FLUX_R][i], f2 = x.p[HG_P_FLUX_T][i],
                        f3 = x.p[HG_P_FLUX_B][i];
            a[it] = in ? r : 0.0f;
            b[it] = in ? d : 0.0f;
            w[it] = in ? t : HG_OOB_HEIGHT;
            fl[it] = in ? f0 : 0.0f; fr[it] = in ? f1 : 0.0f; ft[it] = in ? f2 : 0.0f; fb[it] = in ? f3 : 0.0f;
        });
        stage_walk(T, nx, ny, [&](int it, int s, size_t, bool) {
            t_r[s] = a[it]; t_d[s] = b[it]; t_w[s] = w[it];
            t_fl[s] = fl[it]; t_fr[s] = fr[it]; t_ft[s] = ft[it]; t_fb[s] = fb[it];
        });
    }
    auto rd = [&](int sc) -> float { return t_r[sc] + t_d[sc]; };
    float p0[GIT], p1[GIT], pvx[GIT], pvy[GIT], pvz[GIT], plr[GIT], pld[GIT], pl0[GIT], pl1[GIT];
    float pb[GIT], plf[GIT][4], pld2[GIT], plv[GIT][3];
    double gKc = 0, gKs = 0, gKd = 0;
    tile_gather_late(T, nx, ny,
        [&](int it, int, size_t i) {
            p0[it] = x.p[HG_P_SED_ROCK][i]; p1[it] = x.p[HG_P_SED_DIRT][i];
            pvx[it] = x.p[HG_P_VEL_U][i]; pvy[it] = x.p[HG_P_VEL_V][i]; pvz[it] = x.p[HG_P_VEL_DEPTH][i];
            plr[it] = lo.p[HG_P_ROCK][i]; pld[it] = lo.p[HG_P_DIRT][i];
            pl0[it] = lo.p[HG_P_SED_ROCK][i]; pl1[it] = lo.p[HG_P_SED_DIRT][i];
            pb[it] = B[i];
            plf[it][0] = lo.p[HG_P_FLUX_L][i]; plf[it][1] = lo.p[HG_P_FLUX_R][i];
            plf[it][2] = lo.p[HG_P_FLUX_T][i]; plf[it][3] = lo.p[HG_P_FLUX_B][i];
            pld2[it] = lo.p[HG_P_WATER][i];
            plv[it][0] = lo.p[HG_P_VEL_U][i]; plv[it][1] = lo.p[HG_P_VEL_V][i]; plv[it][2] = lo.p[HG_P_VEL_DEPTH][i];
        },
        [&](int it, int e, int px, int py, bool in, bool tile) {
            ErosionVjp ev{};
            float lfo[4] = {0, 0, 0, 0}, lfin[4] = {0, 0, 0, 0}, lwn[4] = {0, 0, 0, 0};
            float ld1 = 0.0f, lwp = 0.0f;
            if (in) {
                const int sc = sidx(e);
                ev = erosion_vjp(t_r[sc], t_d[sc], p0[it], p1[it], pvx[it], pvy[it], pvz[it], rd(sc - 1),
                                 rd(sc + 1), rd(sc + SRW), rd(sc - SRW), plr[it], pld[it], pl0[it], pl1[it], k);
                if (tile) {                               // coefficient terms counted once
                    gKc += ev.cKc; gKs += ev.cKs; gKd += ev.cKd;
                }
                // the velocity adjoint after the flux pass: the transport's part
                // plus the erosion's (k_adj_erosion adds the latter in memory)
                const float lvx = plv[it][0] + ev.lvx, lvy = plv[it][1] + ev.lvy, lvz = plv[it][2] + ev.lvz;
                float f[4] = {t_fl[sc], t_fr[sc], t_ft[sc], t_fb[sc]};
                const float wn[4] = {t_w[sc - 1], t_w[sc + 1], t_w[sc + SRW], t_w[sc - SRW]};
                const float fin[4] = {t_fr[sc - 1], t_fl[sc + 1], t_fb[sc + SRW], t_ft[sc - SRW]};
                flux_vjp(f, pb[it], t_w[sc], wn, fin, px, py, nx, ny, plf[it], pld2[it], lvx, lvy, lvz, k, lfo,
                         lfin, lwn, ld1, lwp);
            }
            snd[it][0] = ev.ldx;
            snd[it][1] = ev.ldz;
            for (int j = 0; j < 4; ++j) {
                snd[it][2 + j] = lfin[j];
                snd[it][6 + j] = lwn[j];
            }
            if (it < GTIT) {                                  // a tile cell keeps its own part
                o_r[it] = ev.lr;
                o_d[it] = ev.ld;
                o_s0[it] = ev.ls0;
                o_s1[it] = ev.ls1;
                for (int j = 0; j < 4; ++j) o_fo[it][j] = lfo[j];
                o_wp[it] = lwp;
                o_d1[it] = ld1;
            }
        },
        [&](int it, int e) {
            s_dx[e] = snd[it][0];
            s_dz[e] = snd[it][1];
            for (int j = 0; j < 4; ++j) {
                s_fin[j][e] = snd[it][2 + j];
                s_wn[j][e] = snd[it][6 + j];
            }
        },
        [&](int it, int e, int px, int py) {
            const size_t i = at(px, py, nx);
            // rock and dirt after the erosion VJP, as k_adj_erosion writes them
            const float nb = (s_dx[e - 1] - s_dx[e + 1]) + (s_dz[e - GRW] - s_dz[e + GRW]);
            const float lr = o_r[it] + nb, ld = o_d[it] + nb;
            li.p[HG_P_SED_ROCK][i] = o_s0[it];
            li.p[HG_P_SED_DIRT][i] = o_s1[it];
            li.p[HG_P_FLUX_L][i] = o_fo[it][0] + s_fin[1][e - 1];
            li.p[HG_P_FLUX_R][i] = o_fo[it][1] + s_fin[0][e + 1];
            li.p[HG_P_FLUX_T][i] = o_fo[it][2] + s_fin[3][e + GRW];
            li.p[HG_P_FLUX_B][i] = o_fo[it][3] + s_fin[2][e - GRW];
            const float lw = o_wp[it] + s_wn[0][e + 1] + s_wn[1][e - 1] + s_wn[2][e - GRW] + s_wn[3][e + GRW];
            li.p[HG_P_WATER][i] = o_d1[it] + lw;               // w = rock + dirt + water
            li.p[HG_P_ROCK][i] = lr + lw;
            li.p[HG_P_DIRT][i] = ld + lw;
        });
    reduce_coef<GNT>(gpart, T.slot, gKc, gKs, gKd, 0.0);
}

// ---- sediment_transport --------------------------------------------------
// The bilinear gather's adjoint is a scatter to the four taps of each cell's
// back-trace.  A cell whose taps lie within one cell of it ("near") is
// gathered by the taps themselves (k_adj_transport); the others scatter with
// atomics afterwards (k_adj_transport_far).
struct Taps {
    int ix, iy, ix1, iy1;
    float w00, w10, w01, w11;
    bool near;
};
__device__ __forceinline__ Taps taps_of(int qx, int qy, float vx, float vy, int nx, int ny, float dt) {
    Taps t;
    float bx, by;
    back_trace(qx, qy, vx, vy, dt, nx, ny, bx, by);
    t.ix = int(bx); t.iy = int(by);
    const float fx = bx - float(t.ix), fy = by - float(t.iy);
    t.ix1 = min(t.ix + 1, nx - 1); t.iy1 = min(t.iy + 1, ny - 1);
    t.w00 = (1.0f - fx) * (1.0f - fy); t.w10 = fx * (1.0f - fy); t.w01 = (1.0f - fx) * fy; t.w11 = fx * fy;
    t.near = t.ix >= qx - 1 && t.ix1 <= qx + 1 && t.iy >= qy - 1 && t.iy1 <= qy + 1;
    return t;
}

// LDS of one transport tile, in floats.
constexpr int TRANSPORT_LDS = 6 * GNR + (GNR + 3) / 4;

// The transport VJP's near gather on tile T, with `lds` (TRANSPORT_LDS
// floats) as its LDS.
__device__ __forceinline__ void adj_transport_tile(const TileId& T, float* lds, Planes x, Planes lo, Planes li,
                                                   int nx, int ny, const KParams& k, double* gpart) {
    // per tile/ring cell q: its 4 bilinear weights, its taps as offsets from
    // q (bits 0-1: ix - qx + 1, 2-3: ix1 - qx + 1, 4-5: iy - qy + 1, 6-7:
    // iy1 - qy + 1; 0 when q is off the grid or its taps are not near), and the
    // sediment adjoints it spreads over them
    float (*s_w)[GNR] = reinterpret_cast<float (*)[GNR]>(lds);
    float* s_l0 = lds + 4 * GNR;
    float* s_l1 = lds + 5 * GNR;
    unsigned char* s_tap = reinterpret_cast<unsigned char*>(lds + 6 * GNR);
    const float* VX = x.p[HG_P_VEL_U];
    const float* VY = x.p[HG_P_VEL_V];
    double gKe = 0;
    float pvx[GIT], pvy[GIT], pl0[GIT], pl1[GIT];
    // a tile cell's own water and its adjoint, and the sediment at its own
    // back-trace taps (both layers: s00, s10, s01, s11), loaded before the
    // gather phase needs them
    float plb[GTIT], pb[GTIT], ps[GTIT][8];
    tile_gather(T, nx, ny,
        [&](int it, int, size_t i) {
            pvx[it] = VX[i]; pvy[it] = VY[i];
            pl0[it] = lo.p[HG_P_SED_ROCK][i]; pl1[it] = lo.p[HG_P_SED_DIRT][i];
            if (it < GTIT) {
                plb[it] = lo.p[HG_P_WATER][i];
                pb[it] = x.p[HG_P_WATER][i];
            }
        },
        [&](int it, int e, int qx, int qy, bool in, bool) {
            unsigned code = 0;
            float w[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (in) {
                const Taps t = taps_of(qx, qy, pvx[it], pvy[it], nx, ny, k.dt);
                if (t.near) {
                    code = unsigned(t.ix - qx + 1) | (unsigned(t.ix1 - qx + 1) << 2) | (unsigned(t.iy - qy + 1) << 4) |
                           (unsigned(t.iy1 - qy + 1) << 6);
                    w[0] = t.w00; w[1] = t.w10; w[2] = t.w01; w[3] = t.w11;
                }
                if (it < GTIT) {
                    const size_t j00 = at(t.ix, t.iy, nx), j10 = at(t.ix1, t.iy, nx), j01 = at(t.ix, t.iy1, nx),
                                 j11 = at(t.ix1, t.iy1, nx);
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        const float* S = x.p[c == 0 ? HG_P_SED_ROCK : HG_P_SED_DIRT];
                        ps[it][4 * c + 0] = S[j00]; ps[it][4 * c + 1] = S[j10];
                        ps[it][4 * c + 2] = S[j01]; ps[it][4 * c + 3] = S[j11];
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) s_w[j][e] = w[j];
            s_tap[e] = (unsigned char)code;
            s_l0[e] = in ? pl0[it] : 0.0f;
            s_l1[e] = in ? pl1[it] : 0.0f;
        },
        [&](int it, int e, int px, int py) {
            const size_t i = at(px, py, nx);
            // sediment: gather from the near cells whose taps include p.  A
            // neighbour q = p + (dx, dy) has p as a tap column when
            // ix - qx = -dx (or ix1 - qx = -dx), and likewise for rows.  A
            // near cell's code is never 0: that needs ix1 = ix = qx - 1, and
            // ix1 = ix only at ix = nx - 1, i.e. qx = nx.
            float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    const int eq = e + dx + dy * GRW;
                    const unsigned code = s_tap[eq];
                    if (code == 0) continue;
                    const int cx = 1 - dx, cy = 1 - dy;     // p - q + 1
                    const bool x0 = int(code & 3u) == cx, x1 = int((code >> 2) & 3u) == cx;
                    const bool y0 = int((code >> 4) & 3u) == cy, y1 = int((code >> 6) & 3u) == cy;
                    float w = 0.0f;
                    if (y0) {
                        if (x0) w += s_w[0][eq];
                        if (x1) w += s_w[1][eq];
                    }
                    if (y1) {
                        if (x0) w += s_w[2][eq];
                        if (x1) w += s_w[3][eq];
                    }
                    a0 += s_l0[eq] * w;
                    a1 += s_l1[eq] * w;
                }
            li.p[HG_P_SED_ROCK][i] = a0;
            li.p[HG_P_SED_DIRT][i] = a1;
            // own velocity and water
            const float vx = pvx[it], vy = pvy[it];
            const float bxr = float(px) - vx * k.dt;
            const float byr = float(py) - vy * k.dt;
            const Taps t = taps_of(px, py, vx, vy, nx, ny, k.dt);
            float bx, by;
            back_trace(px, py, vx, vy, k.dt, nx, ny, bx, by);
            const float fx = bx - float(t.ix), fy = by - float(t.iy);
            const float l0 = pl0[it], l1 = pl1[it];
            float lbx = 0.0f, lby = 0.0f;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const float s00 = ps[it][4 * c + 0], s10 = ps[it][4 * c + 1], s01 = ps[it][4 * c + 2],
                            s11 = ps[it][4 * c + 3];
                const float l = c == 0 ? l0 : l1;
                lbx += l * ((1.0f - fy) * (s10 - s00) + fy * (s11 - s01));
                lby += l * (lerp1(s01, s11, fx) - lerp1(s00, s10, fx));
            }
            // first writer of the velocity adjoint in a step (velocity is an
            // intermediate of the step: nothing after this step reads it)
            li.p[HG_P_VEL_U][i] = (bxr > 0.0f && bxr < float(nx - 1)) ? -k.dt * lbx : 0.0f;
            li.p[HG_P_VEL_V][i] = (byr > 0.0f && byr < float(ny - 1)) ? -k.dt * lby : 0.0f;
            li.p[HG_P_VEL_DEPTH][i] = 0.0f;
            const float lb = plb[it];
            li.p[HG_P_WATER][i] = lb * k.evap;                   // b' = b (1 - Ke dt)
            gKe += double(lb) * double(pb[it]) * double(-k.dt);
        });
    reduce_coef<GNT>(gpart, T.slot, 0.0, 0.0, 0.0, gKe);
}

__global__ __launch_bounds__(GNT) void k_adj_transport(Planes x, Planes lo, Planes li, int nx, int ny, KParams k,
                                                       double* gpart) {
    __shared__ float lds[TRANSPORT_LDS];
    adj_transport_tile(block_tile(), lds, x, lo, li, nx, ny, k, gpart);
}

// The transport VJP's scatter for cell (px, py) when its back-trace taps are
// not all within one cell (after the near gather has written li).
__device__ __forceinline__ void transport_far_cell(const Planes& x, const Planes& lo, const Planes& li, int px,
                                                   int py, int nx, int ny, const KParams& k) {
    const size_t i = at(px, py, nx);
    const Taps t = taps_of(px, py, x.p[HG_P_VEL_U][i], x.p[HG_P_VEL_V][i], nx, ny, k.dt);
    if (t.near) return;
    const float l0 = lo.p[HG_P_SED_ROCK][i], l1 = lo.p[HG_P_SED_DIRT][i];
    const size_t j00 = at(t.ix, t.iy, nx), j10 = at(t.ix1, t.iy, nx), j01 = at(t.ix, t.iy1, nx),
                 j11 = at(t.ix1, t.iy1, nx);
    aadd(li.p[HG_P_SED_ROCK] + j00, l0 * t.w00);
    aadd(li.p[HG_P_SED_ROCK] + j10, l0 * t.w10);
    aadd(li.p[HG_P_SED_ROCK] + j01, l0 * t.w01);
    aadd(li.p[HG_P_SED_ROCK] + j11, l0 * t.w11);
    aadd(li.p[HG_P_SED_DIRT] + j00, l1 * t.w00);
    aadd(li.p[HG_P_SED_DIRT] + j10, l1 * t.w10);
    aadd(li.p[HG_P_SED_DIRT] + j01, l1 * t.w01);
    aadd(li.p[HG_P_SED_DIRT] + j11, l1 * t.w11);
}

__global__ __launch_bounds__(256) void k_adj_transport_far(Planes x, Planes lo, Planes li, int nx, int ny,
                                                           KParams k) {
    const int px = blockIdx.x * BX + threadIdx.x, py = blockIdx.y * BY + threadIdx.y;
    if (px >= nx || py >= ny) return;
    transport_far_cell(x, lo, li, px, py, nx, ny, k);
}


// ---- thermal layer `layer`: thermal_erosion then thermal_transport --------
// Forward: T_k = outflow(h_in); h_out(p) = h_in(p) + sum_k (T_opp(k)(p+k) - T_k(p)).
// Adjoint: dL/dT_k(p) = lh(p+k) - lh(p) (lh = dL/dh_out of the layer), then the
// outflow's vector-Jacobian product sends -dL/dD_k to neighbour p+k and the
// sum to p (both rock and dirt for layer 1, whose D_k spans both layers).
//
// thermal_vjp: one in-grid cell p.  Dk: p's height differences to its 8
// neighbours (thermal_g's order), h_layer: p's height of the layer, lh and
// lhn: the layer's adjoint at p and at its neighbours, ok: whether each
// neighbour is on the grid.  Gives lD (what p sends to each neighbour, 0 off
// the grid), lself (p's own part) and lh_layer (through H = min(h_layer, .)).
template <int layer>
__device__ __forceinline__ void thermal_vjp(const float Dk[8], float h_layer, float lh, const float lhn[8],
                                            const bool ok[8], const KParams& k, float lD[8], float& lself,
                                            float& lh_layer) {
    ThermalTape tp;
    const float g = thermal_g_t(Dk, h_layer, layer, k, tp);
    if (tp.bk > 0.0f) {
        float lg = 0.0f;
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            const bool m = talus_marked(Dk[kk], kk, layer, k);
            const float lout = lhn[kk] - lh;          // dL/dT_k
            lD[kk] = m ? lout * g : 0.0f;
            if (m) lg += lout * Dk[kk];
        }
        const float rb = rcp(tp.bk);              // g = S / bk
        const float lS = lg * rb, lbk = -lg * g * rb;
#pragma unroll
        for (int kk = 0; kk < 8; ++kk)
            if (talus_marked(Dk[kk], kk, layer, k)) lD[kk] += lbk;
        const float A = k.dt * k.Kspeed[layer] * 0.5f;  // S = A sharp^3 H
        const float sh3 = tp.sharp * (tp.sharp * tp.sharp);
        const float lsh3 = lS * A * tp.H, lH = lS * A * sh3;
        if (tp.sharp_active) {                    // sharp = 1 + atan(rmax) - Kalpha
            const float lrmax = lsh3 * 3.0f * tp.sharp * tp.sharp * rcp(1.0f + tp.rmax * tp.rmax);
            if (tp.mc >= tp.md * 0.70710678118654752f) lD[tp.kc] += lrmax;
            else lD[tp.kd] += lrmax * 0.70710678118654752f;
        }
        if (tp.h_is_layer) lh_layer = lH;         // H = min(h_l, max(0, mc, md))
        else if (fmaxf(tp.mc, tp.md) > 0.0f) lD[tp.mc >= tp.md ? tp.kc : tp.kd] += lH;
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            lself += lD[kk];
            if (!ok[kk]) lD[kk] = 0.0f;           // nothing is sent off the grid
        }
    }
}

// LDS of one thermal tile, in floats: the staged inputs, then (once every
// cell is evaluated) what the cells send their 8 neighbours.
constexpr int THERMAL_LDS = 8 * GNR > 3 * SRN ? 8 * GNR : 3 * SRN;

// A thermal layer's VJP on tile T, with `lds` (THERMAL_LDS floats) as its
// LDS.  x: the layer's input rock and dirt.  Writes li's layer plane (fresh)
// and, for layer 1, adds the rock part to li's rock in place.
template <int layer>
__device__ __forceinline__ void adj_thermal_tile(const TileId& T, float* lds, Planes x, Planes lo, Planes li,
                                                 int nx, int ny, const KParams& k) {
    float* t_r = lds;
    float* t_d = lds + SRN;
    float* t_lh = lds + 2 * SRN;
    float (*s_lD)[GNR] = reinterpret_cast<float (*)[GNR]>(lds);   // sent to the 8 neighbours
    float o_self[GTIT], o_hl[GTIT], o_lh[GTIT];     // kept per thread
    float snd[GIT][8];
    constexpr int DX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
    constexpr int DY[8] = {0, 0, 1, -1, 1, 1, -1, -1};
    const float* R = x.p[HG_P_ROCK];
    const float* Dt = x.p[HG_P_DIRT];
    const float* LH = lo.p[layer == 0 ? HG_P_ROCK : HG_P_DIRT];
    {
        // rock, dirt (out of grid: HG_OOB_HEIGHT) and the layer's adjoint (0)
        // (layer 0 reads no dirt)
        float a[SIT], b[SIT], c[SIT];
        stage_walk(T, nx, ny, [&](int it, int, size_t i, bool in) {
            const float r = R[i], l = LH[i];
            a[it] = in ? r : HG_OOB_HEIGHT;
            if constexpr (layer == 1) {
                const float d = Dt[i];
                b[it] = in ? d : HG_OOB_HEIGHT;
            }
            c[it] = in ? l : 0.0f;
        });
        stage_walk(T, nx, ny, [&](int it, int s, size_t, bool) {
            t_r[s] = a[it];
            if constexpr (layer == 1) t_d[s] = b[it];
            t_lh[s] = c[it];
        });
    }
    tile_gather_late(T, nx, ny,
        [&](int, int, size_t) {},
        [&](int it, int e, int px, int py, bool in, bool) {
            float lD[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            float lself = 0.0f, lh_layer = 0.0f;
            const int sc = sidx(e);
            if (it < GTIT) o_lh[it] = t_lh[sc];
            if (in) {
                const float r = t_r[sc], d = layer == 1 ? t_d[sc] : 0.0f;
                float Dk[8], lhn[8];
                bool ok[8];
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    const int xx = px + DX[kk], yy = py + DY[kk];
                    const int o = sc + DX[kk] + DY[kk] * SRW;
                    ok[kk] = xx >= 0 && yy >= 0 && xx < nx && yy < ny;
                    float v = r - t_r[o];
                    if constexpr (layer == 1) v = v + (d - t_d[o]);
                    Dk[kk] = v;
                    lhn[kk] = t_lh[o];
                }
                thermal_vjp<layer>(Dk, layer == 0 ? r : d, t_lh[sc], lhn, ok, k, lD, lself, lh_layer);
            }
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) snd[it][kk] = lD[kk];
            if (it < GTIT) {                                  // a tile cell keeps its own part
                o_self[it] = lself;
                o_hl[it] = lh_layer;
            }
        },
        [&](int it, int e) {
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) s_lD[kk][e] = snd[it][kk];
        },
        [&](int it, int e, int px, int py) {
            const size_t i = at(px, py, nx);
            float recv = 0.0f;                                // from q = p - dir_k: -dL/dD_k(q)
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) recv -= s_lD[kk][e - (DX[kk] + DY[kk] * GRW)];
            const float shared = o_self[it] + recv;
            const float lh = o_lh[it];
            if constexpr (layer == 0) {
                li.p[HG_P_ROCK][i] = lh + shared + o_hl[it];
            } else {
                li.p[HG_P_DIRT][i] = lh + shared + o_hl[it];
                li.p[HG_P_ROCK][i] += shared;
            }
        });
}

template <int layer>
__global__ __launch_bounds__(GNT) void k_adj_thermal(Planes x, Planes lo, Planes li, int nx, int ny, KParams k) {
    __shared__ float lds[THERMAL_LDS];
    adj_thermal_tile<layer>(block_tile(), lds, x, lo, li, nx, ny, k);
}

// ---- thermal layer 0 and the transport's far taps in one launch ----------
// Run after the transport's near gather (the two passes' adjoints touch
// disjoint planes, so their order does not matter): workgroups [0, tiles)
// take the layer-0 VJP's tiles in row-major order (as k_adj_thermal<0>'s
// grid does), the rest scatter the far taps of 256 cells each (as
// k_adj_transport_far), filling the thermal tiles' tail instead of a launch
// of their own.  x0: layer 0's inputs; xt, lot, lit: the transport's.
__global__ __launch_bounds__(GNT) void k_adj_thermal0_far(Planes x0, Planes lo0, Planes li0, Planes xt, Planes lot,
                                                          Planes lit, int nx, int ny, KParams k) {
    __shared__ float lds[THERMAL_LDS];
    const int ntx = (nx + GTX - 1) / GTX, nty = (ny + GTY - 1) / GTY;
    const int w = int(blockIdx.x);
    if (w < ntx * nty) {
        adj_thermal_tile<0>(tile_at(w, ntx * nty, ntx), lds, x0, lo0, li0, nx, ny, k);
        return;
    }
    const size_t cell = size_t(w - ntx * nty) * GNT + threadIdx.x;
    if (cell >= size_t(nx) * size_t(ny)) return;
    transport_far_cell(xt, lot, lit, int(cell % size_t(nx)), int(cell / size_t(nx)), nx, ny, k);
}

// ---- smoothing -----------------------------------------------------------
// y = x + m (mean5 - x) at an extremum.  Gather: p keeps its own share and
// receives 0.2 m dL/dy(q) from each interior neighbour q that is an extremum.
// The extremum flags of the tile + 1-cell ring are evaluated once into LDS.
// Writes li's rock and dirt (fresh).
__global__ __launch_bounds__(GNT) void k_adj_smooth(Planes x, Planes lo, Planes li, int nx, int ny, KParams k) {
    const TileId T = block_tile();
    // rock, dirt and their adjoints on the tile + 2-cell region; cells off the
    // grid hold the clamped cell's values, which only feed extremum flags of
    // non-interior cells, and those are not used
    __shared__ float t_r[SRN], t_d[SRN], t_lr[SRN], t_ld[SRN];
    // per tile/ring cell: bit 0 rock, bit 1 surface flag of an interior extremum
    __shared__ unsigned char s_f[GNR];
    {
        float a[SIT], b[SIT], c[SIT], d[SIT];
        stage_walk(T, nx, ny, [&](int it, int, size_t i, bool) {
            a[it] = x.p[HG_P_ROCK][i]; b[it] = x.p[HG_P_DIRT][i];
            c[it] = lo.p[HG_P_ROCK][i]; d[it] = lo.p[HG_P_DIRT][i];
        });
        stage_walk(T, nx, ny, [&](int it, int s, size_t, bool) {
            t_r[s] = a[it]; t_d[s] = b[it]; t_lr[s] = c[it]; t_ld[s] = d[it];
        });
    }
    const float m = k.smooth_m;
    float o_lr[GTIT], o_ld[GTIT];
    tile_gather(T, nx, ny,
        [&](int, int, size_t) {},
        [&](int it, int e, int px, int py, bool, bool) {
            const int sc = sidx(e);
            unsigned f = 0;
            if (px > 0 && py > 0 && px < nx - 1 && py < ny - 1) {
                bool er, eg;
                smooth_flags(t_r[sc], t_d[sc], t_r[sc - 1], t_d[sc - 1], t_r[sc + 1], t_d[sc + 1], t_r[sc + SRW],
                             t_d[sc + SRW], t_r[sc - SRW], t_d[sc - SRW], er, eg);
                f = (er ? 1u : 0u) | (eg ? 2u : 0u);
            }
            s_f[e] = (unsigned char)f;
            if (it < GTIT) {                                  // a tile cell keeps its own share
                const float lr = t_lr[sc], ld = t_ld[sc];
                o_lr[it] = (f & 1u) ? lr * (1.0f - 0.8f * m) : lr;
                o_ld[it] = (f & 2u) ? ld * (1.0f - 0.8f * m) : ld;
            }
        },
        [&](int it, int e, int px, int py) {
            const int sp = sidx(e);
            float ar = o_lr[it], ad = o_ld[it];
            constexpr int NXo[4] = {-1, 1, 0, 0}, NYo[4] = {0, 0, 1, -1};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const unsigned f = s_f[e + NXo[j] + NYo[j] * GRW];
                const int sq = sp + NXo[j] + NYo[j] * SRW;
                if (f & 1u) ar += t_lr[sq] * 0.2f * m;
                if (f & 2u) ad += t_ld[sq] * 0.2f * m;
            }
            const size_t i = at(px, py, nx);
            li.p[HG_P_ROCK][i] = ar;
            li.p[HG_P_DIRT][i] = ad;
        });
}

// ---- smoothing and thermal layer 1 in one launch --------------------------
// k_adj_smooth followed by k_adj_thermal<1>, with the same arithmetic in the
// same order.  The layer-1 VJP of a ring cell reads the dirt adjoint of its
// neighbours, so the smoothing VJP is evaluated on the tile + 2-cell region
// (into LDS), from extremum flags on the tile + 3 region and rock and dirt
// staged on the tile + 4 region; the layer-1 VJP then runs as in
// k_adj_thermal<1>.  x: rock after layer 0 (x.p[HG_P_ROCK]) and dirt after
// layer 1 (x.p[HG_P_DIRT]), the smoothing's inputs; d3: dirt after erosion,
// layer 1's input dirt.  lo: the adjoints after smoothing.  Writes li's rock
// and dirt (fresh).

// LDS of one smoothing + layer-1 tile, in floats (see adj_smooth_thermal1_tile).
constexpr int ST1_N4 = (GTX + 8) * (GTY + 8), ST1_N3 = (GTX + 6) * (GTY + 6);
constexpr int ST1_LDS = 2 * ST1_N4 + 2 * ST1_N3 + (ST1_N3 + 3) / 4 > 8 * GNR
                            ? 2 * ST1_N4 + 2 * ST1_N3 + (ST1_N3 + 3) / 4
                            : 8 * GNR;

__device__ __forceinline__ void adj_smooth_thermal1_tile(const TileId& T, float* u, Planes x, Planes lo,
                                                         Planes li, const float* d3, int nx, int ny,
                                                         const KParams& k) {
    constexpr int W4 = GTX + 8, N4 = ST1_N4, IT4 = (N4 + GNT - 1) / GNT;
    constexpr int W3 = GTX + 6, N3 = ST1_N3, IT3 = (N3 + GNT - 1) / GNT;
    // One LDS array (21.1 KB), in turn:
    //  * rock after layer 0 (t_r5) and dirt after layer 1 (t_d6) on tile + 4,
    //    HG_OOB_HEIGHT off the grid (the flags of cells off or on the edge of
    //    the grid are not evaluated, and layer 1 reads the rock as
    //    k_adj_thermal<1> stages it); the adjoints after smoothing on tile + 3
    //    (0 off the grid) and their extremum flags;
    //  * once the flags are set, the dirt adjoint before smoothing on tile + 2
    //    (t_md, 0 off the grid) in t_d6's place; once the smoothing VJP is
    //    done, dirt after erosion on tile + 2 (t_d) in the adjoints' place;
    //  * once every cell's layer-1 VJP is evaluated, what the cells send
    //    their 8 neighbours.
    static_assert(SRN <= N4 && SRN <= 2 * N3, "the tile + 2 arrays fit the places they take");
    float* t_r5 = u;
    float* t_d6 = u + N4;
    float* t_lr = u + 2 * N4;
    float* t_ld = t_lr + N3;
    unsigned char* s_f = reinterpret_cast<unsigned char*>(t_ld + N3);
    float* t_md = t_d6;
    float* t_d = t_lr;
    float (*s_lD)[GNR] = reinterpret_cast<float (*)[GNR]>(u);
    float o_mr[GTIT], o_lh[GTIT], o_self[GTIT], o_hl[GTIT];
    float snd[GIT][8];
    constexpr int DX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
    constexpr int DY[8] = {0, 0, 1, -1, 1, 1, -1, -1};
    const float m = k.smooth_m;
    float e[SIT];                                // dirt after erosion, tile + 2, until its place is free
    {
        float a[IT4], b[IT4], c[IT3], d[IT3];
        region_walk<4>(T, nx, ny, [&](int it, int, size_t i, bool in, int, int) {
            const float r = x.p[HG_P_ROCK][i], dd = x.p[HG_P_DIRT][i];
            a[it] = in ? r : HG_OOB_HEIGHT;
            b[it] = in ? dd : HG_OOB_HEIGHT;
        });
        region_walk<3>(T, nx, ny, [&](int it, int, size_t i, bool in, int, int) {
            const float lr = lo.p[HG_P_ROCK][i], ld = lo.p[HG_P_DIRT][i];
            c[it] = in ? lr : 0.0f;
            d[it] = in ? ld : 0.0f;
        });
        region_walk<2>(T, nx, ny, [&](int it, int, size_t i, bool in, int, int) {
            const float v = d3[i];
            e[it] = in ? v : HG_OOB_HEIGHT;
        });
        region_walk<4>(T, nx, ny, [&](int it, int s, size_t, bool, int, int) { t_r5[s] = a[it]; t_d6[s] = b[it]; });
        region_walk<3>(T, nx, ny, [&](int it, int s, size_t, bool, int, int) { t_lr[s] = c[it]; t_ld[s] = d[it]; });
    }
    __syncthreads();
    // extremum flags on tile + 3, as k_adj_smooth evaluates them
    region_walk<3>(T, nx, ny, [&](int, int s, size_t, bool, int gx, int gy) {
        const int sy = s / W3, sx = s - sy * W3;
        const int c = (sy + 1) * W4 + sx + 1;
        unsigned f = 0;
        if (gx > 0 && gy > 0 && gx < nx - 1 && gy < ny - 1) {
            bool er, eg;
            smooth_flags(t_r5[c], t_d6[c], t_r5[c - 1], t_d6[c - 1], t_r5[c + 1], t_d6[c + 1], t_r5[c + W4],
                         t_d6[c + W4], t_r5[c - W4], t_d6[c - W4], er, eg);
            f = (er ? 1u : 0u) | (eg ? 2u : 0u);
        }
        s_f[s] = (unsigned char)f;
    });
    __syncthreads();
    // the smoothing VJP: its own share, then what each extremum neighbour
    // passes back, in k_adj_smooth's order; c is a cell's index on tile + 3
    constexpr int NXo[4] = {-1, 1, 0, 0}, NYo[4] = {0, 0, 1, -1};
    auto smooth_r = [&](int c) {
        const unsigned f = s_f[c];
        const float lr = t_lr[c];
        float ar = (f & 1u) ? lr * (1.0f - 0.8f * m) : lr;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = c + NXo[j] + NYo[j] * W3;
            if (s_f[q] & 1u) ar += t_lr[q] * 0.2f * m;
        }
        return ar;
    };
    auto smooth_d = [&](int c) {
        const unsigned f = s_f[c];
        const float ld = t_ld[c];
        float ad = (f & 2u) ? ld * (1.0f - 0.8f * m) : ld;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = c + NXo[j] + NYo[j] * W3;
            if (s_f[q] & 2u) ad += t_ld[q] * 0.2f * m;
        }
        return ad;
    };
    // the dirt adjoint on tile + 2 (layer 1's input adjoint)
    region_walk<2>(T, nx, ny, [&](int, int s, size_t, bool in, int, int) {
        const int sy = s / SRW, sx = s - sy * SRW;
        t_md[s] = in ? smooth_d((sy + 1) * W3 + sx + 1) : 0.0f;
    });
    // the rock adjoint of each tile cell, kept by the thread that writes it
#pragma unroll
    for (int it = 0; it < GTIT; ++it) {
        int ex, ey;
        ring_cell(it, ex, ey);
        o_mr[it] = smooth_r((ey + 2) * W3 + ex + 2);
    }
    __syncthreads();                             // the adjoints after smoothing are no longer read
    region_walk<2>(T, nx, ny, [&](int it, int s, size_t, bool, int, int) { t_d[s] = e[it]; });
    // layer 1, as k_adj_thermal<1> with t_lh = t_md (tile_gather_late's first
    // barrier completes t_d)
    tile_gather_late(T, nx, ny,
        [&](int, int, size_t) {},
        [&](int it, int e, int px, int py, bool in, bool) {
            float lD[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            float lself = 0.0f, lh_layer = 0.0f;
            const int sc = sidx(e);
            if (it < GTIT) o_lh[it] = t_md[sc];
            if (in) {
                const int sy = sc / SRW, sx = sc - sy * SRW;
                const int c4 = (sy + 2) * W4 + sx + 2;
                const float r = t_r5[c4], d = t_d[sc];
                float Dk[8], lhn[8];
                bool ok[8];
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    const int xx = px + DX[kk], yy = py + DY[kk];
                    const int o = sc + DX[kk] + DY[kk] * SRW;
                    ok[kk] = xx >= 0 && yy >= 0 && xx < nx && yy < ny;
                    Dk[kk] = (r - t_r5[c4 + DX[kk] + DY[kk] * W4]) + (d - t_d[o]);
                    lhn[kk] = t_md[o];
                }
                thermal_vjp<1>(Dk, d, t_md[sc], lhn, ok, k, lD, lself, lh_layer);
            }
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) snd[it][kk] = lD[kk];
            if (it < GTIT) {
                o_self[it] = lself;
                o_hl[it] = lh_layer;
            }
        },
        [&](int it, int e) {
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) s_lD[kk][e] = snd[it][kk];
        },
        [&](int it, int e, int px, int py) {
            const size_t i = at(px, py, nx);
            float recv = 0.0f;
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) recv -= s_lD[kk][e - (DX[kk] + DY[kk] * GRW)];
            const float shared = o_self[it] + recv;
            li.p[HG_P_DIRT][i] = o_lh[it] + shared + o_hl[it];
            li.p[HG_P_ROCK][i] = o_mr[it] + shared;
        });
}

__global__ __launch_bounds__(GNT) void k_adj_smooth_thermal1(Planes x, Planes lo, Planes li, const float* d3,
                                                             int nx, int ny, KParams k) {
    __shared__ float u[ST1_LDS];
    adj_smooth_thermal1_tile(block_tile(), u, x, lo, li, d3, nx, ny, k);
}

// ---- smoothing + thermal layer 1 and the transport's near gather ----------
// The two touch disjoint planes (rock and dirt; sediment, water and
// velocity), so one launch runs both: workgroups [0, tiles) take the
// smoothing + layer-1 tiles in row-major order, the next `tiles` workgroups
// the transport's tiles (each on the XCD its own kernel would give it, with
// the tile's coefficient slot), starting as the first ones finish.  xs, los,
// lis: the smoothing + layer-1 planes; xt, lot, lit: the transport's.
__global__ __launch_bounds__(GNT) void k_adj_smooth_thermal1_transport(Planes xs, Planes los, Planes lis,
                                                                       const float* d3, Planes xt, Planes lot,
                                                                       Planes lit, int nx, int ny, KParams k,
                                                                       double* gpart) {
    constexpr int NL = ST1_LDS > TRANSPORT_LDS ? ST1_LDS : TRANSPORT_LDS;
    __shared__ float u[NL];
    const int ntx = (nx + GTX - 1) / GTX, nty = (ny + GTY - 1) / GTY, nt = ntx * nty;
    const int w = int(blockIdx.x);
    if (w < nt) {
        adj_smooth_thermal1_tile(tile_at(w, nt, ntx), u, xs, los, lis, d3, nx, ny, k);
    } else {
        adj_transport_tile(tile_at(w - nt, nt, ntx), u, xt, lot, lit, nx, ny, k, gpart);
    }
}

// Copies the 9 state planes of n floats (n = side^2, side % 8 == 0) in one launch.
struct StateCopy {
    const float* s[HG_NUM_STATE_PLANES];
    float* d[HG_NUM_STATE_PLANES];
};
__global__ __launch_bounds__(256) void k_copy_state(StateCopy l, size_t n4) {
    const size_t total = size_t(HG_NUM_STATE_PLANES) * n4;
    for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < total; e += size_t(gridDim.x) * 256) {
        const size_t pl = e / n4, j = e - pl * n4;
        reinterpret_cast<float4*>(l.d[pl])[j] = reinterpret_cast<const float4*>(l.s[pl])[j];
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

// Grouped reverse units (hg_set_adjoint_fusion): the erosion and flux VJPs
// (k_adj_erosion_flux), the smoothing and thermal layer-1 VJPs
// (k_adj_smooth_thermal1), the transport's near gather then thermal layer 0
// with the transport's far taps (k_adj_transport, k_adj_thermal0_far).
constexpr int UNIT_EROSION_FLUX = 1000;
constexpr int UNIT_SMOOTH_THERMAL1 = 1001;
constexpr int UNIT_TRANSPORT_THERMAL0 = 1002;
// With all four fusions: the smoothing + layer-1 and transport tiles in one
// launch (k_adj_smooth_thermal1_transport), then layer 0 and the far taps.
constexpr int UNIT_ST1_TRANSPORT = 1003;
constexpr int UNIT_THERMAL0_FAR = 1004;

// Planes each reverse unit produces: their adjoint before the unit is written
// in full into a fresh buffer (the thermal units stand for outflow + transport
// of a layer; the velocity planes of the flux units are left unwritten, since
// the transport VJP of the step before writes them first).
std::vector<int> produced(int pass) {
    switch (pass) {
        case HG_PASS_RAIN: return {HG_P_WATER};
        case HG_PASS_FLUX: return {HG_P_WATER, HG_P_FLUX_L, HG_P_FLUX_R, HG_P_FLUX_T, HG_P_FLUX_B,
                                   HG_P_VEL_U, HG_P_VEL_V, HG_P_VEL_DEPTH};
        case HG_PASS_EROSION: return {HG_P_ROCK, HG_P_DIRT, HG_P_SED_ROCK, HG_P_SED_DIRT};
        case HG_PASS_TRANSPORT: return {HG_P_SED_ROCK, HG_P_SED_DIRT, HG_P_WATER};
        case HG_PASS_THERMAL_TRANSPORT_0: return {HG_P_ROCK};
        case HG_PASS_THERMAL_TRANSPORT_1: return {HG_P_DIRT};
        case HG_PASS_SMOOTH: return {HG_P_ROCK, HG_P_DIRT};
        case UNIT_SMOOTH_THERMAL1: return {HG_P_ROCK, HG_P_DIRT};
        case UNIT_TRANSPORT_THERMAL0: return {HG_P_SED_ROCK, HG_P_SED_DIRT, HG_P_WATER, HG_P_ROCK};
        case UNIT_ST1_TRANSPORT: return {HG_P_ROCK, HG_P_DIRT, HG_P_SED_ROCK, HG_P_SED_DIRT, HG_P_WATER};
        case UNIT_THERMAL0_FAR: return {HG_P_ROCK};
        case UNIT_EROSION_FLUX: return {HG_P_ROCK, HG_P_DIRT, HG_P_SED_ROCK, HG_P_SED_DIRT, HG_P_WATER,
                                        HG_P_FLUX_L, HG_P_FLUX_R, HG_P_FLUX_T, HG_P_FLUX_B,
                                        HG_P_VEL_U, HG_P_VEL_V, HG_P_VEL_DEPTH};
        default: return {};
    }
}

// Device buffers of the reverse sweep.
struct AdjWork {
    size_t pn = 0, pitch = 0;
    void* block = nullptr;
    std::vector<float*> ckpt;      // (#checkpoints) x 9 planes
    // Two sets (one for a single segment), so that one segment is recomputed
    // while the previous one is reversed: (k + 1) x 9 planes (k the longest
    // segment), each step's start state in a segment
    // (slot 0 is the checkpoint itself), and k_ckpt x NTAPE planes, each
    // step's intermediates.
    std::vector<float*> seg[2];
    std::vector<float*> tape[2];
    std::vector<float*> pool;      // adjoint planes (free list)
    float* zero = nullptr;
    double* gpart = nullptr;       // 4 coefficient partial sums per gather workgroup
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
              bool thermal, bool smooth, hipStream_t st) {
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
    HG_HIP(c, hgk::launch_fused_hydro(c->hydro_variant, io, c->n, c->n, step, 0u, rain, c->xcd_order, c->kp, st));
    if (!th) return HG_OK;
    hgk::ThermIO tio;
    tio.r = tp[T_R3]; tio.d = tp[T_D3];
    tio.ro = out[HG_P_ROCK]; tio.dout = out[HG_P_DIRT];
    tio.tr5 = tp[T_R5]; tio.td6 = tp[T_D6];
    HG_HIP(c, hgk::launch_fused_thermal(c->thermal_variant, tio, c->n, c->n, thermal, smooth, c->xcd_order, c->kp,
                                        st));
    return HG_OK;
}

int copy_state(hg_ctx* c, float* const* src, float* const* dst, size_t pn) {
    hgk::StateCopy l;
    for (int p = 0; p < HG_NUM_STATE_PLANES; ++p) { l.s[p] = src[p]; l.d[p] = dst[p]; }
    const size_t n4 = pn / 4;
    const unsigned grid = unsigned(std::min<size_t>((HG_NUM_STATE_PLANES * n4 + 255) / 256, 8192));
    hgk::k_copy_state<<<grid, 256, 0, c->stream>>>(l, n4);
    HG_HIP(c, hipGetLastError());
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
    if ((flags & HG_STEP_RAIN) && c->rain_mode == HG_RAIN_REFERENCE)
        return fail(c, HG_ERR_UNSUPPORTED, "hg_adjoint_rollout: reference rain is not differentiated");
    DeviceGuard guard(c->device);
    const bool rain = (flags & HG_STEP_RAIN) && c->rain_mode != HG_RAIN_OFF;
    const bool thermal = (flags & HG_STEP_THERMAL) != 0, smooth = (flags & HG_STEP_SMOOTH) != 0;
    const int n = c->n;
    const size_t pn = plane_n(c);
    const uint32_t nck = (nsteps + k_ckpt - 1) / k_ckpt;
    const uint32_t kseg = std::min(k_ckpt, nsteps);        // longest segment
    const int nsets = nck > 1 ? 2 : 1;
    AdjWork w;
    w.pn = pn;
    w.pitch = hgk::plane_stride_bytes(pn);
    const size_t n_pool = 64;
    const size_t nplanes = size_t(nck) * HG_NUM_STATE_PLANES +
                           nsets * (size_t(kseg + 1) * HG_NUM_STATE_PLANES + size_t(kseg) * NTAPE) + n_pool + 2;
    // The workspace stays with the context for the next call (a rollout of
    // the same shape allocates nothing); a larger one replaces it.
    const size_t need = nplanes * w.pitch + 256;
    if (c->adj_bytes < need) {
        if (c->adj_block) {
            HG_HIP(c, hipFree(c->adj_block));
            c->adj_block = nullptr;
            c->adj_bytes = 0;
        }
        void* p = nullptr;
        HG_HIP(c, hipMalloc(&p, need));
        c->adj_block = p;
        c->adj_bytes = need;
    }
    w.block = c->adj_block;
    char* base = static_cast<char*>(w.block);
    size_t used = 0;
    auto plane = [&]() { float* p = reinterpret_cast<float*>(base + used * w.pitch); ++used; return p; };
    for (size_t i = 0; i < size_t(nck) * HG_NUM_STATE_PLANES; ++i) w.ckpt.push_back(plane());
    for (int b = 0; b < nsets; ++b) {
        for (size_t i = 0; i < size_t(kseg + 1) * HG_NUM_STATE_PLANES; ++i) w.seg[b].push_back(plane());
        for (size_t i = 0; i < size_t(kseg) * NTAPE; ++i) w.tape[b].push_back(plane());
    }
    for (size_t i = 0; i < n_pool; ++i) w.pool.push_back(plane());
    w.zero = plane();
    // one slot set per gather workgroup: 32 B per 64 x 8 cells, well within a plane
    const size_t nblk = size_t((c->n + hgk::GTX - 1) / hgk::GTX) * size_t((c->n + hgk::GTY - 1) / hgk::GTY);
    w.gpart = reinterpret_cast<double*>(base + used * w.pitch);
    HG_HIP(c, hipMemsetAsync(w.zero, 0, pn * sizeof(float), c->stream));
    HG_HIP(c, hipMemsetAsync(w.gpart, 0, nblk * 4 * sizeof(double), c->stream));
    const uint64_t step0 = c->step;

    // ---- forward with checkpoints every k_ckpt steps.  Segment `seg` is
    // reversed from plane set (nck - 1 - seg) % 2.  The last two segments (the
    // only one, if there is one) run through the taped kernels straight into
    // the set their reversal reads, so they are never recomputed; the first of
    // them starts from its checkpoint, the second from the first's last slot.
    const uint32_t first_taped = nck - uint32_t(nsets);
    for (uint32_t s = 0; s < nsteps; ++s) {
        const uint32_t seg = s / k_ckpt, j = s - seg * k_ckpt;
        if (seg < first_taped) {
            if (j == 0) {
                const int rc = copy_state(c, c->cur, &w.ckpt[size_t(seg) * HG_NUM_STATE_PLANES], pn);
                if (rc != HG_OK) return rc;
            }
            const int rc = fused_step(c, rain, thermal, smooth, false);
            if (rc != HG_OK) return rc;
        } else {
            std::vector<float*>& slots = w.seg[(nck - 1 - seg) % 2];
            if (j == 0) {
                if (seg == first_taped) {
                    float* const* ck = &w.ckpt[size_t(seg) * HG_NUM_STATE_PLANES];
                    const int rc = copy_state(c, c->cur, ck, pn);
                    if (rc != HG_OK) return rc;
                    for (int p = 0; p < HG_NUM_STATE_PLANES; ++p) slots[size_t(p)] = ck[p];
                } else {
                    const std::vector<float*>& prev = w.seg[(nck - seg) % 2];
                    for (int p = 0; p < HG_NUM_STATE_PLANES; ++p)
                        slots[size_t(p)] = prev[size_t(k_ckpt) * HG_NUM_STATE_PLANES + p];
                }
            }
            const int rc = step_into(c, &slots[size_t(j) * HG_NUM_STATE_PLANES],
                                     &slots[size_t(j + 1) * HG_NUM_STATE_PLANES],
                                     &w.tape[(nck - 1 - seg) % 2][size_t(j) * NTAPE], uint32_t(step0 + s), rain,
                                     thermal, smooth, c->stream);
            if (rc != HG_OK) return rc;
        }
        c->step += 1;
    }
    {
        // the final state, in the last segment's last slot, back into the context
        const size_t last = nsteps - (nck - 1) * k_ckpt;
        const int rc = copy_state(c, &w.seg[0][last * HG_NUM_STATE_PLANES], c->cur, pn);
        if (rc != HG_OK) return rc;
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
    HG_HIP(c, hipMemcpyAsync(lam[HG_P_DIRT], lam[HG_P_ROCK], pn * sizeof(float), hipMemcpyDeviceToDevice, c->stream));

    // reverse units in forward order (they run from the last): a thermal
    // layer (outflow then transport) is one unit, named by its transport
    // pass.  The sediment transport and thermal layer 0 touch disjoint
    // planes, so the grouped units reverse the transport first and layer 0
    // after it.
    std::vector<int> order;
    if (c->adj_fuse & HG_ADJ_FUSE_EROSION_FLUX) {
        order.push_back(UNIT_EROSION_FLUX);
    } else {
        order.push_back(HG_PASS_FLUX);
        order.push_back(HG_PASS_EROSION);
    }
    const bool fuse_st1 = thermal && smooth && (c->adj_fuse & HG_ADJ_FUSE_SMOOTH_THERMAL1);
    const bool fuse_far0 = thermal && (c->adj_fuse & HG_ADJ_FUSE_FAR_THERMAL0);
    if (fuse_st1 && fuse_far0 && (c->adj_fuse & HG_ADJ_FUSE_TRANSPORT_BESIDE)) {
        order.push_back(UNIT_THERMAL0_FAR);
        order.push_back(UNIT_ST1_TRANSPORT);
    } else {
        if (fuse_far0) {
            order.push_back(UNIT_TRANSPORT_THERMAL0);
        } else {
            order.push_back(HG_PASS_TRANSPORT);
            if (thermal) order.push_back(HG_PASS_THERMAL_TRANSPORT_0);
        }
        if (fuse_st1) {
            order.push_back(UNIT_SMOOTH_THERMAL1);
        } else {
            if (thermal) order.push_back(HG_PASS_THERMAL_TRANSPORT_1);
            if (smooth) order.push_back(HG_PASS_SMOOTH);
        }
    }

    const dim3 blk(BX, BY), grd((n + BX - 1) / BX, (n + BY - 1) / BY);
    const dim3 gtile((n + GTX - 1) / GTX, (n + GTY - 1) / GTY);
    // The taped recompute of a segment runs on a second stream (default
    // priority) while the main stream reverses the segment after it: at 1024^2
    // neither kernel sequence fills the chip alone.  Segment i (counting from
    // the last) uses plane set i % 2; its recompute waits for the reversal that
    // last used that set, and its reversal waits for its recompute.
    if (!c->adj_stream) {
        hipStream_t st = nullptr;
        HG_HIP(c, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        c->adj_stream = st;
    }
    for (hipEvent_t& e : c->adj_ev)
        if (!e) {
            hipEvent_t ev = nullptr;
            HG_HIP(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            e = ev;
        }
    hipStream_t rs = c->adj_stream;
    hipEvent_t ev_fwd = c->adj_ev[0], ev_rec[2] = {c->adj_ev[1], c->adj_ev[2]},
               ev_rev[2] = {c->adj_ev[3], c->adj_ev[4]};
    struct Drain {    // on every return, nothing is left running on the recompute stream
        hipStream_t rs;
        ~Drain() { (void)hipStreamSynchronize(rs); }
    } drain{rs};
    HG_HIP(c, hipEventRecord(ev_fwd, c->stream));     // checkpoints and the zeroed adjoint are written
    HG_HIP(c, hipStreamWaitEvent(rs, ev_fwd, 0));
    // recompute segment `seg` from its checkpoint into plane set b, recording
    // every step's start state and intermediates
    auto recompute = [&](int seg, int b, bool wait_reversal) -> int {
        const uint32_t s_begin = uint32_t(seg) * k_ckpt;
        const uint32_t s_end = std::min(nsteps, s_begin + k_ckpt);
        if (wait_reversal) HG_HIP(c, hipStreamWaitEvent(rs, ev_rev[b], 0));
        std::vector<float*>& slots = w.seg[b];
        for (int p = 0; p < HG_NUM_STATE_PLANES; ++p) slots[size_t(p)] = w.ckpt[size_t(seg) * HG_NUM_STATE_PLANES + p];
        for (uint32_t s = s_begin; s < s_end; ++s) {
            const size_t j = s - s_begin;
            const int rc = step_into(c, &slots[j * HG_NUM_STATE_PLANES], &slots[(j + 1) * HG_NUM_STATE_PLANES],
                                     &w.tape[b][j * NTAPE], uint32_t(step0 + s), rain, thermal, smooth, rs);
            if (rc != HG_OK) return rc;
        }
        HG_HIP(c, hipEventRecord(ev_rec[b], rs));
        return HG_OK;
    };
    for (int seg = int(nck) - 1; seg >= 0; --seg) {
        const int i = int(nck) - 1 - seg, b = i % 2;
        const uint32_t s_begin = uint32_t(seg) * k_ckpt;
        const uint32_t s_end = std::min(nsteps, s_begin + k_ckpt);
        if (seg > 0 && uint32_t(seg - 1) < first_taped) {
            const int rc = recompute(seg - 1, 1 - b, i >= 1);
            if (rc != HG_OK) return rc;
        }
        if (uint32_t(seg) < first_taped) HG_HIP(c, hipStreamWaitEvent(c->stream, ev_rec[b], 0));
        std::vector<float*>& slots = w.seg[b];
        // reverse the steps of the segment
        for (int s = int(s_end) - 1; s >= int(s_begin); --s) {
            float* const* in = &slots[size_t(s - int(s_begin)) * HG_NUM_STATE_PLANES];
            float* const* tp = &w.tape[b][size_t(s - int(s_begin)) * NTAPE];
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
                    case UNIT_EROSION_FLUX:
                        X.p[HG_P_ROCK] = in[HG_P_ROCK]; X.p[HG_P_DIRT] = in[HG_P_DIRT]; X.p[HG_P_WATER] = tp[T_B1];
                        for (int p = HG_P_FLUX_L; p <= HG_P_FLUX_B; ++p) X.p[p] = in[p];
                        X.p[HG_P_SED_ROCK] = in[HG_P_SED_ROCK]; X.p[HG_P_SED_DIRT] = in[HG_P_SED_DIRT];
                        X.p[HG_P_VEL_U] = tp[T_VX]; X.p[HG_P_VEL_V] = tp[T_VY]; X.p[HG_P_VEL_DEPTH] = tp[T_VZ];
                        break;
                    case HG_PASS_TRANSPORT:
                        X.p[HG_P_SED_ROCK] = tp[T_S0]; X.p[HG_P_SED_DIRT] = tp[T_S1]; X.p[HG_P_WATER] = tp[T_B2];
                        X.p[HG_P_VEL_U] = tp[T_VX]; X.p[HG_P_VEL_V] = tp[T_VY]; X.p[HG_P_VEL_DEPTH] = tp[T_VZ];
                        break;
                    case HG_PASS_THERMAL_TRANSPORT_0:
                        X.p[HG_P_ROCK] = tp[T_R3]; X.p[HG_P_DIRT] = tp[T_D3];
                        break;
                    case HG_PASS_THERMAL_TRANSPORT_1:
                        X.p[HG_P_ROCK] = tp[T_R5]; X.p[HG_P_DIRT] = tp[T_D3];
                        break;
                    case HG_PASS_SMOOTH:
                    case UNIT_SMOOTH_THERMAL1:             // and tp[T_D3], passed on its own
                        X.p[HG_P_ROCK] = tp[T_R5]; X.p[HG_P_DIRT] = tp[T_D6];
                        break;
                    default: break;
                }
                return X;
            };
            // fresh adjoint buffers of every unit's outputs (each is written
            // in full by its kernel: no zeroing)
            std::vector<std::vector<float*>> fresh_of(order.size());
            for (size_t oi = 0; oi < order.size(); ++oi)
                for (size_t q = 0; q < produced(order[oi]).size(); ++q) fresh_of[oi].push_back(take(w.pool));
            // the transport's adjoint planes (before, after) for the far taps
            // of UNIT_THERMAL0_FAR, set by UNIT_ST1_TRANSPORT
            Planes far_lo{}, far_li{};
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
                        k_adj_flux<<<gtile, GNT, 0, c->stream>>>(X, lo, li, n, n, c->kp);
                        break;
                    case HG_PASS_EROSION:
                        k_adj_erosion<<<gtile, GNT, 0, c->stream>>>(X, lo, li, n, n, c->kp, w.gpart);
                        break;
                    case UNIT_EROSION_FLUX:
                        k_adj_erosion_flux<<<gtile, GNT, 0, c->stream>>>(X, lo, li, n, n, c->kp, w.gpart);
                        break;
                    case HG_PASS_TRANSPORT:
                        k_adj_transport<<<gtile, GNT, 0, c->stream>>>(X, lo, li, n, n, c->kp, w.gpart);
                        HG_HIP(c, hipGetLastError());
                        k_adj_transport_far<<<grd, blk, 0, c->stream>>>(X, lo, li, n, n, c->kp);
                        break;
                    case UNIT_ST1_TRANSPORT: {
                        const Planes Xs = x_of(UNIT_SMOOTH_THERMAL1), Xt = x_of(HG_PASS_TRANSPORT);
                        k_adj_smooth_thermal1_transport<<<2 * gtile.x * gtile.y, GNT, 0, c->stream>>>(
                            Xs, lo, li, tp[T_D3], Xt, lo, li, n, n, c->kp, w.gpart);
                        far_lo = lo;
                        far_li = li;
                        break;
                    }
                    case UNIT_THERMAL0_FAR: {
                        const Planes Xt = x_of(HG_PASS_TRANSPORT), X0 = x_of(HG_PASS_THERMAL_TRANSPORT_0);
                        const unsigned nfar = unsigned((size_t(n) * size_t(n) + GNT - 1) / GNT);
                        k_adj_thermal0_far<<<gtile.x * gtile.y + nfar, GNT, 0, c->stream>>>(X0, lo, li, Xt, far_lo,
                                                                                            far_li, n, n, c->kp);
                        break;
                    }
                    case UNIT_TRANSPORT_THERMAL0: {
                        const Planes Xt = x_of(HG_PASS_TRANSPORT), X0 = x_of(HG_PASS_THERMAL_TRANSPORT_0);
                        k_adj_transport<<<gtile, GNT, 0, c->stream>>>(Xt, lo, li, n, n, c->kp, w.gpart);
                        HG_HIP(c, hipGetLastError());
                        const unsigned nfar = unsigned((size_t(n) * size_t(n) + GNT - 1) / GNT);
                        k_adj_thermal0_far<<<gtile.x * gtile.y + nfar, GNT, 0, c->stream>>>(X0, lo, li, Xt, lo, li,
                                                                                            n, n, c->kp);
                        break;
                    }
                    case HG_PASS_THERMAL_TRANSPORT_0:
                    case HG_PASS_THERMAL_TRANSPORT_1:
                        if (pass == HG_PASS_THERMAL_TRANSPORT_0)
                            k_adj_thermal<0><<<gtile, GNT, 0, c->stream>>>(X, lo, li, n, n, c->kp);
                        else
                            k_adj_thermal<1><<<gtile, GNT, 0, c->stream>>>(X, lo, li, n, n, c->kp);
                        break;
                    case HG_PASS_SMOOTH:
                        k_adj_smooth<<<gtile, GNT, 0, c->stream>>>(X, lo, li, n, n, c->kp);
                        break;
                    case UNIT_SMOOTH_THERMAL1:
                        k_adj_smooth_thermal1<<<gtile, GNT, 0, c->stream>>>(X, lo, li, tp[T_D3], n, n, c->kp);
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
        HG_HIP(c, hipEventRecord(ev_rev[b], c->stream));   // plane set b is free again
    }
    c->step = step_end;
    {
        const int rc = read_pinned(c, lam[HG_P_ROCK], dL_dH0, pn);
        if (rc != HG_OK) return rc;
    }
    std::vector<double> part(nblk * 4);
    HG_HIP(c, hipMemcpyAsync(part.data(), w.gpart, nblk * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HG_HIP(c, hipStreamSynchronize(c->stream));
    for (int j = 0; j < 4; ++j) {
        double sum = 0.0;
        for (size_t b = 0; b < nblk; ++b) sum += part[b * 4 + j];
        dL_dcoef[j] = sum;
    }
    return HG_OK;
}
