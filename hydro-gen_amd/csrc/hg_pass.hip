// hg_pass.hip — one kernel per reference compute dispatch (the "plain",
// unfused form of Erosion::dispatch_grid, src/erosion.cpp:158-200).
//
// Each kernel reads the input planes and writes the output planes (Jacobi, like
// the reference's Tex_pair ping-pong).  One thread per cell; 64x4 blocks so a
// wavefront covers 64 consecutive cells of a row (coalesced SoA streams);
// neighbour reads go through the caches.
#include "hg_kernels.hpp"

namespace hgk {

namespace {

constexpr int BX = 64, BY = 4;

__device__ __forceinline__ size_t at(int x, int y, int nx) { return size_t(y) * size_t(nx) + size_t(x); }

__global__ __launch_bounds__(BX * BY) void k_rain(Planes in, Planes out, int nx, int ny,
                                                  uint32_t step, uint32_t terrain, KParams k) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const size_t i = at(x, y, nx);
    out.p[HG_P_WATER][i] = in.p[HG_P_WATER][i] + rain_incr(x, y, step, terrain, k);
}

// glsl/rain.glsl:32-56 (HG_RAIN_REFERENCE)
__global__ __launch_bounds__(BX * BY) void k_rain_reference(Planes in, Planes out, int nx, int ny, KParams k) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const size_t i = at(x, y, nx);
    const float b = in.p[HG_P_WATER][i];
    const float w = in.p[HG_P_ROCK][i] + in.p[HG_P_DIRT][i] + b;
    out.p[HG_P_WATER][i] = b + ref_rain_incr(x, y, w, k);
}

// glsl/hydro_flux.glsl:77-166
__global__ __launch_bounds__(BX * BY) void k_flux(Planes in, Planes out, int nx, int ny, KParams k) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const size_t i = at(x, y, nx);
    const float* R = in.p[HG_P_ROCK];
    const float* D = in.p[HG_P_DIRT];
    const float* B = in.p[HG_P_WATER];
    auto total = [&](int xx, int yy) -> float {
        if (xx < 0 || yy < 0 || xx >= nx || yy >= ny) return HG_OOB_HEIGHT;
        const size_t j = at(xx, yy, nx);
        return R[j] + D[j] + B[j];
    };
    float f[4] = {in.p[HG_P_FLUX_L][i], in.p[HG_P_FLUX_R][i], in.p[HG_P_FLUX_T][i], in.p[HG_P_FLUX_B][i]};
    const float w = R[i] + D[i] + B[i];
    const float wn[4] = {total(x - 1, y), total(x + 1, y), total(x, y + 1), total(x, y - 1)};
    const float fin[4] = {x > 0 ? in.p[HG_P_FLUX_R][i - 1] : 0.0f,
                          x < nx - 1 ? in.p[HG_P_FLUX_L][i + 1] : 0.0f,
                          y < ny - 1 ? in.p[HG_P_FLUX_B][i + nx] : 0.0f,
                          y > 0 ? in.p[HG_P_FLUX_T][i - nx] : 0.0f};
    const unsigned wall = (x == 0 ? 1u : 0u) | (x == nx - 1 ? 2u : 0u) | (y == 0 ? 4u : 0u) |
                          (y == ny - 1 ? 8u : 0u);
    float vx, vy, vz;
    const float d2 = flux_cell(f, B[i], w, wn, fin, wall, k, vx, vy, vz);
    out.p[HG_P_WATER][i] = d2;
    out.p[HG_P_FLUX_L][i] = f[0];
    out.p[HG_P_FLUX_R][i] = f[1];
    out.p[HG_P_FLUX_T][i] = f[2];
    out.p[HG_P_FLUX_B][i] = f[3];
    out.p[HG_P_VEL_U][i] = vx;
    out.p[HG_P_VEL_V][i] = vy;
    out.p[HG_P_VEL_DEPTH][i] = vz;
}

// glsl/hydro_erosion.glsl:37-92
__global__ __launch_bounds__(BX * BY) void k_erosion(Planes in, Planes out, int nx, int ny, KParams k) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const size_t i = at(x, y, nx);
    const float* R = in.p[HG_P_ROCK];
    const float* D = in.p[HG_P_DIRT];
    auto rd = [&](int xx, int yy) -> float {  // out-of-grid imageLoad reads 0
        if (xx < 0 || yy < 0 || xx >= nx || yy >= ny) return 0.0f;
        const size_t j = at(xx, yy, nx);
        return R[j] + D[j];
    };
    float r = R[i], d = D[i], s0 = in.p[HG_P_SED_ROCK][i], s1 = in.p[HG_P_SED_DIRT][i];
    erosion_cell(r, d, s0, s1, in.p[HG_P_VEL_U][i], in.p[HG_P_VEL_V][i], in.p[HG_P_VEL_DEPTH][i],
                 rd(x - 1, y), rd(x + 1, y), rd(x, y + 1), rd(x, y - 1), k);
    out.p[HG_P_ROCK][i] = r;
    out.p[HG_P_DIRT][i] = d;
    out.p[HG_P_SED_ROCK][i] = s0;
    out.p[HG_P_SED_DIRT][i] = s1;
}

// glsl/sediment_transport.glsl:66-93 with img_bilinear (glsl/img_interpolation.glsl:3-22)
__global__ __launch_bounds__(BX * BY) void k_transport(Planes in, Planes out, int nx, int ny,
                                                       KParams k) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const size_t i = at(x, y, nx);
    float bx, by;
    back_trace(x, y, in.p[HG_P_VEL_U][i], in.p[HG_P_VEL_V][i], k.dt, nx, ny, bx, by);
    const int ix = int(bx), iy = int(by);
    const float fx = bx - float(ix), fy = by - float(iy);
    // the +1 taps carry weight 0 at the last row/column; clamp them in range
    const int ix1 = min(ix + 1, nx - 1), iy1 = min(iy + 1, ny - 1);
    const float* S0 = in.p[HG_P_SED_ROCK];
    const float* S1 = in.p[HG_P_SED_DIRT];
    const size_t j00 = at(ix, iy, nx), j10 = at(ix1, iy, nx), j01 = at(ix, iy1, nx), j11 = at(ix1, iy1, nx);
    out.p[HG_P_SED_ROCK][i] = lerp1(lerp1(S0[j00], S0[j10], fx), lerp1(S0[j01], S0[j11], fx), fy);
    out.p[HG_P_SED_DIRT][i] = lerp1(lerp1(S1[j00], S1[j10], fx), lerp1(S1[j01], S1[j11], fx), fy);
    out.p[HG_P_WATER][i] = in.p[HG_P_WATER][i] * k.evap;
}

__device__ const int TDX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
__device__ const int TDY[8] = {0, 0, 1, -1, 1, 1, -1, -1};

// glsl/thermal_erosion.glsl:28-115
__global__ __launch_bounds__(BX * BY) void k_thermal_flux(Planes in, Planes out, int nx, int ny,
                                                          int layer, KParams k) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const size_t i = at(x, y, nx);
    const float* R = in.p[HG_P_ROCK];
    const float* Dt = in.p[HG_P_DIRT];
    const float r = R[i], d = Dt[i];
    float Dk[8];
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        const int xx = x + TDX[kk], yy = y + TDY[kk];
        const bool ok = xx >= 0 && yy >= 0 && xx < nx && yy < ny;
        const size_t j = ok ? at(xx, yy, nx) : 0;
        const float rn = ok ? R[j] : HG_OOB_HEIGHT;
        float v = r - rn;
        if (layer == 1) v = v + (d - (ok ? Dt[j] : HG_OOB_HEIGHT));
        Dk[kk] = v;
    }
    const float g = thermal_g(Dk, layer == 0 ? r : d, layer, k);
#pragma unroll
    for (int kk = 0; kk < 8; ++kk)
        out.p[HG_P_THERM_L + kk][i] = talus_marked(Dk[kk], kk, layer, k) ? g * Dk[kk] : 0.0f;
}

// glsl/thermal_transport.glsl:146-180
__global__ __launch_bounds__(BX * BY) void k_thermal_transport(Planes in, Planes out, int nx, int ny,
                                                               int layer) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const size_t i = at(x, y, nx);
    // neighbour k sends its outflow in the opposite direction
    const int opp[8] = {1, 0, 3, 2, 7, 6, 5, 4};
    float acc = 0.0f;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        const int xx = x + TDX[kk], yy = y + TDY[kk];
        const bool ok = xx >= 0 && yy >= 0 && xx < nx && yy < ny;
        const float fin = ok ? in.p[HG_P_THERM_L + opp[kk]][at(xx, yy, nx)] : 0.0f;
        acc = acc + (fin - in.p[HG_P_THERM_L + kk][i]);
    }
    const int hp = layer == 0 ? HG_P_ROCK : HG_P_DIRT;
    out.p[hp][i] = in.p[hp][i] + acc;
}

// glsl/smoothing.glsl:22-74,97-103 (grid branch)
__global__ __launch_bounds__(BX * BY) void k_smooth(Planes in, Planes out, int nx, int ny, KParams k) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const size_t i = at(x, y, nx);
    const float* R = in.p[HG_P_ROCK];
    const float* D = in.p[HG_P_DIRT];
    float r = R[i], d = D[i];
    if (x > 0 && y > 0 && x < nx - 1 && y < ny - 1)
        smooth_cell(r, d, R[i - 1], D[i - 1], R[i + 1], D[i + 1], R[i + nx], D[i + nx], R[i - nx],
                    D[i - nx], k.smooth_m);
    out.p[HG_P_ROCK][i] = r;
    out.p[HG_P_DIRT][i] = d;
}

}  // namespace

hipError_t launch_pass(int pass, const Planes& in, const Planes& out, int nx, int ny, uint32_t step,
                       uint32_t terrain, const KParams& k, hipStream_t s) {
    const dim3 block(BX, BY), grid((nx + BX - 1) / BX, (ny + BY - 1) / BY);
    switch (pass) {
        case HG_PASS_RAIN:
            if (k.rain_mode == HG_RAIN_REFERENCE) k_rain_reference<<<grid, block, 0, s>>>(in, out, nx, ny, k);
            else k_rain<<<grid, block, 0, s>>>(in, out, nx, ny, step, terrain, k);
            break;
        case HG_PASS_FLUX: k_flux<<<grid, block, 0, s>>>(in, out, nx, ny, k); break;
        case HG_PASS_EROSION: k_erosion<<<grid, block, 0, s>>>(in, out, nx, ny, k); break;
        case HG_PASS_TRANSPORT: k_transport<<<grid, block, 0, s>>>(in, out, nx, ny, k); break;
        case HG_PASS_THERMAL_FLUX_0: k_thermal_flux<<<grid, block, 0, s>>>(in, out, nx, ny, 0, k); break;
        case HG_PASS_THERMAL_TRANSPORT_0: k_thermal_transport<<<grid, block, 0, s>>>(in, out, nx, ny, 0); break;
        case HG_PASS_THERMAL_FLUX_1: k_thermal_flux<<<grid, block, 0, s>>>(in, out, nx, ny, 1, k); break;
        case HG_PASS_THERMAL_TRANSPORT_1: k_thermal_transport<<<grid, block, 0, s>>>(in, out, nx, ny, 1); break;
        case HG_PASS_SMOOTH: k_smooth<<<grid, block, 0, s>>>(in, out, nx, ny, k); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace hgk
