// hg_kernels.hpp — kernel launch interface shared by the runtime and the
// kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "hg_cell.hpp"
#include "../../include/hg.h"

namespace hgk {

// Device plane pointers, indexed by HG_P_*.
struct Planes {
    float* p[HG_NUM_PLANES];
};

// Sum of 8 values in a fixed pairwise order (thermal in/out totals).
__device__ __forceinline__ float sum8(const float v[8]) {
    return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

// Thermal height change of one cell: inflow minus own outflow
// (glsl/thermal_transport.glsl:146-172).
__device__ __forceinline__ float thermal_delta(const float in[8], const float out[8]) {
    return sum8(in) - sum8(out);
}

// Rain increment of cell (x, y) at `step` (config rain modes).
__device__ __forceinline__ float rain_incr(int x, int y, uint32_t step, uint32_t terrain,
                                           const KParams& k) {
    if (k.rain_mode == HG_RAIN_CONSTANT) return k.rain_amount;
    if (k.rain_mode == HG_RAIN_UNIFORM) {
        uint32_t o[4];
        philox4x32_10(uint32_t(x) >> 2, uint32_t(y), step, terrain, k.seed_lo, k.seed_hi, o);
        const uint32_t w = (x & 3) == 0 ? o[0] : (x & 3) == 1 ? o[1] : (x & 3) == 2 ? o[2] : o[3];
        return k.rain_amount * u24(w);
    }
    return 0.0f;
}

// ---- launchers (all enqueue on `s`; return hipError_t of the launch) -------

// per-pass kernels (hg_pass.hip): one per reference dispatch
hipError_t launch_pass(int pass, const Planes& in, const Planes& out, int nx, int ny, uint32_t step,
                       const KParams& k, hipStream_t s);

// fused kernels (hg_fused.hip)
hipError_t launch_fused_hydro(const Planes& in, const Planes& out, int nx, int ny, uint32_t step,
                              bool rain, bool keep_velocity, const KParams& k, hipStream_t s);
hipError_t launch_fused_thermal(const Planes& in, const Planes& out, int nx, int ny, bool thermal,
                                bool smooth, const KParams& k, hipStream_t s);

// heightmap init (hg_init.hip)
void fbm_permutation(uint64_t seed, uint8_t perm[256]);
hipError_t launch_fbm(float* out, int nx, int ny, const uint8_t* perm_host, int octaves,
                      float lacunarity, float gain, float base_freq, hipStream_t s);
hipError_t launch_minmax(const float* v, size_t n, unsigned* mm_dev, hipStream_t s);
hipError_t launch_rescale(float* v, size_t n, const unsigned* mm_dev, float lo, float hi,
                          hipStream_t s);
hipError_t launch_fill(float* p, size_t n, float value, hipStream_t s);

}  // namespace hgk
