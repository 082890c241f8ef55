// oracle_capi.cpp — extern "C" wrappers of the CPU oracle for ctypes.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg, never by the product library.
//
// State is passed as one contiguous array of ORC_NUM_PLANES planes of n = nx*ny
// values each, in the order of hg.h's HG_P_* enum:
//   rock, dirt, water, flux L/R/T/B, sed rock, sed dirt, vel u/v/depth,
//   thermal L/R/T/B, thermal LT/RT/LB/RB.
#include <cstdint>
#include <cstring>
#include <vector>
#include "hg_oracle.hpp"
#include "hg_dual.hpp"
#include "../include/hg.h"

static_assert(sizeof(hg_erosion_settings) == 96, "Erosion_data is 96 bytes in std140");

namespace {
constexpr int ORC_NUM_PLANES = HG_NUM_PLANES;

template <class T> orc::Settings<T> to_settings(const hg_erosion_settings* s) {
    orc::Settings<T> k;
    k.particle_count = s->particle_count;
    k.Kc = T(s->Kc);
    k.Kalpha[0] = T(s->Kalpha[0]); k.Kalpha[1] = T(s->Kalpha[1]);
    k.Kconv = T(s->Kconv);
    k.Ks[0] = T(s->Ks[0]); k.Ks[1] = T(s->Ks[1]);
    k.Kd[0] = T(s->Kd[0]); k.Kd[1] = T(s->Kd[1]);
    k.Ke = T(s->Ke);
    k.E = T(s->energy_kept);
    k.Kspeed[0] = T(s->Kspeed[0]); k.Kspeed[1] = T(s->Kspeed[1]);
    k.G = T(s->G);
    k.dt = T(s->d_t);
    return k;
}

template <class T> void load(orc::State<T>& st, const T* planes, int nx, int ny) {
    st.resize(nx, ny);
    size_t n = size_t(nx) * ny;
    auto ps = st.planes();
    for (int p = 0; p < ORC_NUM_PLANES; ++p) std::memcpy(ps[p]->data(), planes + p * n, n * sizeof(T));
}

template <class T> void store(orc::State<T>& st, T* planes) {
    size_t n = size_t(st.nx) * st.ny;
    auto ps = st.planes();
    for (int p = 0; p < ORC_NUM_PLANES; ++p) std::memcpy(planes + p * n, ps[p]->data(), n * sizeof(T));
}

template <class T>
int run(T* planes, int nx, int ny, const hg_erosion_settings* s, int rain_mode, double rain_amount,
        uint64_t seed, uint32_t terrain, uint64_t step0, int nsteps, unsigned flags, int single_pass) {
    if (!planes || !s || nx < 2 || ny < 2) return -1;
    orc::State<T> st;
    load(st, planes, nx, ny);
    orc::Settings<T> k = to_settings<T>(s);
    orc::RainCfg rc;
    rc.mode = rain_mode;
    rc.amount = rain_amount;
    rc.seed = seed;
    rc.terrain = terrain;
    if (single_pass >= 0) {
        orc::run_pass(st, k, rc, step0, single_pass);
    } else {
        for (int i = 0; i < nsteps; ++i) orc::grid_step(st, k, rc, step0 + uint64_t(i), flags);
    }
    store(st, planes);
    return 0;
}
// Loss of a rollout: L = mean over cells of (rock + dirt - target)^2 after
// nsteps GRID steps from `planes`, with (Kc, Ks, Kd, Ke) = coef (Ks, Kd for
// both layers) and the other settings from s.
template <class T>
T rollout_loss(orc::State<T>& st, const hg_erosion_settings* s, const T coef[4], int rain_mode,
               double rain_amount, uint64_t seed, int nsteps, unsigned flags, const double* target) {
    orc::Settings<T> k = to_settings<T>(s);
    k.Kc = coef[0];
    k.Ks[0] = coef[1]; k.Ks[1] = coef[1];
    k.Kd[0] = coef[2]; k.Kd[1] = coef[2];
    k.Ke = coef[3];
    orc::RainCfg rc;
    rc.mode = rain_mode;
    rc.amount = rain_amount;
    rc.seed = seed;
    for (int i = 0; i < nsteps; ++i) orc::grid_step(st, k, rc, uint64_t(i), flags);
    const size_t n = size_t(st.nx) * st.ny;
    T acc = T(0);
    for (size_t i = 0; i < n; ++i) {
        const T e = st.r[i] + st.d[i] - T(target[i]);
        acc += e * e;
    }
    return acc / T(double(n));
}

constexpr int DUAL_N = 36;   // 4 coefficients + up to 32 cells

template <class T> orc::RefParticleParams<T> ref_particle_params(const hg_erosion_settings* s) {
    orc::RefParticleParams<T> k;
    k.particle_count = s->particle_count;
    k.ttl = s->ttl;
    k.Kc = T(s->Kc); k.Kconv = T(s->Kconv); k.Ke = T(s->Ke); k.G = T(s->G); k.dt = T(s->d_t);
    for (int l = 0; l < 2; ++l) { k.Ks[l] = T(s->Ks[l]); k.Kd[l] = T(s->Kd[l]); }
    k.init_volume = T(s->init_volume); k.friction = T(s->friction); k.inertia = T(s->inertia);
    k.min_volume = T(s->min_volume); k.min_velocity = T(s->min_velocity);
    return k;
}

// `iterations` reference particle dispatches; planes as for orc_run (the
// FLUX planes hold the momentum map x, y, z, w), particles in the std430
// layout (converted to T for the run and rounded back to FP32 at the end).
template <class T>
int ref_particle_run(T* planes, int n, hg_ref_particle* parts, uint32_t count, const hg_erosion_settings* s,
                     uint32_t iterations, int should_rain, float time0, float time_step, unsigned flags,
                     int dims_x, int dims_y) {
    if (!planes || !s || !parts || n < 8) return -1;
    orc::State<T> st;
    load(st, planes, n, n);
    const orc::Settings<T> k = to_settings<T>(s);
    orc::RefParticleParams<T> pk = ref_particle_params<T>(s);
    pk.dims_x = dims_x;
    pk.dims_y = dims_y;
    std::vector<orc::RefParticle<T>> ps(count);
    for (uint32_t i = 0; i < count; ++i) {
        const hg_ref_particle& q = parts[i];
        auto& p = ps[i];
        p.sc = T(q.sc); p.iters = q.iters; p.px = T(q.position[0]); p.py = T(q.position[1]);
        p.vx = T(q.velocity[0]); p.vy = T(q.velocity[1]); p.volume = T(q.volume);
        p.sed[0] = T(q.sediment[0]); p.sed[1] = T(q.sediment[1]); p.to_kill = q.to_kill != 0;
    }
    std::vector<T> mom[4] = {st.fl, st.fr, st.ft, st.fb};
    for (uint32_t it = 0; it < iterations; ++it) {
        const float time = time0 + float(it) * time_step;
        orc::ref_particle_dispatch(st, mom, ps, pk, k, T(time), should_rain != 0, (flags & 2u) != 0, (flags & 4u) != 0,
                                   (flags & 0x100u) != 0);
    }
    st.fl = mom[0]; st.fr = mom[1]; st.ft = mom[2]; st.fb = mom[3];
    for (uint32_t i = 0; i < count; ++i) {
        const auto& p = ps[i];
        hg_ref_particle& q = parts[i];
        q.sc = float(p.sc); q.iters = p.iters; q.position[0] = float(p.px); q.position[1] = float(p.py);
        q.velocity[0] = float(p.vx); q.velocity[1] = float(p.vy); q.volume = float(p.volume);
        q.sediment[0] = float(p.sed[0]); q.sediment[1] = float(p.sed[1]); q.to_kill = p.to_kill ? 1u : 0u;
    }
    store(st, planes);
    return 0;
}
}  // namespace

extern "C" {

int orc_num_planes(void) { return ORC_NUM_PLANES; }

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    orc::Philox::gen(ctr, key, out);
}

double orc_rain_u(uint64_t seed, uint64_t step, uint32_t terrain, int x, int y) {
    return orc::rain_uniform_u(seed, step, terrain, x, y);
}

void orc_fbm_permutation(uint64_t seed, uint8_t perm[256]) { orc::fbm_permutation(seed, perm); }

double orc_simplex2_f64(uint64_t seed, double x, double y) {
    uint8_t perm[256];
    orc::fbm_permutation(seed, perm);
    orc::Simplex2<double> sx{perm};
    return sx(x, y);
}

int orc_fbm_init_f64(double* out, int nx, int ny, const hg_fbm_params* p) {
    if (!out || !p || nx < 1 || ny < 1) return -1;
    std::vector<double> v;
    orc::fbm_init<double>(v, nx, ny, p->seed, p->octaves, double(p->lacunarity), double(p->gain),
                          double(p->base_freq), double(p->height_lo), double(p->height_hi));
    std::memcpy(out, v.data(), v.size() * sizeof(double));
    return 0;
}

int orc_fbm_init_f32(float* out, int nx, int ny, const hg_fbm_params* p) {
    if (!out || !p || nx < 1 || ny < 1) return -1;
    std::vector<float> v;
    orc::fbm_init<float>(v, nx, ny, p->seed, p->octaves, p->lacunarity, p->gain, p->base_freq,
                         p->height_lo, p->height_hi);
    std::memcpy(out, v.data(), v.size() * sizeof(float));
    return 0;
}

// nsteps GRID steps (flags: 1 rain, 2 thermal, 4 smoothing), step counter
// starting at step0.  rain_amount is the FP32 setting value.
int orc_run_f64(double* planes, int nx, int ny, const hg_erosion_settings* s, int rain_mode,
                float rain_amount, uint64_t seed, uint32_t terrain, uint64_t step0, int nsteps,
                unsigned flags) {
    return run<double>(planes, nx, ny, s, rain_mode, double(rain_amount), seed, terrain, step0, nsteps,
                       flags, -1);
}

int orc_run_f32(float* planes, int nx, int ny, const hg_erosion_settings* s, int rain_mode,
                float rain_amount, uint64_t seed, uint32_t terrain, uint64_t step0, int nsteps,
                unsigned flags) {
    return run<float>(planes, nx, ny, s, rain_mode, double(rain_amount), seed, terrain, step0, nsteps,
                      flags, -1);
}

// One reference pass (HG_PASS_* numbering).
int orc_pass_f64(double* planes, int nx, int ny, const hg_erosion_settings* s, int rain_mode,
                 float rain_amount, uint64_t seed, uint32_t terrain, uint64_t step, int pass) {
    if (pass < 0 || pass >= HG_NUM_PASSES) return -1;
    return run<double>(planes, nx, ny, s, rain_mode, double(rain_amount), seed, terrain, step, 0, 0,
                       pass);
}

int orc_pass_f32(float* planes, int nx, int ny, const hg_erosion_settings* s, int rain_mode,
                 float rain_amount, uint64_t seed, uint32_t terrain, uint64_t step, int pass) {
    if (pass < 0 || pass >= HG_NUM_PASSES) return -1;
    return run<float>(planes, nx, ny, s, rain_mode, double(rain_amount), seed, terrain, step, 0, 0,
                      pass);
}

}  // extern "C"

// Smoothing extremum flags of every cell (glsl/smoothing.glsl:40-73): bit 0
// rock, bit 1 surface; 0 on the border, which the pass skips.
template <class T> int smooth_flags(const T* planes, int nx, int ny, uint8_t* flags) {
    if (!planes || !flags || nx < 3 || ny < 3) return -1;
    orc::State<T> st;
    load(st, planes, nx, ny);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            bool er = false, eg = false;
            if (x > 0 && y > 0 && x < nx - 1 && y < ny - 1) orc::smooth_flags_at(st, x, y, er, eg);
            flags[size_t(y) * nx + x] = uint8_t((er ? 1 : 0) | (eg ? 2 : 0));
        }
    return 0;
}

extern "C" {
int orc_smooth_flags_f64(const double* planes, int nx, int ny, uint8_t* flags) {
    return smooth_flags<double>(planes, nx, ny, flags);
}
int orc_smooth_flags_f32(const float* planes, int nx, int ny, uint8_t* flags) {
    return smooth_flags<float>(planes, nx, ny, flags);
}

// One droplet iteration (serial, droplets in id order) on rock/dirt planes.
// Returns the mass exported over the map edge.
double orc_droplets_f64(double* rock, double* dirt, int n, const hg_droplet_settings* s, uint32_t iteration) {
    orc::DropletParams p{s->count, s->max_lifetime, s->inertia, s->capacity, s->deposition, s->erosion,
                         s->evaporation, s->gravity, s->min_slope, s->brush_radius, s->init_water,
                         s->init_speed, s->seed};
    std::vector<double> r(rock, rock + size_t(n) * n), d(dirt, dirt + size_t(n) * n);
    const double ex = orc::droplets<double>(r, d, n, p, iteration);
    std::memcpy(rock, r.data(), r.size() * sizeof(double));
    std::memcpy(dirt, d.data(), d.size() * sizeof(double));
    return ex;
}

double orc_droplets_f32(float* rock, float* dirt, int n, const hg_droplet_settings* s, uint32_t iteration) {
    orc::DropletParams p{s->count, s->max_lifetime, s->inertia, s->capacity, s->deposition, s->erosion,
                         s->evaporation, s->gravity, s->min_slope, s->brush_radius, s->init_water,
                         s->init_speed, s->seed};
    std::vector<float> r(rock, rock + size_t(n) * n), d(dirt, dirt + size_t(n) * n);
    const double ex = orc::droplets<float>(r, d, n, p, iteration);
    std::memcpy(rock, r.data(), r.size() * sizeof(float));
    std::memcpy(dirt, d.data(), d.size() * sizeof(float));
    return ex;
}

// Reference rain (one pass) on the state planes; rs is the Rain_data block.
int orc_rain_reference_f64(double* planes, int nx, int ny, const hg_rain_settings* rs, float time,
                           float max_height) {
    if (!planes || !rs || nx < 1 || ny < 1) return -1;
    orc::State<double> st;
    load(st, planes, nx, ny);
    orc::RefRain rr{rs->amount, rs->mountain_thresh, rs->mountain_multip, rs->drops, time, max_height};
    orc::pass_rain_reference(st, rr);
    store(st, planes);
    return 0;
}

int orc_rain_reference_f32(float* planes, int nx, int ny, const hg_rain_settings* rs, float time,
                           float max_height) {
    if (!planes || !rs || nx < 1 || ny < 1) return -1;
    orc::State<float> st;
    load(st, planes, nx, ny);
    orc::RefRain rr{rs->amount, rs->mountain_thresh, rs->mountain_multip, rs->drops, time, max_height};
    orc::pass_rain_reference(st, rr);
    store(st, planes);
    return 0;
}

double orc_ref_simplex_f64(double x, double y) { return orc::ref_simplex<double>(x, y); }
float orc_ref_simplex_f32(float x, float y) { return orc::ref_simplex<float>(x, y); }

double orc_ref_sfbm_f64(double x, double y, double seed, double scale, int octaves) {
    return orc::ref_sfbm<double>(x, y, seed, 0.5, 2.0, scale, octaves);
}

static orc::MapCfg map_cfg(const hg_map_settings* s) {
    orc::MapCfg m;
    m.max_height = s->max_height; m.max_dirt = s->max_dirt; m.height_mult = s->height_mult; m.seed = s->seed;
    m.persistance = s->persistance; m.lacunarity = s->lacunarity; m.scale = s->scale;
    m.uplift_scale = s->uplift_scale; m.domain_warp_scale = s->domain_warp_scale;
    m.terrace_scale = s->terrace_scale; m.octaves = s->octaves; m.domain_warp = s->domain_warp;
    m.terrace = s->terrace; m.mask_round = s->mask_round; m.mask_exp = s->mask_exp;
    m.mask_power = s->mask_power; m.mask_slope = s->mask_slope; m.uplift = s->uplift;
    return m;
}

// Reference heightmap (rock and dirt planes of n x n).
int orc_gen_heightmap_f64(double* rock, double* dirt, int n, const hg_map_settings* s) {
    if (!rock || !dirt || !s || n < 1) return -1;
    std::vector<double> r, d;
    orc::gen_heightmap<double>(r, d, n, map_cfg(s));
    std::memcpy(rock, r.data(), r.size() * sizeof(double));
    std::memcpy(dirt, d.data(), d.size() * sizeof(double));
    return 0;
}

int orc_gen_heightmap_f32(float* rock, float* dirt, int n, const hg_map_settings* s) {
    if (!rock || !dirt || !s || n < 1) return -1;
    std::vector<float> r, d;
    orc::gen_heightmap<float>(r, d, n, map_cfg(s));
    std::memcpy(rock, r.data(), r.size() * sizeof(float));
    std::memcpy(dirt, d.data(), d.size() * sizeof(float));
    return 0;
}

void orc_elias_hash(int px, int py, double out[2]) { orc::elias_hash<double>(px, py, out[0], out[1]); }

void orc_noised_f64(double x, double y, double out[3]) { orc::noised<double>(x, y, out[0], out[1], out[2]); }

// FP64 rollout loss (see rollout_loss); planes are the initial state.
int orc_rollout_loss_f64(const double* planes, int nx, int ny, const hg_erosion_settings* s,
                         const double coef[4], int rain_mode, float rain_amount, uint64_t seed,
                         int nsteps, unsigned flags, const double* target, double* loss) {
    if (!planes || !s || !coef || !target || !loss || nx < 2 || ny < 2) return -1;
    orc::State<double> st;
    load(st, planes, nx, ny);
    *loss = rollout_loss<double>(st, s, coef, rain_mode, double(rain_amount), seed, nsteps, flags, target);
    return 0;
}

// Dual-number FP64 rollout: the loss and its derivatives with respect to
// (Kc, Ks, Kd, Ke) and to the initial rock height at `ncells` cells
// (grad[0..3], then grad[4 + i] for cells[i]).  The dual width is the
// smallest of 8, 16 and 36 that holds them: every arithmetic operation of
// the rollout costs one value plus the width in derivatives.
}  // extern "C"
template <int N>
int rollout_grad_dual(const double* planes, int nx, int ny, const hg_erosion_settings* s,
                      const double coef[4], int rain_mode, float rain_amount, uint64_t seed,
                      int nsteps, unsigned flags, const double* target, const int* cells,
                      int ncells, double* loss, double* grad) {
    using D = orc::Dual<N>;
    orc::State<D> st;
    st.resize(nx, ny);
    const size_t n = size_t(nx) * ny;
    auto ps = st.planes();
    for (int p = 0; p < ORC_NUM_PLANES; ++p)
        for (size_t i = 0; i < n; ++i) (*ps[p])[i] = D(planes[p * n + i]);
    for (int c = 0; c < ncells; ++c) {
        if (cells[c] < 0 || size_t(cells[c]) >= n) return -1;
        st.r[size_t(cells[c])].d[4 + c] = 1.0;
    }
    D cf[4];
    for (int j = 0; j < 4; ++j) cf[j] = D::seed(coef[j], j);
    const D L = rollout_loss<D>(st, s, cf, rain_mode, double(rain_amount), seed, nsteps, flags, target);
    *loss = L.v;
    for (int j = 0; j < 4 + ncells; ++j) grad[j] = L.d[j];
    return 0;
}
extern "C" {

int orc_rollout_grad_dual(const double* planes, int nx, int ny, const hg_erosion_settings* s,
                          const double coef[4], int rain_mode, float rain_amount, uint64_t seed,
                          int nsteps, unsigned flags, const double* target, const int* cells,
                          int ncells, double* loss, double* grad) {
    if (!planes || !s || !coef || !target || !loss || !grad || ncells < 0 || ncells > DUAL_N - 4) return -1;
    if (ncells <= 4)
        return rollout_grad_dual<8>(planes, nx, ny, s, coef, rain_mode, rain_amount, seed, nsteps, flags, target,
                                    cells, ncells, loss, grad);
    if (ncells <= 12)
        return rollout_grad_dual<16>(planes, nx, ny, s, coef, rain_mode, rain_amount, seed, nsteps, flags, target,
                                     cells, ncells, loss, grad);
    return rollout_grad_dual<DUAL_N>(planes, nx, ny, s, coef, rain_mode, rain_amount, seed, nsteps, flags, target,
                                     cells, ncells, loss, grad);
}


// Central differences of the rollout loss with every branch frozen at the
// unperturbed run (orc::Shadow): fd[j] = (L(+h_j) - L(-h_j)) / (2 h_j) for
// the directions (Kc, Ks, Kd, Ke) with h_j = eps_coef * coef[j], then the
// initial rock height at cells[i] with h = eps_cell.  Their limit as h -> 0
// is the derivative of the branch taken, what orc_rollout_grad_dual returns.
}  // extern "C"
template <int NC>
int rollout_fd_frozen(const double* planes, int nx, int ny, const hg_erosion_settings* s,
                      const double coef[4], int rain_mode, float rain_amount, uint64_t seed, int nsteps,
                      unsigned flags, const double* target, const int* cells, int ncells, double eps_coef,
                      double eps_cell, double* fd) {
    constexpr int K = 2 * (4 + NC);
    using S = orc::Shadow<K>;
    orc::State<S> st;
    st.resize(nx, ny);
    const size_t n = size_t(nx) * ny;
    auto ps = st.planes();
    for (int p = 0; p < ORC_NUM_PLANES; ++p)
        for (size_t i = 0; i < n; ++i) (*ps[p])[i] = S(planes[p * n + i]);
    if (ncells > NC) return -1;
    for (int c = 0; c < NC && c < ncells; ++c) {
        if (cells[c] < 0 || size_t(cells[c]) >= n) return -1;
        S& r = st.r[size_t(cells[c])];
        r.c[2 * (4 + c)] += eps_cell;
        r.c[2 * (4 + c) + 1] -= eps_cell;
    }
    S cf[4];
    for (int j = 0; j < 4; ++j) {
        cf[j] = S(coef[j]);
        cf[j].c[2 * j] = coef[j] + eps_coef * coef[j];
        cf[j].c[2 * j + 1] = coef[j] - eps_coef * coef[j];
    }
    const S L = rollout_loss<S>(st, s, cf, rain_mode, double(rain_amount), seed, nsteps, flags, target);
    for (int j = 0; j < 4 + NC && j < 4 + ncells; ++j) {
        const double h = j < 4 ? (cf[j].c[2 * j] - cf[j].c[2 * j + 1]) / 2 : eps_cell;
        fd[j] = (L.c[2 * j] - L.c[2 * j + 1]) / (2 * h);
    }
    return 0;
}
extern "C" {

int orc_rollout_fd_frozen(const double* planes, int nx, int ny, const hg_erosion_settings* s,
                          const double coef[4], int rain_mode, float rain_amount, uint64_t seed, int nsteps,
                          unsigned flags, const double* target, const int* cells, int ncells, double eps_coef,
                          double eps_cell, double* fd) {
    if (!planes || !s || !coef || !target || !fd || ncells < 0 || ncells > 12 || !(eps_coef > 0) ||
        !(eps_cell > 0))
        return -1;
    if (ncells == 0)
        return rollout_fd_frozen<0>(planes, nx, ny, s, coef, rain_mode, rain_amount, seed, nsteps, flags, target,
                                    cells, ncells, eps_coef, eps_cell, fd);
    if (ncells <= 4)
        return rollout_fd_frozen<4>(planes, nx, ny, s, coef, rain_mode, rain_amount, seed, nsteps, flags, target,
                                    cells, ncells, eps_coef, eps_cell, fd);
    return rollout_fd_frozen<12>(planes, nx, ny, s, coef, rain_mode, rain_amount, seed, nsteps, flags, target,
                                 cells, ncells, eps_coef, eps_cell, fd);
}

// The reference's particle mode (orc::ref_particle_dispatch); dims_x, dims_y
// = Map_settings_data.hmap_dims as the shader reads it (0 = n).
int orc_ref_particle_run_f64(double* planes, int n, hg_ref_particle* parts, uint32_t count,
                             const hg_erosion_settings* s, uint32_t iterations, int should_rain, float time0,
                             float time_step, unsigned flags, int dims_x, int dims_y) {
    return ref_particle_run<double>(planes, n, parts, count, s, iterations, should_rain, time0, time_step, flags,
                                    dims_x, dims_y);
}
int orc_ref_particle_run_f32(float* planes, int n, hg_ref_particle* parts, uint32_t count,
                             const hg_erosion_settings* s, uint32_t iterations, int should_rain, float time0,
                             float time_step, unsigned flags, int dims_x, int dims_y) {
    return ref_particle_run<float>(planes, n, parts, count, s, iterations, should_rain, time0, time_step, flags,
                                   dims_x, dims_y);
}
float orc_ref_particle_rand_f32(float x, float y) { return orc::ref_particle_rand<float>(x, y); }
// FP32 exp of the generator's exp mask (orc::gen_exp).
float orc_gen_exp_f32(float x) { return orc::gen_exp(x); }
// FP32 sin of the particle spawn hash (orc::hash_sin).
float orc_hash_sin_f32(float x) { return orc::hash_sin(x); }
double orc_ref_particle_rand_f64(double x, double y) { return orc::ref_particle_rand<double>(x, y); }
}  // extern "C"
