// extern "C" definitions of the launcher half of include/stencil_capi.h.
//
// Validation, the host-side scalars (formed exactly as the reference
// launchers form them, src/kernels/stencil_naive.cu:96-99), and dispatch to the
// gfx950 kernels.  Nothing here allocates, synchronises or exits.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "launchers.hpp"
#include "stencil/grid.hpp"
#include "stencil_capi.h"

using namespace stencil::hipk;

namespace {
thread_local int g_last_hip = 0;
}  // namespace

int stencil::hipk::hip_status(hipError_t e) {
    if (e == hipSuccess) return STENCIL_OK;
    g_last_hip = static_cast<int>(e);
    return STENCIL_ERR_HIP;
}

namespace {

// Dims must be non-zero and the field must fit size_t indexing; dims of at
// most 2R on an axis leave no interior and make the launch a no-op (the
// reference kernels then write nothing either).
bool valid_grid(const stencil_grid* g) {
    if (!g || g->nx == 0 || g->ny == 0 || g->nz == 0) return false;
    if (g->nx > (1u << 30) || g->ny > (1u << 30) || g->nz > (1u << 30)) return false;
    const double cells = double(g->nx) * double(g->ny) * double(g->nz);
    return cells < 9.0e18;
}

Dims dims_of(const stencil_grid* g) {
    Dims d;
    d.nx = static_cast<int64_t>(g->nx);
    d.ny = static_cast<int64_t>(g->ny);
    d.nz = static_cast<int64_t>(g->nz);
    return d;
}

template <class T>
Scalars<T> scalars_of(const stencil_grid* g) {
    return make_scalars<T>(static_cast<T>(g->dx), static_cast<T>(g->dy), static_cast<T>(g->dz),
                           static_cast<T>(g->dt));
}

template <class T>
bool wave_args_ok(const T* cur, const T* prev, const T* vel2, const T* next,
                  const stencil_grid* g) {
    if (!cur || !prev || !vel2 || !next || !valid_grid(g)) return false;
    return next != cur && next != prev && next != vel2;
}

}  // namespace

extern "C" {

const char* stencil_strerror(int status) {
    switch (status) {
        case STENCIL_OK: return "ok";
        case STENCIL_ERR_INVALID: return "invalid argument";
        case STENCIL_ERR_HIP: return "HIP runtime error";
        case STENCIL_ERR_UNSUPPORTED: return "unsupported shape for this kernel";
        case STENCIL_ERR_NOMEM: return "out of device memory";
        case STENCIL_ERR_TIMEOUT: return "a device-side wait timed out (results invalid)";
        default: return "unknown status";
    }
}

int stencil_last_hip_error(void) { return g_last_hip; }

int stencil_upload_fd_coefficients(void) {
    int dev = -1;
    return hip_status(hipGetDevice(&dev));
}

int stencil_launch_naive_f32(const float* p_cur, const float* p_prev, const float* vel2,
                             float* p_next, const stencil_grid* grid, void* stream) {
    if (!wave_args_ok(p_cur, p_prev, vel2, p_next, grid)) return STENCIL_ERR_INVALID;
    return hip_status(wave_naive_f32(p_cur, p_prev, vel2, p_next, dims_of(grid),
                                     scalars_of<float>(grid), static_cast<hipStream_t>(stream)));
}

namespace {
// An explicit shape, else the autotuned one for these dims (kind 0 wave,
// 1 heat), else the library default.
bool shape_of(const stencil_tile_shape* t, const Dims& d, int kind, ZqShape* out) {
    ZqShape s;
    if (!t) {
        if (zq_cached_shape(d, kind, &s)) {
            *out = s;
            return true;
        }
    } else {
        if ((t->tile_x4 != 0 || t->tile_y != 0) && !zq_shape_valid(t->tile_x4, t->tile_y)) return false;
        if (t->n_chunks < 0) return false;
        s.txv = t->tile_x4;
        s.ty = t->tile_y;
        s.n_chunks = t->n_chunks;
        s.xcd_remap = t->xcd_remap;
        s.nt = t->nontemporal;
    }
    *out = s;
    return true;
}
}  // namespace

int stencil_launch_tiled_ex_f32(const float* p_cur, const float* p_prev, const float* vel2,
                                float* p_next, const stencil_grid* grid, int exact,
                                const stencil_tile_shape* tile, void* stream) {
    if (!wave_args_ok(p_cur, p_prev, vel2, p_next, grid)) return STENCIL_ERR_INVALID;
    const Dims d = dims_of(grid);
    const void* ptrs[4] = {p_cur, p_prev, vel2, p_next};
    ZqShape shape;
    if (!shape_of(tile, d, 0, &shape)) return STENCIL_ERR_INVALID;
    if (!zq_supported(d, ptrs, 4)) return STENCIL_ERR_UNSUPPORTED;
    return hip_status(wave_zq_f32(p_cur, p_prev, vel2, p_next, d, scalars_of<float>(grid),
                                  exact != 0, shape, static_cast<hipStream_t>(stream)));
}

int stencil_launch_tiled_f32(const float* p_cur, const float* p_prev, const float* vel2,
                             float* p_next, const stencil_grid* grid, void* stream) {
    const int st =
        stencil_launch_tiled_ex_f32(p_cur, p_prev, vel2, p_next, grid, 1, nullptr, stream);
    if (st != STENCIL_ERR_UNSUPPORTED) return st;
    // nx % 4 != 0 or unaligned bases: the coalesced kernel takes any shape.
    return stencil_launch_naive_f32(p_cur, p_prev, vel2, p_next, grid, stream);
}

int stencil_launch_f32(int variant, const float* p_cur, const float* p_prev, const float* vel2,
                       float* p_next, const stencil_grid* grid, void* stream) {
    switch (variant) {
        case STENCIL_VARIANT_NAIVE:
            return stencil_launch_naive_f32(p_cur, p_prev, vel2, p_next, grid, stream);
        case STENCIL_VARIANT_TILED:
        case STENCIL_VARIANT_MULTIGPU:
            return stencil_launch_tiled_f32(p_cur, p_prev, vel2, p_next, grid, stream);
        default: return STENCIL_ERR_INVALID;
    }
}

int stencil_launch_naive_f64(const double* p_cur, const double* p_prev, const double* vel2,
                             double* p_next, const stencil_grid* grid, void* stream) {
    if (!wave_args_ok(p_cur, p_prev, vel2, p_next, grid)) return STENCIL_ERR_INVALID;
    return hip_status(wave_naive_f64(p_cur, p_prev, vel2, p_next, dims_of(grid),
                                     scalars_of<double>(grid), static_cast<hipStream_t>(stream)));
}

// FP64 is the correctness path: the tiled variant runs the same kernel, as the
// reference's CPU build routes launch_tiled to launch_naive (stencil_naive.cu:173-176).
int stencil_launch_tiled_f64(const double* p_cur, const double* p_prev, const double* vel2,
                             double* p_next, const stencil_grid* grid, void* stream) {
    return stencil_launch_naive_f64(p_cur, p_prev, vel2, p_next, grid, stream);
}

int stencil_launch_f64(int variant, const double* p_cur, const double* p_prev,
                       const double* vel2, double* p_next, const stencil_grid* grid,
                       void* stream) {
    if (variant < STENCIL_VARIANT_NAIVE || variant > STENCIL_VARIANT_MULTIGPU)
        return STENCIL_ERR_INVALID;
    return stencil_launch_naive_f64(p_cur, p_prev, vel2, p_next, grid, stream);
}

int stencil_heat_step_ex_f32(const float* t, float* t_next, float alpha, const stencil_grid* grid,
                             int exact, const stencil_tile_shape* tile, void* stream) {
    if (!t || !t_next || t == t_next || !valid_grid(grid)) return STENCIL_ERR_INVALID;
    const Dims d = dims_of(grid);
    const void* ptrs[2] = {t, t_next};
    ZqShape shape;
    if (!shape_of(tile, d, 1, &shape)) return STENCIL_ERR_INVALID;
    if (!zq_supported(d, ptrs, 2)) return STENCIL_ERR_UNSUPPORTED;
    const Scalars<float> s = scalars_of<float>(grid);
    return hip_status(heat_zq_f32(t, t_next, d, s, s.dt2 * alpha, exact != 0, shape,
                                  static_cast<hipStream_t>(stream)));
}

int stencil_heat_step_f32(const float* t, float* t_next, float alpha, const stencil_grid* grid,
                          int variant, void* stream) {
    if (!t || !t_next || t == t_next || !valid_grid(grid)) return STENCIL_ERR_INVALID;
    if (variant < STENCIL_VARIANT_NAIVE || variant > STENCIL_VARIANT_MULTIGPU)
        return STENCIL_ERR_INVALID;
    const Scalars<float> s = scalars_of<float>(grid);
    // dt2 * vel2[c] of the reference with vel2 == alpha, formed in float.
    const float coef = s.dt2 * alpha;
    if (variant != STENCIL_VARIANT_NAIVE) {
        const int st = stencil_heat_step_ex_f32(t, t_next, alpha, grid, 1, nullptr, stream);
        if (st != STENCIL_ERR_UNSUPPORTED) return st;
    }
    return hip_status(
        heat_naive_f32(t, t_next, dims_of(grid), s, coef, static_cast<hipStream_t>(stream)));
}

int stencil_heat_step_f64(const double* t, double* t_next, double alpha, const stencil_grid* grid,
                          int variant, void* stream) {
    if (!t || !t_next || t == t_next || !valid_grid(grid)) return STENCIL_ERR_INVALID;
    if (variant < STENCIL_VARIANT_NAIVE || variant > STENCIL_VARIANT_MULTIGPU)
        return STENCIL_ERR_INVALID;
    const Scalars<double> s = scalars_of<double>(grid);
    return hip_status(heat_naive_f64(t, t_next, dims_of(grid), s, s.dt2 * alpha,
                                     static_cast<hipStream_t>(stream)));
}

int stencil_field16_bytes(int storage, const stencil_grid* grid, uint64_t* data_bytes,
                          uint64_t* exp_bytes) {
    if (!valid_grid(grid) || storage < STENCIL_STORE_BF16 || storage > STENCIL_STORE_BFP16)
        return STENCIL_ERR_INVALID;
    const uint64_t n = grid->nx * grid->ny * grid->nz;
    if (data_bytes) *data_bytes = 2 * n;
    if (exp_bytes) *exp_bytes = storage == STENCIL_STORE_BFP16 ? (n + 15) / 16 : 0;
    return STENCIL_OK;
}

namespace {
bool pow2(float s) {
    int e = 0;
    return s > 0.f && std::isfinite(s) && std::frexp(s, &e) == 0.5f;
}
bool field16_ok(int storage, const stencil_field16* f) {
    if (!f || !f->data || (reinterpret_cast<uintptr_t>(f->data) % 16) != 0) return false;
    return storage != STENCIL_STORE_BFP16 || f->exp != nullptr;
}
Field16 as_field(const stencil_field16* f) { return Field16{f->data, f->exp}; }
}  // namespace

int stencil_encode16(int storage, const float* src, const stencil_field16* dst, uint64_t n,
                     float scale, void* stream) {
    if (storage < STENCIL_STORE_BF16 || storage > STENCIL_STORE_BFP16) return STENCIL_ERR_INVALID;
    if (!src || !field16_ok(storage, dst) || n == 0 || n % 16 != 0 || !pow2(scale))
        return STENCIL_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(src) % 16 != 0) return STENCIL_ERR_INVALID;
    return hip_status(encode16(storage, src, as_field(dst), n, scale, static_cast<hipStream_t>(stream)));
}

int stencil_decode16(int storage, const stencil_field16* src, float* dst, uint64_t n, float scale,
                     void* stream) {
    if (storage < STENCIL_STORE_BF16 || storage > STENCIL_STORE_BFP16) return STENCIL_ERR_INVALID;
    if (!dst || !field16_ok(storage, src) || n == 0 || n % 16 != 0 || !pow2(scale))
        return STENCIL_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(dst) % 16 != 0) return STENCIL_ERR_INVALID;
    return hip_status(decode16(storage, as_field(src), dst, n, scale, static_cast<hipStream_t>(stream)));
}

int stencil_launch_tiled_16(int storage, const stencil_field16* p_cur, const stencil_field16* p_prev,
                            const stencil_field16* vel2, const stencil_field16* p_next,
                            const stencil_grid* grid, float p_scale, float v_scale, int exact,
                            const stencil_tile_shape* tile, void* stream) {
    if (storage < STENCIL_STORE_BF16 || storage > STENCIL_STORE_BFP16) return STENCIL_ERR_INVALID;
    if (!valid_grid(grid) || !field16_ok(storage, p_cur) || !field16_ok(storage, p_prev) ||
        !field16_ok(storage, vel2) || !field16_ok(storage, p_next))
        return STENCIL_ERR_INVALID;
    if (p_next->data == p_cur->data || p_next->data == p_prev->data || p_next->data == vel2->data)
        return STENCIL_ERR_INVALID;
    if (!pow2(p_scale) || !pow2(v_scale)) return STENCIL_ERR_INVALID;
    const Dims d = dims_of(grid);
    if (d.nx % 16 != 0 || d.nx * d.ny * 4 * 30 >= (int64_t(1) << 32)) return STENCIL_ERR_UNSUPPORTED;
    ZqShape shape;
    if (!shape_of(tile, d, 2 + storage, &shape)) return STENCIL_ERR_INVALID;
    return hip_status(wave_zq_16(storage, as_field(p_cur), as_field(p_prev), as_field(vel2),
                                 as_field(p_next), d, scalars_of<float>(grid), p_scale, v_scale, exact != 0, shape,
                                 static_cast<hipStream_t>(stream)));
}

double stencil_wave_bytes_per_step(const stencil_grid* g, int bytes_per_value) {
    if (!valid_grid(g) || g->nx <= 8 || g->ny <= 8 || g->nz <= 8) return 0.0;
    return 4.0 * bytes_per_value * double(g->nx - 8) * double(g->ny - 8) * double(g->nz - 8);
}

double stencil_heat_bytes_per_step(const stencil_grid* g, int bytes_per_value) {
    if (!valid_grid(g) || g->nx <= 8 || g->ny <= 8 || g->nz <= 8) return 0.0;
    return 2.0 * bytes_per_value * double(g->nx - 8) * double(g->ny - 8) * double(g->nz - 8);
}

int stencil_host_init_gaussian_f32(float* host, uint64_t nx, uint64_t ny, uint64_t nz, float cx,
                                   float cy, float cz, float sigma) {
    if (!host || nx == 0 || ny == 0 || nz == 0) return STENCIL_ERR_INVALID;
    stencil::BasicGridDims<float> g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    // init_gaussian writes the interior only; the halo must read zero.
    stencil::BasicHostGrid<float> h(g);
    stencil::init_gaussian(h, cx, cy, cz, sigma);
    std::copy(h.data.begin(), h.data.end(), host);
    return STENCIL_OK;
}

int stencil_host_layered_noise_vel2_f32(float* host, uint64_t nx, uint64_t ny, uint64_t nz,
                                        int n_layers, float vmin, float vmax, float noise,
                                        uint64_t seed) {
    if (!host || nx == 0 || ny == 0 || nz == 0 || n_layers <= 0) return STENCIL_ERR_INVALID;
    stencil::init_layered_noise_vel2(host, nx, ny, nz, n_layers, vmin, vmax, noise, seed);
    return STENCIL_OK;
}

int stencil_host_noise_cube_f32(float* host, uint64_t nx, uint64_t ny, uint64_t nz,
                                uint64_t side, float hot, uint64_t seed) {
    if (!host || nx == 0 || ny == 0 || nz == 0 || side > nx || side > ny || side > nz)
        return STENCIL_ERR_INVALID;
    stencil::init_noise_cube(host, nx, ny, nz, side, hot, seed);
    return STENCIL_OK;
}

}  // extern "C"
