#pragma once

/**
 * include/stencil/kernels.hpp -- the launcher API of the reference
 * (include/stencil/kernels.cuh:33-83) in C++, over the C-ABI in
 * include/stencil_capi.h.
 *
 * Same names, arguments and behaviour: asynchronous on `stream`, no host
 * synchronisation, a failed HIP launch prints and exits like the reference's
 * CUDA_CHECK_LAST (types.hpp:64-74).  Invalid arguments (null pointers, zero
 * dims, p_next aliasing an input) throw std::invalid_argument instead of
 * faulting.  Overloaded for float and double (the STENCIL_FP64 build).
 */

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

#include "stencil/config.hpp"
#include "stencil/types.hpp"
#include "stencil_capi.h"

namespace stencil {

namespace detail {
template <class T>
inline stencil_grid to_capi(const BasicGridDims<T>& g) {
    stencil_grid c;
    c.nx = g.nx;
    c.ny = g.ny;
    c.nz = g.nz;
    c.dx = static_cast<double>(g.dx);
    c.dy = static_cast<double>(g.dy);
    c.dz = static_cast<double>(g.dz);
    c.dt = static_cast<double>(g.dt);
    return c;
}

inline void check(int status, const char* what) {
    if (status == STENCIL_OK) return;
    if (status == STENCIL_ERR_HIP) {
        std::fprintf(stderr, "[HIP ERROR] %s: %s (hip error %d)\n", what,
                     stencil_strerror(status), stencil_last_hip_error());
        std::exit(EXIT_FAILURE);
    }
    throw std::invalid_argument(std::string(what) + ": " + stencil_strerror(status));
}
}  // namespace detail

/// Reference kernels.cuh:33.  The coefficients are compile-time constants of
/// the gfx950 kernels; this checks that a device is current and usable.
inline void upload_fd_coefficients() {
    detail::check(stencil_upload_fd_coefficients(), "upload_fd_coefficients");
}

// Reference kernels.cuh:42-43.
inline void launch_naive(const float* p_cur, const float* p_prev, const float* vel2, float* p_next,
                         const BasicGridDims<float>& grid, hipStream_t stream = nullptr) {
    const stencil_grid g = detail::to_capi(grid);
    detail::check(stencil_launch_naive_f32(p_cur, p_prev, vel2, p_next, &g, stream), "launch_naive");
}
inline void launch_naive(const double* p_cur, const double* p_prev, const double* vel2,
                         double* p_next, const BasicGridDims<double>& grid,
                         hipStream_t stream = nullptr) {
    const stencil_grid g = detail::to_capi(grid);
    detail::check(stencil_launch_naive_f64(p_cur, p_prev, vel2, p_next, &g, stream), "launch_naive");
}

// Reference kernels.cuh:54-55.
inline void launch_tiled(const float* p_cur, const float* p_prev, const float* vel2, float* p_next,
                         const BasicGridDims<float>& grid, hipStream_t stream = nullptr) {
    const stencil_grid g = detail::to_capi(grid);
    detail::check(stencil_launch_tiled_f32(p_cur, p_prev, vel2, p_next, &g, stream), "launch_tiled");
}
inline void launch_tiled(const double* p_cur, const double* p_prev, const double* vel2,
                         double* p_next, const BasicGridDims<double>& grid,
                         hipStream_t stream = nullptr) {
    const stencil_grid g = detail::to_capi(grid);
    detail::check(stencil_launch_tiled_f64(p_cur, p_prev, vel2, p_next, &g, stream), "launch_tiled");
}

/// Reference kernels.cuh:63-64: the tiled kernel on one Z-slab whose Z-face
/// halos the caller has filled.  One GPU here: identical to launch_tiled.
template <class T>
inline void launch_slab(const T* p_cur, const T* p_prev, const T* vel2, T* p_next,
                        const BasicGridDims<T>& dims, hipStream_t stream = nullptr) {
    launch_tiled(p_cur, p_prev, vel2, p_next, dims, stream);
}

/// Reference kernels.cuh:68-83.
template <class T>
inline void launch(KernelVariant variant, const T* d_p_cur, const T* d_p_prev, const T* d_vel2,
                   T* d_p_next, const BasicGridDims<T>& grid, hipStream_t stream = nullptr) {
    switch (variant) {
        case KernelVariant::Naive:
            launch_naive(d_p_cur, d_p_prev, d_vel2, d_p_next, grid, stream);
            break;
        case KernelVariant::Tiled:
        default:
            launch_tiled(d_p_cur, d_p_prev, d_vel2, d_p_next, grid, stream);
            break;
    }
}

/// Dedicated single-level heat step t_next = t + (dt^2 alpha) Lap(t), where
/// grid.dt is the heat solver's dt_eff (replaces the wave-kernel + D2D copy of
/// the reference's HeatDiffusionSolver::step_impl, heat_diffusion.cpp:103-122).
inline void launch_heat(KernelVariant variant, const float* t, float* t_next, float alpha,
                        const BasicGridDims<float>& grid, hipStream_t stream = nullptr) {
    const stencil_grid g = detail::to_capi(grid);
    const int v = variant == KernelVariant::Naive ? STENCIL_VARIANT_NAIVE : STENCIL_VARIANT_TILED;
    detail::check(stencil_heat_step_f32(t, t_next, alpha, &g, v, stream), "launch_heat");
}

}  // namespace stencil
