#pragma once

/**
 * include/stencil/types.hpp -- core types of the MI355X stencil library.
 *
 * Mirrors the reference's include/stencil/types.hpp: real_t, STENCIL_RADIUS,
 * GridDims, FDCoeffs<R> and the roofline byte/flop models, with the same
 * names and meaning, so code written against the reference compiles here.
 * The device side is HIP for gfx950; nothing here pulls in HIP headers.
 */

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

// ── Floating-point precision (reference types.hpp:33-37) ─────────────────────
#ifdef STENCIL_FP64
using real_t = double;
#else
using real_t = float;
#endif

// ── Stencil radius (reference types.hpp:42-44) ───────────────────────────────
#ifndef STENCIL_RADIUS
#define STENCIL_RADIUS 4
#endif
static_assert(STENCIL_RADIUS == 4, "the gfx950 kernels implement the 8th-order (R = 4) stencil");

// Tracing ranges (the reference's NVTX_* macros, types.hpp:85-101): roctx when
// built with STENCIL_ENABLE_ROCTX, otherwise nothing.
#ifdef STENCIL_ENABLE_ROCTX
#include <roctracer/roctx.h>
#define STENCIL_PUSH(name) roctxRangePushA(name)
#define STENCIL_POP() roctxRangePop()
#else
#define STENCIL_PUSH(name) \
    do {                   \
    } while (0)
#define STENCIL_POP() \
    do {              \
    } while (0)
#endif

namespace stencil {

/**
 * GridDims -- the reference's grid descriptor (types.hpp:115-140).  Dims
 * include an R-wide halo on every face; the interior is [R, n-R)^3; layout is
 * z-slowest, idx = z*nx*ny + y*nx + x.
 */
template <class T>
struct BasicGridDims {
    std::size_t nx = 0;
    std::size_t ny = 0;
    std::size_t nz = 0;
    T dx = 1;
    T dy = 1;
    T dz = 1;
    T dt = 1;

    std::size_t size() const noexcept { return nx * ny * nz; }
    std::size_t idx(std::size_t x, std::size_t y, std::size_t z) const noexcept {
        return z * nx * ny + y * nx + x;
    }
    std::size_t interior_nx() const noexcept { return nx - 2 * STENCIL_RADIUS; }
    std::size_t interior_ny() const noexcept { return ny - 2 * STENCIL_RADIUS; }
    std::size_t interior_nz() const noexcept { return nz - 2 * STENCIL_RADIUS; }
    std::size_t interior_size() const noexcept {
        return interior_nx() * interior_ny() * interior_nz();
    }

    /// CFL limit used by Config::make_grid (reference types.hpp:139).
    static constexpr T CFL_LIMIT = static_cast<T>(0.45);
};
using GridDims = BasicGridDims<real_t>;

/**
 * FDCoeffs<R> -- centred second-derivative weights (reference
 * types.hpp:151-181): u'' ~ c0 u[i] + sum_r c[r-1] (u[i+r] + u[i-r]).
 */
template <int R>
struct FDCoeffs;

template <>
struct FDCoeffs<1> {
    static constexpr real_t c[1] = {static_cast<real_t>(1)};
    static constexpr real_t c0 = static_cast<real_t>(-2);
};

template <>
struct FDCoeffs<2> {
    static constexpr real_t c[2] = {static_cast<real_t>(4.0 / 3.0),
                                    static_cast<real_t>(-1.0 / 12.0)};
    static constexpr real_t c0 = static_cast<real_t>(-5.0 / 2.0);
};

template <>
struct FDCoeffs<3> {
    static constexpr real_t c[3] = {static_cast<real_t>(3.0 / 2.0),
                                    static_cast<real_t>(-3.0 / 20.0),
                                    static_cast<real_t>(1.0 / 90.0)};
    static constexpr real_t c0 = static_cast<real_t>(-49.0 / 18.0);
};

template <>
struct FDCoeffs<4> {
    static constexpr real_t c[4] = {static_cast<real_t>(8.0 / 5.0), static_cast<real_t>(-1.0 / 5.0),
                                    static_cast<real_t>(8.0 / 315.0),
                                    static_cast<real_t>(-1.0 / 560.0)};
    static constexpr real_t c0 = static_cast<real_t>(-205.0 / 72.0);
};

// ── Roofline models (reference types.hpp:198-227) ────────────────────────────
/// No-reuse model: (6R+4) values per interior point (tracks L2, not HBM).
template <class T>
inline double bandwidth_bytes_per_step(const BasicGridDims<T>& g) {
    const double N = static_cast<double>(g.interior_size());
    return N * static_cast<double>(6 * STENCIL_RADIUS + 4) * sizeof(T);
}

/// Compulsory HBM traffic: read p_cur, p_prev, vel2 and write p_next once per
/// interior point = 4 * sizeof(T).  The algorithmic bytes of the roofline.
template <class T>
inline double compulsory_bytes_per_step(const BasicGridDims<T>& g) {
    const double N = static_cast<double>(g.interior_size());
    return N * 4.0 * sizeof(T);
}

/// 3*(4R+1) + 5 = 56 flops per interior point.
template <class T>
inline double flops_per_step(const BasicGridDims<T>& g) {
    const double N = static_cast<double>(g.interior_size());
    return N * static_cast<double>(3 * (4 * STENCIL_RADIUS + 1) + 5);
}

}  // namespace stencil
