#pragma once

/**
 * include/stencil/grid.hpp -- host fields, initial conditions, the device
 * triple-buffer ring, validation and error norms.
 *
 * Mirrors the reference's include/stencil/grid.hpp (HostGrid :33-47, init_*
 * :57-99, DeviceGrid :110-183, utilities :220-229) on HIP.  Additions: seeded
 * counter-based input generators (u01, init_layered_noise_vel2,
 * init_noise_cube) and rel_l2_interior.
 */

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "stencil/types.hpp"

#define HIP_CHECK(call)                                                                     \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess) {                                                             \
            std::fprintf(stderr, "[HIP ERROR] %s:%d  %s\n", __FILE__, __LINE__,             \
                         hipGetErrorString(_e));                                            \
            std::exit(EXIT_FAILURE);                                                        \
        }                                                                                   \
    } while (0)

#define HIP_CHECK_LAST() HIP_CHECK(hipGetLastError())

namespace stencil {

namespace detail {
/// Runs fn(i) for i in [0, n) on up to 16 host threads (static blocks).
template <class F>
inline void parallel_for(std::ptrdiff_t n, F&& fn) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 64 || nt == 1) {
        for (std::ptrdiff_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    const std::ptrdiff_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const std::ptrdiff_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([b, e, &fn] {
            for (std::ptrdiff_t i = b; i < e; ++i) fn(i);
        });
    }
    for (auto& th : pool) th.join();
}
}  // namespace detail

// ── Host grid (reference grid.hpp:33-47) ─────────────────────────────────────
template <class T>
struct BasicHostGrid {
    std::vector<T> data;
    BasicGridDims<T> dims;

    explicit BasicHostGrid(const BasicGridDims<T>& g, T fill = T{0}) : data(g.size(), fill), dims(g) {}

    T& at(std::size_t x, std::size_t y, std::size_t z) { return data[dims.idx(x, y, z)]; }
    const T& at(std::size_t x, std::size_t y, std::size_t z) const { return data[dims.idx(x, y, z)]; }

    T* ptr() noexcept { return data.data(); }
    const T* ptr() const noexcept { return data.data(); }
    std::size_t bytes() const noexcept { return data.size() * sizeof(T); }
};
using HostGrid = BasicHostGrid<real_t>;

// ── Initial conditions (reference grid.hpp:57-99) ────────────────────────────
/// Gaussian pulse, interior only, exp evaluated in T, centre at c*(n-1).
template <class T>
inline void init_gaussian(BasicHostGrid<T>& g, T cx, T cy, T cz, T sigma) {
    const auto& d = g.dims;
    const T icx = cx * static_cast<T>(d.nx - 1);
    const T icy = cy * static_cast<T>(d.ny - 1);
    const T icz = cz * static_cast<T>(d.nz - 1);
    const T inv2s2 = T{1} / (T{2} * sigma * sigma);
    const std::size_t R = STENCIL_RADIUS;
    if (d.nz <= 2 * R) return;
    detail::parallel_for(static_cast<std::ptrdiff_t>(d.nz - 2 * R), [&](std::ptrdiff_t k) {
        const std::size_t iz = R + static_cast<std::size_t>(k);
        for (std::size_t iy = R; iy < d.ny - R; ++iy)
            for (std::size_t ix = R; ix < d.nx - R; ++ix) {
                const T dx = static_cast<T>(ix) - icx;
                const T dy = static_cast<T>(iy) - icy;
                const T dz = static_cast<T>(iz) - icz;
                g.at(ix, iy, iz) = std::exp(-(dx * dx + dy * dy + dz * dz) * inv2s2);
            }
    });
}

template <class T>
inline void init_constant_velocity(BasicHostGrid<T>& g, T v_ms) {
    std::fill(g.data.begin(), g.data.end(), v_ms * v_ms);
}

template <class T>
inline void init_layered_velocity(BasicHostGrid<T>& g, T v_top, T v_bot) {
    const auto& d = g.dims;
    for (std::size_t iz = 0; iz < d.nz; ++iz)
        for (std::size_t iy = 0; iy < d.ny; ++iy)
            for (std::size_t ix = 0; ix < d.nx; ++ix) {
                const T v = (iz < d.nz / 2) ? v_top : v_bot;
                g.at(ix, iy, iz) = v * v;
            }
}

// ── Seeded, counter-based inputs (additions) ─────────────────────────────────
/// SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014).
inline std::uint64_t splitmix64(std::uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/// U[0,1) with 24 random bits; a pure function of (seed, stream, idx).
inline float u01(std::uint64_t seed, std::uint64_t stream, std::uint64_t idx) {
    const std::uint64_t k = splitmix64(seed ^ splitmix64(stream)) + idx;
    return static_cast<float>(splitmix64(k) >> 40) * (1.0f / 16777216.0f);
}

/// vel2 over every cell: n_layers equal z-bands with v_band ~ U[vmin, vmax],
/// times (1 + noise * U[-1, 1]) per cell, squared.  Float arithmetic.
inline void init_layered_noise_vel2(float* vel2, std::size_t nx, std::size_t ny, std::size_t nz,
                                    int n_layers, float vmin, float vmax, float noise,
                                    std::uint64_t seed) {
    detail::parallel_for(static_cast<std::ptrdiff_t>(nz), [&](std::ptrdiff_t iz) {
        const std::size_t layer = static_cast<std::size_t>(iz) * static_cast<std::size_t>(n_layers) / nz;
        const float vl = vmin + (vmax - vmin) * u01(seed, 1, layer);
        for (std::size_t iy = 0; iy < ny; ++iy)
            for (std::size_t ix = 0; ix < nx; ++ix) {
                const std::size_t i = static_cast<std::size_t>(iz) * nx * ny + iy * nx + ix;
                const float f = 1.0f + noise * (2.0f * u01(seed, 2, i) - 1.0f);
                const float v = vl * f;
                vel2[i] = v * v;
            }
    });
}

/// Heat initial field: interior U[0,1] plus a centred cube of side `side`
/// set to `hot`; halo 0.
inline void init_noise_cube(float* t, std::size_t nx, std::size_t ny, std::size_t nz,
                            std::size_t side, float hot, std::uint64_t seed) {
    const std::size_t R = STENCIL_RADIUS;
    const std::size_t x0 = nx / 2 - side / 2, y0 = ny / 2 - side / 2, z0 = nz / 2 - side / 2;
    detail::parallel_for(static_cast<std::ptrdiff_t>(nz), [&](std::ptrdiff_t izs) {
        const std::size_t iz = static_cast<std::size_t>(izs);
        for (std::size_t iy = 0; iy < ny; ++iy)
            for (std::size_t ix = 0; ix < nx; ++ix) {
                const std::size_t i = iz * nx * ny + iy * nx + ix;
                const bool interior = ix >= R && ix < nx - R && iy >= R && iy < ny - R &&
                                      iz >= R && iz < nz - R;
                const bool in_cube = ix >= x0 && ix < x0 + side && iy >= y0 && iy < y0 + side &&
                                     iz >= z0 && iz < z0 + side;
                t[i] = !interior ? 0.0f : (in_cube ? hot : u01(seed, 3, i));
            }
    });
}

// ── Device grid (reference grid.hpp:110-183) ─────────────────────────────────
/**
 * DeviceGrid -- three time levels (prev/cur/next) and vel2 in HBM, zeroed at
 * construction (which is what keeps the Dirichlet halos at zero).  rotate()
 * advances the ring without copies: prev=buf[i], cur=buf[i+1], next=buf[i+2].
 * Each field is one hipMalloc (256-byte aligned), so the kernels' 16-byte
 * vector accesses are always aligned.
 */
template <class T>
class BasicDeviceGrid {
public:
    explicit BasicDeviceGrid(const BasicGridDims<T>& g) : dims_(g) {
        const std::size_t n = g.size() * sizeof(T);
        for (auto& p : buf_) HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), n));
        HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&vel2_), n));
        clear();
    }
    ~BasicDeviceGrid() {
        for (auto& p : buf_)
            if (p) (void)hipFree(p);
        if (vel2_) (void)hipFree(vel2_);
    }
    BasicDeviceGrid(const BasicDeviceGrid&) = delete;
    BasicDeviceGrid& operator=(const BasicDeviceGrid&) = delete;
    BasicDeviceGrid(BasicDeviceGrid&& o) noexcept : dims_(o.dims_), idx_(o.idx_), vel2_(o.vel2_) {
        for (int i = 0; i < 3; ++i) {
            buf_[i] = o.buf_[i];
            o.buf_[i] = nullptr;
        }
        o.vel2_ = nullptr;
    }

    const BasicGridDims<T>& dims() const noexcept { return dims_; }

    T* prev() noexcept { return buf_[idx_]; }
    T* cur() noexcept { return buf_[(idx_ + 1) % 3]; }
    T* next() noexcept { return buf_[(idx_ + 2) % 3]; }
    T* vel2() noexcept { return vel2_; }
    const T* prev() const noexcept { return buf_[idx_]; }
    const T* cur() const noexcept { return buf_[(idx_ + 1) % 3]; }

    void rotate() noexcept { idx_ = (idx_ + 1) % 3; }

    void clear() {
        const std::size_t n = dims_.size() * sizeof(T);
        for (auto& p : buf_) HIP_CHECK(hipMemset(p, 0, n));
    }

    void upload_cur(const BasicHostGrid<T>& h, hipStream_t s = nullptr) {
        HIP_CHECK(hipMemcpyAsync(cur(), h.ptr(), h.bytes(), hipMemcpyHostToDevice, s));
    }
    void upload_prev(const BasicHostGrid<T>& h, hipStream_t s = nullptr) {
        HIP_CHECK(hipMemcpyAsync(prev(), h.ptr(), h.bytes(), hipMemcpyHostToDevice, s));
    }
    void upload_vel2(const BasicHostGrid<T>& h, hipStream_t s = nullptr) {
        HIP_CHECK(hipMemcpyAsync(vel2_, h.ptr(), h.bytes(), hipMemcpyHostToDevice, s));
    }
    void download_cur(BasicHostGrid<T>& h, hipStream_t s = nullptr) const {
        HIP_CHECK(hipMemcpyAsync(h.ptr(), cur(), h.bytes(), hipMemcpyDeviceToHost, s));
    }

private:
    BasicGridDims<T> dims_;
    int idx_ = 0;
    T* buf_[3] = {};
    T* vel2_ = nullptr;
};
using DeviceGrid = BasicDeviceGrid<real_t>;

// ── Utilities (reference grid.hpp:220-229, src/utils/grid.cpp) ───────────────
/// Throws std::invalid_argument on dims <= 2R, non-positive spacing/dt, or a
/// CFL violation v_max*dt/min(h) > CFL_LIMIT when v_max > 0.
template <class T>
void validate_grid(const BasicGridDims<T>& g, T v_max = T{0});

template <class T>
void print_grid_info(const BasicGridDims<T>& g, const char* label = "");

/// RMS difference over ALL cells (reference grid.cpp:86-97).
template <class T>
double l2_error(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b);

/// L-infinity difference (reference grid.cpp:102-115).
template <class T>
double max_abs_error(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b);

/// ||a-b||_2 / ||b||_2 over interior cells, FP64 accumulation (addition).
template <class T>
double rel_l2_interior(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b);

}  // namespace stencil
