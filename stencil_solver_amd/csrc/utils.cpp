// Host utilities: validate_grid, print_grid_info, the error norms
// (reference src/utils/grid.cpp:29-115) and format_perf_table (reference
// src/utils/timer.cpp:36-49).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <sstream>
#include <stdexcept>
#include <string>

#include "stencil/grid.hpp"
#include "stencil/timer.hpp"

namespace stencil {

template <class T>
void validate_grid(const BasicGridDims<T>& g, T v_max) {
    const int R = STENCIL_RADIUS;
    std::ostringstream err;
    bool bad = false;
    auto check = [&](bool cond, const char* msg) {
        if (!cond) {
            err << "  " << msg << "\n";
            bad = true;
        }
    };
    check(g.nx > static_cast<std::size_t>(2 * R), "nx must be > 2 * STENCIL_RADIUS");
    check(g.ny > static_cast<std::size_t>(2 * R), "ny must be > 2 * STENCIL_RADIUS");
    check(g.nz > static_cast<std::size_t>(2 * R), "nz must be > 2 * STENCIL_RADIUS");
    check(g.dx > T{0}, "dx must be positive");
    check(g.dy > T{0}, "dy must be positive");
    check(g.dz > T{0}, "dz must be positive");
    check(g.dt > T{0}, "dt must be positive");
    if (!bad && v_max > T{0}) {
        const T min_h = std::min({g.dx, g.dy, g.dz});
        const T cfl = v_max * g.dt / min_h;
        if (cfl > BasicGridDims<T>::CFL_LIMIT) {
            err << "  CFL violated: v*dt/dx = " << cfl << " > " << BasicGridDims<T>::CFL_LIMIT
                << "\n";
            bad = true;
        }
    }
    if (bad) throw std::invalid_argument("Invalid grid:\n" + err.str());
}

template <class T>
void print_grid_info(const BasicGridDims<T>& g, const char* label) {
    const int R = STENCIL_RADIUS;
    std::printf("\n── Grid: %s ───────────────────────────────\n", label);
    std::printf("  Total cells   : %zu × %zu × %zu  =  %.2f M\n", g.nx, g.ny, g.nz,
                static_cast<double>(g.size()) / 1e6);
    std::printf("  Interior      : %zu × %zu × %zu  =  %.2f M\n", g.interior_nx(), g.interior_ny(),
                g.interior_nz(), static_cast<double>(g.interior_size()) / 1e6);
    std::printf("  Spacing       : dx=%.2f m  dy=%.2f m  dz=%.2f m\n", static_cast<double>(g.dx),
                static_cast<double>(g.dy), static_cast<double>(g.dz));
    std::printf("  Time step     : dt=%.4e s\n", static_cast<double>(g.dt));
    std::printf("  Stencil radius: R=%d  (%d-th order)\n", R, 2 * R);
    std::printf("  Memory/field  : %.2f MB\n", static_cast<double>(g.size() * sizeof(T)) / 1e6);
    std::printf("  Fields (3×p + vel2): %.2f MB\n\n",
                static_cast<double>(g.size() * sizeof(T) * 4) / 1e6);
}

template <class T>
double l2_error(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b) {
    if (a.dims.size() != b.dims.size()) throw std::invalid_argument("l2_error: grid size mismatch");
    double sum = 0.0;
    const std::size_t N = a.dims.size();
    for (std::size_t i = 0; i < N; ++i) {
        const double d = static_cast<double>(a.data[i]) - static_cast<double>(b.data[i]);
        sum += d * d;
    }
    return std::sqrt(sum / static_cast<double>(N));
}

template <class T>
double max_abs_error(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b) {
    if (a.dims.size() != b.dims.size())
        throw std::invalid_argument("max_abs_error: grid size mismatch");
    double m = 0.0;
    const std::size_t N = a.dims.size();
    for (std::size_t i = 0; i < N; ++i)
        m = std::max(m, std::abs(static_cast<double>(a.data[i]) - static_cast<double>(b.data[i])));
    return m;
}

template <class T>
double rel_l2_interior(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b) {
    if (a.dims.size() != b.dims.size())
        throw std::invalid_argument("rel_l2_interior: grid size mismatch");
    const auto& d = a.dims;
    const std::size_t R = STENCIL_RADIUS;
    double num = 0.0, den = 0.0;
    for (std::size_t z = R; z < d.nz - R; ++z)
        for (std::size_t y = R; y < d.ny - R; ++y)
            for (std::size_t x = R; x < d.nx - R; ++x) {
                const double va = static_cast<double>(a.at(x, y, z));
                const double vb = static_cast<double>(b.at(x, y, z));
                num += (va - vb) * (va - vb);
                den += vb * vb;
            }
    return den > 0.0 ? std::sqrt(num / den) : std::sqrt(num);
}

#define STENCIL_INSTANTIATE(T)                                                   \
    template void validate_grid<T>(const BasicGridDims<T>&, T);                  \
    template void print_grid_info<T>(const BasicGridDims<T>&, const char*);      \
    template double l2_error<T>(const BasicHostGrid<T>&, const BasicHostGrid<T>&); \
    template double max_abs_error<T>(const BasicHostGrid<T>&, const BasicHostGrid<T>&); \
    template double rel_l2_interior<T>(const BasicHostGrid<T>&, const BasicHostGrid<T>&);
STENCIL_INSTANTIATE(float)
STENCIL_INSTANTIATE(double)
#undef STENCIL_INSTANTIATE

void format_perf_table(const std::vector<PerfResult>& results) {
    if (results.empty()) return;
    std::printf("\n%-26s  %8s  %10s  %10s  %12s\n", "Kernel", "Grid", "ms/step", "GB/s", "GFLOP/s");
    std::printf("%-26s  %8s  %10s  %10s  %12s\n", std::string(26, '-').c_str(), "--------",
                "----------", "----------", "------------");
    for (const auto& r : results)
        std::printf("  %-24s  %6zu³  %8.3f ms  %8.1f     %10.1f\n", r.label.c_str(), r.grid_size,
                    r.time_ms, r.bandwidth_gbs, r.gflops);
    std::printf("\n");
}

}  // namespace stencil
