// Host utilities around the grids: argument checks of a grid, a summary
// print, the error norms the tests use, and the results table.
//
// Contract kept from the reference (src/utils/grid.cpp:29-115,
// src/utils/timer.cpp:36-49): validate_grid throws std::invalid_argument for
// an axis of <= 2R cells, a non-positive spacing or time step, or (v_max > 0)
// a CFL number v_max dt / min(h) above GridDims::CFL_LIMIT; l2_error is the
// RMS of the difference over every cell, halos included, and max_abs_error
// its largest magnitude, both in double and both throwing
// std::invalid_argument on a size mismatch.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <string>

#include "stencil/grid.hpp"
#include "stencil/timer.hpp"

namespace stencil {
namespace {

template <class T>
void same_size(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b, const char* fn) {
    if (a.dims.size() != b.dims.size() || a.data.size() != b.data.size())
        throw std::invalid_argument(std::string(fn) + ": the two grids differ in size (" +
                                    std::to_string(a.dims.size()) + " vs " + std::to_string(b.dims.size()) +
                                    " cells)");
}

// Folds f(a_i, b_i) over every cell pair in double.
template <class T, class Op, class F>
double fold_pairs(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b, double init, Op op, F f) {
    return std::inner_product(a.data.begin(), a.data.end(), b.data.begin(), init, op,
                              [&](T x, T y) { return f(static_cast<double>(x), static_cast<double>(y)); });
}

}  // namespace

template <class T>
void validate_grid(const BasicGridDims<T>& g, T v_max) {
    const std::size_t min_cells = 2 * static_cast<std::size_t>(STENCIL_RADIUS) + 1;
    std::string problems;
    auto note = [&problems](const std::string& p) { problems += (problems.empty() ? "" : "; ") + p; };
    const std::size_t n[3] = {g.nx, g.ny, g.nz};
    const T h[3] = {g.dx, g.dy, g.dz};
    const char axis[3] = {'x', 'y', 'z'};
    for (int a = 0; a < 3; ++a) {
        if (n[a] < min_cells)
            note(std::string("n") + axis[a] + " = " + std::to_string(n[a]) + " leaves no interior (need >= " +
                 std::to_string(min_cells) + ")");
        if (!(h[a] > T{0})) note(std::string("d") + axis[a] + " must be > 0");
    }
    if (!(g.dt > T{0})) note("dt must be > 0");
    if (problems.empty() && v_max > T{0}) {
        const T cfl = v_max * g.dt / std::min({g.dx, g.dy, g.dz});
        if (cfl > BasicGridDims<T>::CFL_LIMIT)
            note("CFL number " + std::to_string(static_cast<double>(cfl)) + " exceeds the limit " +
                 std::to_string(static_cast<double>(BasicGridDims<T>::CFL_LIMIT)));
    }
    if (!problems.empty()) throw std::invalid_argument("validate_grid: " + problems);
}

template <class T>
void print_grid_info(const BasicGridDims<T>& g, const char* label) {
    const double mb = static_cast<double>(g.size() * sizeof(T)) / 1e6;
    std::printf("grid %s: %zu x %zu x %zu cells (%zu x %zu x %zu interior, %.2f M points)\n",
                label ? label : "", g.nx, g.ny, g.nz, g.interior_nx(), g.interior_ny(), g.interior_nz(),
                static_cast<double>(g.interior_size()) / 1e6);
    std::printf("  h = (%g, %g, %g) m, dt = %.4e s, radius %d (order %d)\n", static_cast<double>(g.dx),
                static_cast<double>(g.dy), static_cast<double>(g.dz), static_cast<double>(g.dt), STENCIL_RADIUS,
                2 * STENCIL_RADIUS);
    std::printf("  %.2f MB per field, %.2f MB for the ring of three levels and vel2\n", mb, 4.0 * mb);
}

template <class T>
double l2_error(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b) {
    same_size(a, b, "l2_error");
    if (a.data.empty()) return 0.0;
    const double ss = fold_pairs(a, b, 0.0, std::plus<double>(), [](double x, double y) { return (x - y) * (x - y); });
    return std::sqrt(ss / static_cast<double>(a.data.size()));
}

template <class T>
double max_abs_error(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b) {
    same_size(a, b, "max_abs_error");
    return fold_pairs(a, b, 0.0, [](double m, double v) { return v > m ? v : m; },
                      [](double x, double y) { return std::fabs(x - y); });
}

template <class T>
double rel_l2_interior(const BasicHostGrid<T>& a, const BasicHostGrid<T>& b) {
    same_size(a, b, "rel_l2_interior");
    const auto& d = a.dims;
    const std::size_t R = STENCIL_RADIUS;
    double num = 0.0, den = 0.0;
    for (std::size_t z = R; z + R < d.nz; ++z)
        for (std::size_t y = R; y + R < d.ny; ++y) {
            const std::size_t row = d.idx(0, y, z);
            for (std::size_t x = R; x + R < d.nx; ++x) {
                const double va = static_cast<double>(a.data[row + x]);
                const double vb = static_cast<double>(b.data[row + x]);
                num += (va - vb) * (va - vb);
                den += vb * vb;
            }
        }
    return den > 0.0 ? std::sqrt(num / den) : std::sqrt(num);
}

template void validate_grid<float>(const BasicGridDims<float>&, float);
template void validate_grid<double>(const BasicGridDims<double>&, double);
template void print_grid_info<float>(const BasicGridDims<float>&, const char*);
template void print_grid_info<double>(const BasicGridDims<double>&, const char*);
template double l2_error<float>(const BasicHostGrid<float>&, const BasicHostGrid<float>&);
template double l2_error<double>(const BasicHostGrid<double>&, const BasicHostGrid<double>&);
template double max_abs_error<float>(const BasicHostGrid<float>&, const BasicHostGrid<float>&);
template double max_abs_error<double>(const BasicHostGrid<double>&, const BasicHostGrid<double>&);
template double rel_l2_interior<float>(const BasicHostGrid<float>&, const BasicHostGrid<float>&);
template double rel_l2_interior<double>(const BasicHostGrid<double>&, const BasicHostGrid<double>&);

void format_perf_table(const std::vector<PerfResult>& results) {
    if (results.empty()) return;
    std::size_t w = 6;  // label column width
    for (const auto& r : results) w = std::max(w, r.label.size());
    const int lw = static_cast<int>(w);
    const std::string rule(w + 46, '=');
    std::printf("\n%s\n", rule.c_str());
    std::printf("%-*s | %6s | %9s | %9s | %9s\n", lw, "kernel", "Grid", "ms/step", "GB/s", "GFLOP/s");
    std::printf("%s\n", rule.c_str());
    for (const auto& r : results) {
        const std::string grid = std::to_string(r.grid_size) + "^3";
        std::printf("%-*s | %6s | %9.3f | %9.1f | %9.1f\n", lw, r.label.c_str(), grid.c_str(), r.time_ms,
                    r.bandwidth_gbs, r.gflops);
    }
    std::printf("%s\n\n", rule.c_str());
}

}  // namespace stencil
