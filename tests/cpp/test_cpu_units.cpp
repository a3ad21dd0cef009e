// Host-side unit tests re-hosting the reference's CPU-labelled suites on this
// library's C++ API (no device needed):
//   tests/unit/test_fd_coeffs.cpp, test_grid.cpp, test_solver_cpu.cpp,
//   test_timer.cpp, test_io.cpp  (the gtest assertions, same names and
//   tolerances).
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <unistd.h>

#include "mini_test.hpp"
#include "stencil/config.hpp"
#include "stencil/io.hpp"
#include "stencil_capi.h"
#include "stencil/grid.hpp"
#include "stencil/timer.hpp"
#include "stencil/types.hpp"

using namespace stencil;

// ── test_fd_coeffs.cpp:17-62 ─────────────────────────────────────────────────
template <int R>
static void check_consistency() {
    FDCoeffs<R> fd;
    double sum = static_cast<double>(fd.c0);
    for (int r = 0; r < R; ++r) sum += 2.0 * static_cast<double>(fd.c[r]);
    EXPECT_NEAR(sum, 0.0, 1e-6) << "FDCoeffs<" << R << "> consistency";
}
TEST(FDCoeffs, ConsistencyRadius1) { check_consistency<1>(); }
TEST(FDCoeffs, ConsistencyRadius2) { check_consistency<2>(); }
TEST(FDCoeffs, ConsistencyRadius3) { check_consistency<3>(); }
TEST(FDCoeffs, ConsistencyRadius4) { check_consistency<4>(); }

TEST(FDCoeffs, KnownValuesRadius1) {
    FDCoeffs<1> fd;
    EXPECT_NEAR(static_cast<double>(fd.c0), -2.0, 1e-6);
    EXPECT_NEAR(static_cast<double>(fd.c[0]), 1.0, 1e-6);
}
TEST(FDCoeffs, KnownValuesRadius2) {
    FDCoeffs<2> fd;
    EXPECT_NEAR(static_cast<double>(fd.c0), -5.0 / 2.0, 1e-6);
    EXPECT_NEAR(static_cast<double>(fd.c[0]), 4.0 / 3.0, 1e-6);
    EXPECT_NEAR(static_cast<double>(fd.c[1]), -1.0 / 12.0, 1e-6);
}
TEST(FDCoeffs, KnownValuesRadius4) {
    FDCoeffs<4> fd;
    EXPECT_NEAR(static_cast<double>(fd.c0), -205.0 / 72.0, 1e-6);
    EXPECT_NEAR(static_cast<double>(fd.c[0]), 8.0 / 5.0, 1e-6);
    EXPECT_NEAR(static_cast<double>(fd.c[1]), -1.0 / 5.0, 1e-6);
    EXPECT_NEAR(static_cast<double>(fd.c[2]), 8.0 / 315.0, 1e-6);
    EXPECT_NEAR(static_cast<double>(fd.c[3]), -1.0 / 560.0, 1e-6);
}

// test_fd_coeffs.cpp:64-126: measured convergence order on sin(x) with the
// double rationals (float weights would swamp the truncation error), the R = 4
// case asserting > 3.8 as there.
template <int R>
struct FDCoeffsDouble;
template <>
struct FDCoeffsDouble<1> {
    static constexpr double c0 = -2.0;
    static constexpr double c[1] = {1.0};
};
template <>
struct FDCoeffsDouble<2> {
    static constexpr double c0 = -5.0 / 2.0;
    static constexpr double c[2] = {4.0 / 3.0, -1.0 / 12.0};
};
template <>
struct FDCoeffsDouble<4> {
    static constexpr double c0 = -205.0 / 72.0;
    static constexpr double c[4] = {8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
};

template <int R>
static double fd_laplacian_error(int N) {
    const double h = 2.0 * M_PI / N;
    std::vector<double> u(N);
    for (int i = 0; i < N; ++i) u[i] = std::sin(i * h);
    FDCoeffsDouble<R> fd;
    double max_err = 0.0;
    for (int i = R; i < N - R; ++i) {
        double lap = fd.c0 * u[i];
        for (int r = 1; r <= R; ++r) lap += fd.c[r - 1] * (u[i + r] + u[i - r]);
        lap /= (h * h);
        max_err = std::max(max_err, std::abs(lap + std::sin(i * h)));
    }
    return max_err;
}
TEST(FDCoeffs, ConvergenceOrder2) {
    const double rate = std::log2(fd_laplacian_error<1>(64) / fd_laplacian_error<1>(128));
    EXPECT_GT(rate, 1.8);
}
TEST(FDCoeffs, ConvergenceOrder4) {
    const double rate = std::log2(fd_laplacian_error<2>(16) / fd_laplacian_error<2>(32));
    EXPECT_GT(rate, 3.8);
}
TEST(FDCoeffs, ConvergenceOrder8) {
    const double rate = std::log2(fd_laplacian_error<4>(32) / fd_laplacian_error<4>(64));
    EXPECT_GT(rate, 3.8);
}

// ── test_grid.cpp:16-167 ─────────────────────────────────────────────────────
TEST(GridDims, SizeAndIndex) {
    GridDims g;
    g.nx = 10;
    g.ny = 20;
    g.nz = 30;
    EXPECT_EQ(g.size(), 10u * 20u * 30u);
    EXPECT_EQ(g.idx(0, 0, 0), 0u);
    EXPECT_EQ(g.idx(1, 0, 0), 1u);
    EXPECT_EQ(g.idx(0, 1, 0), 10u);
    EXPECT_EQ(g.idx(0, 0, 1), 200u);
    EXPECT_EQ(g.idx(9, 19, 29), g.size() - 1);
}
TEST(GridDims, InteriorSize) {
    GridDims g;
    g.nx = g.ny = g.nz = 64;
    const unsigned R = STENCIL_RADIUS;
    EXPECT_EQ(g.interior_nx(), 64u - 2u * R);
    EXPECT_EQ(g.interior_size(), (64u - 2u * R) * (64u - 2u * R) * (64u - 2u * R));
}
TEST(HostGrid, ConstructAndFill) {
    GridDims g;
    g.nx = g.ny = g.nz = 4;
    HostGrid h(g, real_t{3.14f});
    EXPECT_EQ(h.data.size(), 64u);
    for (auto v : h.data) EXPECT_FLOAT_EQ(v, real_t{3.14f});
}
TEST(HostGrid, AtAccessor) {
    GridDims g;
    g.nx = 4;
    g.ny = 5;
    g.nz = 6;
    HostGrid h(g);
    h.at(1, 2, 3) = real_t{99};
    EXPECT_FLOAT_EQ(h.data[g.idx(1, 2, 3)], real_t{99});
}
TEST(GridInit, GaussianPeak) {
    GridDims g;
    g.nx = g.ny = g.nz = 32;
    HostGrid h(g);
    init_gaussian(h, 0.5f, 0.5f, 0.5f, 3.0f);
    EXPECT_NEAR(static_cast<double>(h.at(16, 16, 16)), 1.0, 0.05);
    EXPECT_LT(static_cast<double>(h.at(0, 0, 0)), 0.01);
}
TEST(GridInit, GaussianHaloStaysZero) {
    GridDims g;
    g.nx = g.ny = g.nz = 20;
    HostGrid h(g);
    init_gaussian(h, 0.5f, 0.5f, 0.5f, 30.0f);
    for (std::size_t z = 0; z < g.nz; ++z)
        for (std::size_t y = 0; y < g.ny; ++y)
            for (std::size_t x = 0; x < g.nx; ++x) {
                const bool halo = x < 4 || y < 4 || z < 4 || x >= 16 || y >= 16 || z >= 16;
                if (halo) {
                    EXPECT_EQ(h.at(x, y, z), real_t{0});
                }
            }
}
TEST(GridInit, ConstantVelocity) {
    GridDims g;
    g.nx = g.ny = g.nz = 8;
    HostGrid h(g);
    init_constant_velocity(h, 1500.0f);
    for (auto v : h.data) EXPECT_FLOAT_EQ(v, 1500.0f * 1500.0f);
}
TEST(GridInit, LayeredVelocity) {
    GridDims g;
    g.nx = 4;
    g.ny = 4;
    g.nz = 8;
    HostGrid h(g);
    init_layered_velocity(h, 1000.0f, 2000.0f);
    EXPECT_FLOAT_EQ(h.at(0, 0, 0), 1000.0f * 1000.0f);
    EXPECT_FLOAT_EQ(h.at(0, 0, 7), 2000.0f * 2000.0f);
}
TEST(GridValidation, ValidGrid) {
    GridDims g;
    g.nx = g.ny = g.nz = 64;
    g.dx = g.dy = g.dz = 10.0f;
    g.dt = 0.001f;
    EXPECT_NO_THROW(validate_grid(g, 1500.0f));
}
TEST(GridValidation, TooSmall) {
    GridDims g;
    g.nx = 2;
    g.ny = g.nz = 64;
    g.dx = g.dy = g.dz = 10.0f;
    g.dt = 0.001f;
    EXPECT_THROW(validate_grid(g), std::invalid_argument);
}
TEST(GridValidation, CFLViolation) {
    GridDims g;
    g.nx = g.ny = g.nz = 64;
    g.dx = g.dy = g.dz = 10.0f;
    g.dt = 1.0f;
    EXPECT_THROW(validate_grid(g, 1500.0f), std::invalid_argument);
}
TEST(GridError, L2ErrorIdentical) {
    GridDims g;
    g.nx = g.ny = g.nz = 8;
    HostGrid a(g, real_t{1}), b(g, real_t{1});
    EXPECT_DOUBLE_EQ(l2_error(a, b), 0.0);
}
TEST(GridError, L2ErrorKnown) {
    GridDims g;
    g.nx = g.ny = g.nz = 4;
    HostGrid a(g, real_t{0}), b(g, real_t{1});
    EXPECT_NEAR(l2_error(a, b), 1.0, 1e-6);
}
TEST(GridError, MaxAbsError) {
    GridDims g;
    g.nx = g.ny = g.nz = 4;
    HostGrid a(g, real_t{0}), b(g, real_t{0});
    b.at(2, 2, 2) = real_t{5};
    EXPECT_NEAR(max_abs_error(a, b), 5.0, 1e-6);
}
TEST(GridError, RelL2InteriorKnown) {
    GridDims g;
    g.nx = g.ny = g.nz = 10;  // interior 2^3
    HostGrid a(g, real_t{0}), b(g, real_t{0});
    for (std::size_t z = 4; z < 6; ++z)
        for (std::size_t y = 4; y < 6; ++y)
            for (std::size_t x = 4; x < 6; ++x) {
                b.at(x, y, z) = real_t{2};
                a.at(x, y, z) = real_t{2} + real_t{0.5};
            }
    a.at(0, 0, 0) = real_t{100};  // halo: ignored
    EXPECT_NEAR(rel_l2_interior(a, b), 0.25, 1e-12);
}

// ── test_solver_cpu.cpp:17-127 ───────────────────────────────────────────────
TEST(Config, DefaultValues) {
    Config cfg;
    EXPECT_EQ(cfg.nx, 256u);
    EXPECT_EQ(cfg.ny, 256u);
    EXPECT_EQ(cfg.nz, 256u);
    EXPECT_EQ(cfg.steps, 100);
    EXPECT_TRUE(cfg.physics == Physics::WavePropagation);
    EXPECT_TRUE(cfg.variant == KernelVariant::Naive);
}
TEST(Config, ParseGridSize) {
    const char* argv[] = {"prog", "--nx", "512", "--ny", "512", "--nz", "512"};
    Config cfg = parse_args(7, const_cast<char**>(argv));
    EXPECT_EQ(cfg.nx, 512u);
    EXPECT_EQ(cfg.ny, 512u);
    EXPECT_EQ(cfg.nz, 512u);
}
TEST(Config, NonCubicGridThrows) {
    const char* argv[] = {"prog", "--nx", "512", "--ny", "128", "--nz", "64"};
    EXPECT_THROW(parse_args(7, const_cast<char**>(argv)), std::runtime_error);
}
TEST(Config, ParseSteps) {
    const char* argv[] = {"prog", "--steps", "200", "--warmup", "5"};
    Config cfg = parse_args(5, const_cast<char**>(argv));
    EXPECT_EQ(cfg.steps, 200);
    EXPECT_EQ(cfg.warmup_steps, 5);
}
TEST(Config, ParsePhysicsHeat) {
    const char* argv[] = {"prog", "--physics", "heat"};
    EXPECT_TRUE(parse_args(3, const_cast<char**>(argv)).physics == Physics::HeatDiffusion);
}
TEST(Config, ParseKernelTiled) {
    const char* argv[] = {"prog", "--kernel", "tiled"};
    EXPECT_TRUE(parse_args(3, const_cast<char**>(argv)).variant == KernelVariant::Tiled);
}
TEST(Config, ParseVerbose) {
    const char* argv[] = {"prog", "--verbose"};
    EXPECT_TRUE(parse_args(2, const_cast<char**>(argv)).verbose);
}
TEST(Config, ParseOutputPath) {
    const char* argv[] = {"prog", "--output", "results/out.json"};
    EXPECT_EQ(parse_args(3, const_cast<char**>(argv)).output_path, "results/out.json");
}
TEST(Config, UnknownFlagThrows) {
    const char* argv[] = {"prog", "--unknown"};
    EXPECT_THROW(parse_args(2, const_cast<char**>(argv)), std::runtime_error);
}
TEST(Config, UnknownPhysicsThrows) {
    const char* argv[] = {"prog", "--physics", "quantum"};
    EXPECT_THROW(parse_args(3, const_cast<char**>(argv)), std::runtime_error);
}
TEST(Config, MissingValueThrows) {
    const char* argv[] = {"prog", "--nx"};
    EXPECT_THROW(parse_args(2, const_cast<char**>(argv)), std::runtime_error);
}
TEST(Config, ParseExplicitDt) {
    const char* argv[] = {"prog", "--dt", "0.001", "--vmodel", "layered-noise", "--seed", "7"};
    Config cfg = parse_args(7, const_cast<char**>(argv));
    EXPECT_FLOAT_EQ(cfg.make_grid().dt, 0.001f);
    EXPECT_TRUE(cfg.vmodel == VelocityModel::LayeredNoise);
    EXPECT_EQ(cfg.seed, 7u);
}
TEST(Config, MakeGridDimensions) {
    Config cfg;
    cfg.nx = 64;
    cfg.ny = 128;
    cfg.nz = 32;
    cfg.dx = cfg.dy = cfg.dz = 5.0f;
    const GridDims g = cfg.make_grid();
    EXPECT_EQ(g.nx, 64u);
    EXPECT_EQ(g.ny, 128u);
    EXPECT_EQ(g.nz, 32u);
    EXPECT_FLOAT_EQ(g.dx, 5.0f);
}
TEST(Config, MakeGridCFLStable) {
    Config cfg;
    cfg.nx = cfg.ny = cfg.nz = 64;
    cfg.dx = cfg.dy = cfg.dz = 10.0f;
    cfg.velocity = 1500.0f;
    const GridDims g = cfg.make_grid();
    const real_t expected_dt = GridDims::CFL_LIMIT * cfg.dx / cfg.velocity;
    EXPECT_FLOAT_EQ(g.dt, expected_dt);
    EXPECT_NO_THROW(validate_grid(g, cfg.velocity));
}

// ── test_timer.cpp:15-166 ────────────────────────────────────────────────────
TEST(CpuTimer, ElapsedPositive) {
    CpuTimer t;
    t.start();
    std::this_thread::sleep_for(std::chrono::milliseconds(10));
    t.stop();
    EXPECT_GT(t.elapsed_ms(), 5.0);
    EXPECT_LT(t.elapsed_ms(), 500.0);
}
TEST(CpuTimer, ElapsedZero) {
    CpuTimer t;
    t.start();
    t.stop();
    EXPECT_GE(t.elapsed_ms(), 0.0);
    EXPECT_LT(t.elapsed_ms(), 5.0);
}
TEST(PerfResult, DefaultValues) {
    PerfResult r;
    EXPECT_EQ(r.label, "");
    EXPECT_DOUBLE_EQ(r.time_ms, 0.0);
    EXPECT_DOUBLE_EQ(r.bandwidth_gbs, 0.0);
    EXPECT_DOUBLE_EQ(r.gflops, 0.0);
}
static std::vector<PerfResult> two_rows() {
    PerfResult a, b;
    a.label = b.label = "naive";
    a.time_ms = 1.403;
    a.bandwidth_gbs = 1217.4;
    a.gflops = 608.7;
    a.grid_size = 256;
    b.time_ms = 11.462;
    b.bandwidth_gbs = 1250.9;
    b.gflops = 625.5;
    b.grid_size = 512;
    return {a, b};
}
TEST(FormatPerfTable, LabelsEachRowByGridSize) {
    const auto rows = two_rows();
    const std::string out = mt::capture_stdout([&] { format_perf_table(rows); });
    EXPECT_NE(out.find("256"), std::string::npos);
    EXPECT_NE(out.find("512"), std::string::npos);
}
TEST(FormatPerfTable, HasGridColumnNoSpeedupColumn) {
    PerfResult a;
    a.label = "tiled";
    a.time_ms = 2.0;
    a.grid_size = 128;
    const std::string out = mt::capture_stdout([&] { format_perf_table({a}); });
    EXPECT_NE(out.find("Grid"), std::string::npos);
    EXPECT_EQ(out.find("Speedup"), std::string::npos);
}
TEST(FormatPerfTable, EmptyResultsProducesNoOutput) {
    const std::string out = mt::capture_stdout([&] { format_perf_table({}); });
    EXPECT_TRUE(out.empty());
}
TEST(FormatPerfTable, DoesNotCompareDifferentSizedRows) {
    const auto rows = two_rows();
    const std::string out = mt::capture_stdout([&] { format_perf_table(rows); });
    EXPECT_NE(out.find("1.403"), std::string::npos);
    EXPECT_NE(out.find("11.462"), std::string::npos);
}
TEST(BandwidthFlops, ScalesWithGrid) {
    GridDims g1, g2;
    g1.nx = g1.ny = g1.nz = 64;
    g2.nx = g2.ny = g2.nz = 128;
    EXPECT_NEAR(bandwidth_bytes_per_step(g2) / bandwidth_bytes_per_step(g1), 8.0, 2.5);
    EXPECT_NEAR(flops_per_step(g2) / flops_per_step(g1), 8.0, 2.5);
}
TEST(BandwidthFlops, CompulsoryBytes) {
    GridDims g;
    g.nx = g.ny = g.nz = 64;
    const double n = 64.0 - 2.0 * STENCIL_RADIUS;
    EXPECT_DOUBLE_EQ(compulsory_bytes_per_step(g), n * n * n * 4.0 * sizeof(real_t));
    EXPECT_LT(compulsory_bytes_per_step(g), bandwidth_bytes_per_step(g));
}
TEST(BandwidthFlops, PositiveValues) {
    GridDims g;
    g.nx = g.ny = g.nz = 32;
    EXPECT_GT(bandwidth_bytes_per_step(g), 0.0);
    EXPECT_GT(flops_per_step(g), 0.0);
}

MT_MAIN()

// ── test_io.cpp:45-132 ───────────────────────────────────────────────────────
// The reference checks tile_x / tile_y / pencil_z against its compile-time
// macros; here they are the launch shape the tiled kernel uses on the largest
// grid of the run (cells, rows, z-chunk planes), read back through the C-ABI.
namespace {

std::string io_path() { return "/tmp/stencil_test_io_" + std::to_string(getpid()) + ".json"; }

std::string read_file(const std::string& path) {
    std::ifstream f(path);
    EXPECT_TRUE(f.good());
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::vector<PerfResult> one_result(const char* label, std::size_t n, double ms) {
    PerfResult r;
    r.label = label;
    r.grid_size = n;
    r.time_ms = ms;
    return {r};
}

}  // namespace

TEST(WriteResultsJson, RecordsTileConfigFromTheLaunchShape) {
    const std::string path = io_path();
    write_results_json(path, one_result("naive", 512, 11.5), 0, 200, 20);
    const std::string content = read_file(path);
    stencil_grid g{512, 512, 512, 1.0, 1.0, 1.0, 1.0};
    stencil_tile_shape sh{};
    EXPECT_GE(stencil_tuned_shape(&g, 0, &sh), 0);
    const int chunk = ((504 + sh.n_chunks - 1) / sh.n_chunks + 8) / 9 * 9;
    EXPECT_NE(content.find("\"tile_x\": " + std::to_string(4 * sh.tile_x4)), std::string::npos);
    EXPECT_NE(content.find("\"tile_y\": " + std::to_string(sh.tile_y)), std::string::npos);
    EXPECT_NE(content.find("\"pencil_z\": " + std::to_string(chunk < 504 ? chunk : 504)), std::string::npos);
    EXPECT_NE(content.find("\"stencil_radius\": 4"), std::string::npos);
    EXPECT_NE(content.find("\"precision\": \"fp32\""), std::string::npos);
    std::remove(path.c_str());
}

TEST(WriteResultsJson, RecordsStepsAndWarmupFromArguments) {
    const std::string path = io_path();
    write_results_json(path, one_result("tiled", 256, 2.0), 0, 200, 20);
    const std::string content = read_file(path);
    EXPECT_NE(content.find("\"steps\": 200"), std::string::npos);
    EXPECT_NE(content.find("\"warmup\": 20"), std::string::npos);
    std::remove(path.c_str());
}

TEST(WriteResultsJson, DifferentStepsWarmupProduceDifferentOutput) {
    const std::string path = io_path();
    const auto results = one_result("naive", 64, 0.1);
    write_results_json(path, results, 0, 50, 5);
    const std::string a = read_file(path);
    EXPECT_NE(a.find("\"steps\": 50"), std::string::npos);
    EXPECT_NE(a.find("\"warmup\": 5"), std::string::npos);
    write_results_json(path, results, 0, 999, 99);
    const std::string b = read_file(path);
    EXPECT_NE(b.find("\"steps\": 999"), std::string::npos);
    EXPECT_NE(b.find("\"warmup\": 99"), std::string::npos);
    EXPECT_EQ(b.find("\"steps\": 50"), std::string::npos);  // overwritten, not appended
    std::remove(path.c_str());
}

TEST(WriteResultsJson, EmptyPathIsNoOp) {
    EXPECT_NO_THROW(write_results_json("", one_result("naive", 64, 0.1), 0, 10, 1));
}

TEST(WriteResultsJson, PerResultFieldsStillPresent) {
    const std::string path = io_path();
    std::vector<PerfResult> results = one_result("tiled", 384, 6.743);
    results[0].bandwidth_gbs = 882.9;
    results[0].gflops = 441.4;
    write_results_json(path, results, 0, 200, 20);
    const std::string content = read_file(path);
    EXPECT_NE(content.find("\"label\": \"tiled\""), std::string::npos);
    EXPECT_NE(content.find("\"grid_size\": 384"), std::string::npos);
    EXPECT_NE(content.find("6.743"), std::string::npos);
    EXPECT_NE(content.find("882.9"), std::string::npos);
    EXPECT_NE(content.find("441.4"), std::string::npos);
    std::remove(path.c_str());
}

TEST(WriteScalingJson, RecordsEveryPoint) {
    const std::string path = io_path();
    std::vector<ScalingResult> rs(2);
    rs[0].mode = "weak";
    rs[0].ngpu = 1;
    rs[0].time_ms = 10.5;
    rs[0].efficiency = 1.0;
    rs[1].mode = "strong";
    rs[1].ngpu = 8;
    rs[1].time_ms = 1.75;
    rs[1].efficiency = 0.75;
    write_scaling_json(path, rs);
    const std::string content = read_file(path);
    EXPECT_NE(content.find("\"mode\": \"weak\""), std::string::npos);
    EXPECT_NE(content.find("\"mode\": \"strong\""), std::string::npos);
    EXPECT_NE(content.find("\"ngpu\": 8"), std::string::npos);
    EXPECT_NE(content.find("1.75"), std::string::npos);
    EXPECT_NE(content.find("0.75"), std::string::npos);
    std::remove(path.c_str());
}
