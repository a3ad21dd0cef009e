// Device tests re-hosting the reference's kernel / solver / physics suites on
// this library's C++ API and the gfx950 kernels (same configs, same
// tolerances):
//   tests/unit/test_kernels.cpp, tests/unit/test_solver.cpp,
//   tests/integration/test_kernel_correctness.cpp, test_wave_3d.cpp,
//   test_heat_3d.cpp.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>

#include "mini_test.hpp"
#include "stencil/config.hpp"
#include "stencil/grid.hpp"
#include "stencil/kernels.hpp"
#include "stencil/solver.hpp"
#include "stencil/types.hpp"

using namespace stencil;

namespace {

struct DevBuf {
    real_t* ptr = nullptr;
    explicit DevBuf(std::size_t n) { HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(real_t))); }
    ~DevBuf() {
        if (ptr) (void)hipFree(ptr);
    }
    void upload(const real_t* h, std::size_t n) {
        HIP_CHECK(hipMemcpy(ptr, h, n * sizeof(real_t), hipMemcpyHostToDevice));
    }
    void download(real_t* h, std::size_t n) const {
        HIP_CHECK(hipMemcpy(h, ptr, n * sizeof(real_t), hipMemcpyDeviceToHost));
    }
};

GridDims make_dims(std::size_t N = 24) {
    GridDims d;
    d.nx = d.ny = d.nz = N;
    d.dx = d.dy = d.dz = real_t{10.0f};
    d.dt = real_t{0.001f};
    return d;
}

template <class Launch>
void run_kernel(Launch launch_fn, const HostGrid& cur, const HostGrid& prev, const HostGrid& vel,
                HostGrid& nxt) {
    const std::size_t N = cur.dims.size();
    DevBuf d_cur(N), d_prev(N), d_vel(N), d_nxt(N);
    d_cur.upload(cur.ptr(), N);
    d_prev.upload(prev.ptr(), N);
    d_vel.upload(vel.ptr(), N);
    d_nxt.upload(nxt.ptr(), N);
    launch_fn(d_cur.ptr, d_prev.ptr, d_vel.ptr, d_nxt.ptr, cur.dims, nullptr);
    HIP_CHECK(hipDeviceSynchronize());
    d_nxt.download(nxt.data.data(), N);
}

void run_naive(const HostGrid& c, const HostGrid& p, const HostGrid& v, HostGrid& n) {
    run_kernel([](auto... a) { launch_naive(a...); }, c, p, v, n);
}
void run_tiled(const HostGrid& c, const HostGrid& p, const HostGrid& v, HostGrid& n) {
    run_kernel([](auto... a) { launch_tiled(a...); }, c, p, v, n);
}

// tests/integration/test_kernel_correctness.cpp:43-75, serial.
void cpu_reference_step(const HostGrid& cur, const HostGrid& prev, const HostGrid& vel2,
                        HostGrid& next) {
    const auto& d = cur.dims;
    const int R = STENCIL_RADIUS;
    FDCoeffs<STENCIL_RADIUS> fd;
    const real_t inv_dx2 = real_t{1} / (d.dx * d.dx);
    const real_t inv_dy2 = real_t{1} / (d.dy * d.dy);
    const real_t inv_dz2 = real_t{1} / (d.dz * d.dz);
    const real_t dt2 = d.dt * d.dt;
    for (int iz = R; iz < int(d.nz) - R; ++iz)
        for (int iy = R; iy < int(d.ny) - R; ++iy)
            for (int ix = R; ix < int(d.nx) - R; ++ix) {
                const std::size_t c = d.idx(ix, iy, iz);
                const std::size_t sy = d.nx, sz = d.nx * d.ny;
                real_t lap_x = fd.c0 * cur.data[c] * inv_dx2;
                real_t lap_y = fd.c0 * cur.data[c] * inv_dy2;
                real_t lap_z = fd.c0 * cur.data[c] * inv_dz2;
                for (int r = 1; r <= R; ++r) {
                    lap_x += fd.c[r - 1] * (cur.data[c + r] + cur.data[c - r]) * inv_dx2;
                    lap_y += fd.c[r - 1] * (cur.data[c + r * sy] + cur.data[c - r * sy]) * inv_dy2;
                    lap_z += fd.c[r - 1] * (cur.data[c + r * sz] + cur.data[c - r * sz]) * inv_dz2;
                }
                next.data[c] = real_t{2} * cur.data[c] - prev.data[c] +
                               dt2 * vel2.data[c] * (lap_x + lap_y + lap_z);
            }
}

Config small_wave_cfg() {
    Config cfg;
    cfg.nx = cfg.ny = cfg.nz = 32;
    cfg.steps = 10;
    cfg.warmup_steps = 2;
    cfg.physics = Physics::WavePropagation;
    cfg.variant = KernelVariant::Naive;
    cfg.velocity = 1500.0f;
    return cfg;
}
Config small_heat_cfg() {
    Config cfg = small_wave_cfg();
    cfg.physics = Physics::HeatDiffusion;
    cfg.diffusivity = 1.0f;
    return cfg;
}
Config correctness_cfg(std::size_t N = 32) {
    Config cfg;
    cfg.nx = cfg.ny = cfg.nz = N;
    cfg.steps = 1;
    cfg.warmup_steps = 0;
    cfg.physics = Physics::WavePropagation;
    cfg.velocity = 1500.0f;
    cfg.dx = cfg.dy = cfg.dz = 10.0f;
    cfg.pulse_cx = cfg.pulse_cy = cfg.pulse_cz = 0.5f;
    cfg.pulse_sigma = 4.0f;
    return cfg;
}
Config wave3d_cfg(std::size_t N = 64) {
    Config cfg = correctness_cfg(N);
    cfg.steps = 20;
    cfg.variant = KernelVariant::Naive;
    return cfg;
}
Config heat3d_cfg(std::size_t N = 48) {
    Config cfg;
    cfg.nx = cfg.ny = cfg.nz = N;
    cfg.steps = 20;
    cfg.warmup_steps = 0;
    cfg.physics = Physics::HeatDiffusion;
    cfg.variant = KernelVariant::Naive;
    cfg.diffusivity = 1.0e-3f;
    cfg.velocity = 500.0f;
    cfg.dx = cfg.dy = cfg.dz = 1.0f;
    cfg.pulse_cx = cfg.pulse_cy = cfg.pulse_cz = 0.5f;
    cfg.pulse_sigma = 5.0f;
    return cfg;
}

double peak(const HostGrid& g) { return double(*std::max_element(g.data.begin(), g.data.end())); }
double total(const HostGrid& g) {
    double s = 0;
    for (auto v : g.data) s += double(v);
    return s;
}

}  // namespace

// ── tests/unit/test_kernels.cpp ──────────────────────────────────────────────
TEST(KernelTest, UploadCoefficientsDoesNotThrow) { EXPECT_NO_THROW(upload_fd_coefficients()); }

TEST(KernelTest, NaiveFlatFieldFixedPoint) {
    const GridDims d = make_dims(24);
    HostGrid cur(d, real_t{2.0f}), prev(d, real_t{2.0f}), vel(d, real_t{1.0f}), nxt(d);
    run_naive(cur, prev, vel, nxt);
    const std::size_t R = STENCIL_RADIUS;
    for (std::size_t iz = R; iz < d.nz - R; ++iz)
        for (std::size_t iy = R; iy < d.ny - R; ++iy)
            for (std::size_t ix = R; ix < d.nx - R; ++ix)
                EXPECT_NEAR(double(nxt.at(ix, iy, iz)), 2.0, 1e-5);
}

TEST(KernelTest, NaiveLinearFieldZeroLaplacian) {
    const GridDims d = make_dims(24);
    HostGrid cur(d), prev(d), vel(d, real_t{1.0f}), nxt(d);
    for (std::size_t iz = 0; iz < d.nz; ++iz)
        for (std::size_t iy = 0; iy < d.ny; ++iy)
            for (std::size_t ix = 0; ix < d.nx; ++ix)
                cur.at(ix, iy, iz) = prev.at(ix, iy, iz) =
                    real_t{0.1f} * ix + real_t{0.2f} * iy + real_t{0.3f} * iz;
    run_naive(cur, prev, vel, nxt);
    const std::size_t R = STENCIL_RADIUS;
    for (std::size_t iz = R; iz < d.nz - R; ++iz)
        for (std::size_t iy = R; iy < d.ny - R; ++iy)
            for (std::size_t ix = R; ix < d.nx - R; ++ix) {
                const double e = double(cur.at(ix, iy, iz));
                EXPECT_NEAR(double(nxt.at(ix, iy, iz)), e, std::abs(e) * 1e-4 + 1e-6);
            }
}

TEST(KernelTest, NaiveGaussianMatchesSolver) {
    Config cfg;
    cfg.nx = cfg.ny = cfg.nz = 32;
    cfg.variant = KernelVariant::Naive;
    cfg.velocity = 1500.0f;
    auto solver = make_solver(cfg);
    solver->step();
    HostGrid out(solver->dims());
    solver->download(out);
    EXPECT_GT(peak(out), 0.0);
    EXPECT_LT(peak(out), 2.0);
}

TEST(KernelTest, TiledFlatFieldFixedPoint) {
    const GridDims d = make_dims(32);
    HostGrid cur(d, real_t{3.0f}), prev(d, real_t{3.0f}), vel(d, real_t{1.0f}), nxt(d);
    run_tiled(cur, prev, vel, nxt);
    const std::size_t R = STENCIL_RADIUS;
    for (std::size_t iz = R; iz < d.nz - R; ++iz)
        for (std::size_t iy = R; iy < d.ny - R; ++iy)
            for (std::size_t ix = R; ix < d.nx - R; ++ix)
                EXPECT_NEAR(double(nxt.at(ix, iy, iz)), 3.0, 1e-5);
}

TEST(KernelTest, TiledAgreesWithNaiveSingleStep) {
    Config cfg;
    cfg.nx = cfg.ny = cfg.nz = 32;
    cfg.velocity = 1500.0f;
    cfg.variant = KernelVariant::Naive;
    auto a = make_solver(cfg);
    a->step();
    HostGrid fa(a->dims());
    a->download(fa);
    cfg.variant = KernelVariant::Tiled;
    auto b = make_solver(cfg);
    b->step();
    HostGrid fb(b->dims());
    b->download(fb);
    EXPECT_LT(max_abs_error(fa, fb), 1e-4);
}

TEST(KernelTest, NaiveHaloCellsRemainZero) {
    const GridDims d = make_dims(24);
    HostGrid cur(d), prev(d), vel(d, real_t{1.0f}), nxt(d);
    init_gaussian(cur, 0.5f, 0.5f, 0.5f, 3.0f);
    init_gaussian(prev, 0.5f, 0.5f, 0.5f, 3.0f);
    run_naive(cur, prev, vel, nxt);
    for (std::size_t iz = 0; iz < d.nz; ++iz)
        for (std::size_t iy = 0; iy < d.ny; ++iy)
            for (std::size_t ix = 0; ix < STENCIL_RADIUS; ++ix) EXPECT_FLOAT_EQ(nxt.at(ix, iy, iz), real_t{0});
}

TEST(KernelTest, NaiveDeterministic) {
    Config cfg;
    cfg.nx = cfg.ny = cfg.nz = 24;
    cfg.variant = KernelVariant::Naive;
    cfg.velocity = 1500.0f;
    cfg.steps = 3;
    auto s1 = make_solver(cfg);
    s1->run(cfg.steps);
    HostGrid f1(s1->dims());
    s1->download(f1);
    auto s2 = make_solver(cfg);
    s2->run(cfg.steps);
    HostGrid f2(s2->dims());
    s2->download(f2);
    EXPECT_DOUBLE_EQ(l2_error(f1, f2), 0.0);
}

// ── tests/unit/test_solver.cpp ───────────────────────────────────────────────
TEST(SolverFactory, CreatesWaveSolver) {
    auto s = make_solver(small_wave_cfg());
    ASSERT_TRUE(s != nullptr);
    EXPECT_NE(s->name().find("wave"), std::string::npos);
}
TEST(SolverFactory, CreatesHeatSolver) {
    auto s = make_solver(small_heat_cfg());
    ASSERT_TRUE(s != nullptr);
    EXPECT_NE(s->name().find("heat"), std::string::npos);
}
TEST(SolverFactory, DimsMatchConfig) {
    const Config cfg = small_wave_cfg();
    auto s = make_solver(cfg);
    EXPECT_EQ(s->dims().nx, cfg.nx);
    EXPECT_EQ(s->dims().ny, cfg.ny);
    EXPECT_EQ(s->dims().nz, cfg.nz);
}
TEST(SolverName, WaveNaive) {
    Config cfg = small_wave_cfg();
    cfg.variant = KernelVariant::Naive;
    EXPECT_EQ(make_solver(cfg)->name(), "wave/naive");
}
TEST(SolverName, WaveTiled) {
    Config cfg = small_wave_cfg();
    cfg.variant = KernelVariant::Tiled;
    EXPECT_EQ(make_solver(cfg)->name(), "wave/tiled");
}
TEST(SolverName, HeatNaive) {
    Config cfg = small_heat_cfg();
    cfg.variant = KernelVariant::Naive;
    EXPECT_EQ(make_solver(cfg)->name(), "heat/naive");
}
TEST(WaveSolver, StepChangesField) {
    auto s = make_solver(small_wave_cfg());
    HostGrid before(s->dims());
    s->download(before);
    s->step();
    HostGrid after(s->dims());
    s->download(after);
    EXPECT_GT(max_abs_error(before, after), 0.0);
}
TEST(HeatSolver, StepChangesField) {
    auto s = make_solver(small_heat_cfg());
    HostGrid before(s->dims());
    s->download(before);
    s->step();
    HostGrid after(s->dims());
    s->download(after);
    EXPECT_GT(max_abs_error(before, after), 0.0);
}
TEST(WaveSolver, PreservesXSymmetry) {
    auto s = make_solver(small_wave_cfg());
    for (int i = 0; i < 2; ++i) s->step();
    HostGrid f(s->dims());
    s->download(f);
    const auto& d = s->dims();
    double m = 0;
    for (std::size_t iz = STENCIL_RADIUS; iz < d.nz - STENCIL_RADIUS; ++iz)
        for (std::size_t iy = STENCIL_RADIUS; iy < d.ny - STENCIL_RADIUS; ++iy)
            for (std::size_t ix = STENCIL_RADIUS; ix < d.nx / 2; ++ix)
                m = std::max(m, std::abs(double(f.at(ix, iy, iz)) - double(f.at(d.nx - 1 - ix, iy, iz))));
    EXPECT_LT(m, 0.5);
}
TEST(SolverRun, ReturnsPositiveMetrics) {
    Config cfg = small_wave_cfg();
    cfg.steps = 5;
    cfg.warmup_steps = 1;
    auto s = make_solver(cfg);
    const auto r = s->run(cfg.steps, cfg.warmup_steps);
    EXPECT_GT(r.time_ms, 0.0);
    EXPECT_GT(r.bandwidth_gbs, 0.0);
    EXPECT_GT(r.gflops, 0.0);
    EXPECT_EQ(r.grid_size, cfg.nx);
}
TEST(SolverRun, NameMatchesVariant) {
    Config cfg = small_wave_cfg();
    auto s = make_solver(cfg);
    EXPECT_EQ(s->run(2).label, s->name());
}
TEST(HeatSolver, PeakDecreases) {
    auto s = make_solver(small_heat_cfg());
    HostGrid f(s->dims());
    s->download(f);
    const double p0 = peak(f);
    for (int i = 0; i < 100; ++i) s->step();
    s->download(f);
    EXPECT_LT(peak(f), p0);
}

// ── tests/integration/test_kernel_correctness.cpp ────────────────────────────
TEST(KernelCorrectness, NaiveMatchesCPUReference) {
    Config cfg = correctness_cfg(32);
    cfg.variant = KernelVariant::Naive;
    auto s = make_solver(cfg);
    s->step();
    HostGrid gpu(s->dims());
    s->download(gpu);
    const GridDims d = cfg.make_grid();
    HostGrid h_cur(d), h_prev(d), h_vel(d), cpu(d);
    init_gaussian(h_cur, cfg.pulse_cx, cfg.pulse_cy, cfg.pulse_cz, cfg.pulse_sigma);
    init_gaussian(h_prev, cfg.pulse_cx, cfg.pulse_cy, cfg.pulse_cz, cfg.pulse_sigma);
    init_constant_velocity(h_vel, cfg.velocity);
    cpu_reference_step(h_cur, h_prev, h_vel, cpu);
    double m = 0;
    for (std::size_t iz = STENCIL_RADIUS; iz < d.nz - STENCIL_RADIUS; ++iz)
        for (std::size_t iy = STENCIL_RADIUS; iy < d.ny - STENCIL_RADIUS; ++iy)
            for (std::size_t ix = STENCIL_RADIUS; ix < d.nx - STENCIL_RADIUS; ++ix)
                m = std::max(m, std::abs(double(gpu.at(ix, iy, iz)) - double(cpu.at(ix, iy, iz))));
    EXPECT_LT(m, 1e-4);
}
TEST(KernelCorrectness, NaiveAndTiledAgreeSingleStep) {
    Config cfg = correctness_cfg(32);
    cfg.variant = KernelVariant::Naive;
    auto a = make_solver(cfg);
    a->step();
    HostGrid fa(a->dims());
    a->download(fa);
    cfg.variant = KernelVariant::Tiled;
    auto b = make_solver(cfg);
    b->step();
    HostGrid fb(b->dims());
    b->download(fb);
    EXPECT_LT(max_abs_error(fa, fb), 1e-5);
}
TEST(KernelCorrectness, NaiveAndTiledAgreeMultiStep) {
    Config cfg = correctness_cfg(32);
    cfg.variant = KernelVariant::Naive;
    auto a = make_solver(cfg);
    cfg.variant = KernelVariant::Tiled;
    auto b = make_solver(cfg);
    for (int i = 0; i < 20; ++i) {
        a->step();
        b->step();
    }
    HostGrid fa(a->dims()), fb(b->dims());
    a->download(fa);
    b->download(fb);
    EXPECT_LT(l2_error(fa, fb), 1e-4);
}
TEST(KernelCorrectness, FlatFieldIsFixedPoint) {
    const GridDims d = correctness_cfg(24).make_grid();
    HostGrid c(d, real_t{1.0f}), p(d, real_t{1.0f}), v(d, real_t{1.0f}), n(d);
    run_tiled(c, p, v, n);
    for (std::size_t iz = STENCIL_RADIUS; iz < d.nz - STENCIL_RADIUS; ++iz)
        for (std::size_t iy = STENCIL_RADIUS; iy < d.ny - STENCIL_RADIUS; ++iy)
            for (std::size_t ix = STENCIL_RADIUS; ix < d.nx - STENCIL_RADIUS; ++ix)
                EXPECT_NEAR(double(n.at(ix, iy, iz)), 1.0, 1e-5);
}
TEST(KernelCorrectness, LinearFieldZeroLaplacian) {
    const GridDims d = correctness_cfg(24).make_grid();
    HostGrid c(d), p(d), v(d, real_t{1.0f}), n(d);
    for (std::size_t iz = 0; iz < d.nz; ++iz)
        for (std::size_t iy = 0; iy < d.ny; ++iy)
            for (std::size_t ix = 0; ix < d.nx; ++ix)
                c.at(ix, iy, iz) = p.at(ix, iy, iz) = real_t{0.1f} * ix + real_t{0.2f} * iy + real_t{0.3f} * iz;
    run_tiled(c, p, v, n);
    for (std::size_t iz = STENCIL_RADIUS; iz < d.nz - STENCIL_RADIUS; ++iz)
        for (std::size_t iy = STENCIL_RADIUS; iy < d.ny - STENCIL_RADIUS; ++iy)
            for (std::size_t ix = STENCIL_RADIUS; ix < d.nx - STENCIL_RADIUS; ++ix) {
                const double e = double(c.at(ix, iy, iz));
                EXPECT_NEAR(double(n.at(ix, iy, iz)), e, std::abs(e) * 1e-4 + 1e-6);
            }
}

// ── tests/integration/test_wave_3d.cpp ───────────────────────────────────────
static double energy(const HostGrid& a, const HostGrid& b) {
    double e = 0;
    for (std::size_t i = 0; i < a.dims.size(); ++i)
        e += double(a.data[i]) * double(a.data[i]) + double(b.data[i]) * double(b.data[i]);
    return 0.5 * e;
}
TEST(WaveSolver3DTest, EnergyDoesNotGrow) {
    for (auto v : {KernelVariant::Naive, KernelVariant::Tiled}) {
        Config cfg = wave3d_cfg(48);
        cfg.variant = v;
        cfg.steps = 50;
        auto s = make_solver(cfg);
        HostGrid cur(s->dims());
        s->download(cur);
        const double E0 = energy(cur, cur);
        ASSERT_GT(E0, 0.0);
        s->run(cfg.steps);
        s->download(cur);
        const double E1 = energy(cur, cur);
        EXPECT_LT(E1, E0 * 10.0);
        EXPECT_GT(E1, 0.0);
    }
}
TEST(WaveSolver3DTest, PulseSpreadsOutward) {
    Config cfg = wave3d_cfg(64);
    cfg.steps = 10;
    auto s = make_solver(cfg);
    HostGrid f0(s->dims());
    s->download(f0);
    for (int i = 0; i < cfg.steps; ++i) s->step();
    HostGrid f1(s->dims());
    s->download(f1);
    EXPECT_LT(peak(f1), peak(f0));
}
TEST(WaveSolver3DTest, SymmetryPreservedAllAxes) {
    for (auto v : {KernelVariant::Naive, KernelVariant::Tiled}) {
        Config cfg = wave3d_cfg(48);
        cfg.variant = v;
        cfg.steps = 8;
        auto s = make_solver(cfg);
        for (int i = 0; i < cfg.steps; ++i) s->step();
        HostGrid f(s->dims());
        s->download(f);
        const auto& d = s->dims();
        double mx = 0, my = 0, mz = 0;
        for (std::size_t iz = STENCIL_RADIUS; iz < d.nz - STENCIL_RADIUS; ++iz)
            for (std::size_t iy = STENCIL_RADIUS; iy < d.ny - STENCIL_RADIUS; ++iy)
                for (std::size_t ix = STENCIL_RADIUS; ix < d.nx / 2; ++ix) {
                    mx = std::max(mx, std::abs(double(f.at(ix, iy, iz)) - double(f.at(d.nx - 1 - ix, iy, iz))));
                    my = std::max(my, std::abs(double(f.at(iy, ix, iz)) - double(f.at(d.ny - 1 - iy, ix, iz))));
                    mz = std::max(mz, std::abs(double(f.at(ix, iy, iz)) - double(f.at(ix, iy, d.nz - 1 - iz))));
                }
        EXPECT_LT(mx, 1e-4);
        EXPECT_LT(my, 1e-4);
        EXPECT_LT(mz, 1e-4);
    }
}
TEST(WaveSolver3DTest, HaloCellsRemainZero) {
    for (auto v : {KernelVariant::Naive, KernelVariant::Tiled}) {
        Config cfg = wave3d_cfg(48);
        cfg.variant = v;
        cfg.steps = 10;
        auto s = make_solver(cfg);
        s->run(cfg.steps);
        HostGrid f(s->dims());
        s->download(f);
        const auto& d = s->dims();
        for (std::size_t iz = 0; iz < d.nz; ++iz)
            for (std::size_t iy = 0; iy < d.ny; ++iy)
                for (std::size_t ix = 0; ix < d.nx; ++ix) {
                    const bool halo = ix < 4 || iy < 4 || iz < 4 || ix >= d.nx - 4 || iy >= d.ny - 4 ||
                                      iz >= d.nz - 4;
                    if (halo) {
                        EXPECT_FLOAT_EQ(f.at(ix, iy, iz), real_t{0});
                    }
                }
    }
}
TEST(WaveSolver3DTest, BitIdenticalReproducibility) {
    for (auto v : {KernelVariant::Naive, KernelVariant::Tiled}) {
        Config cfg = wave3d_cfg(32);
        cfg.variant = v;
        cfg.steps = 5;
        auto s1 = make_solver(cfg);
        auto s2 = make_solver(cfg);
        s1->run(cfg.steps);
        s2->run(cfg.steps);
        HostGrid f1(s1->dims()), f2(s2->dims());
        s1->download(f1);
        s2->download(f2);
        EXPECT_DOUBLE_EQ(l2_error(f1, f2), 0.0);
    }
}

// ── tests/integration/test_heat_3d.cpp ───────────────────────────────────────
TEST(HeatSolver3DTest, PeakDecreaseMonotonically) {
    for (auto v : {KernelVariant::Naive, KernelVariant::Tiled}) {
        Config cfg = heat3d_cfg();
        cfg.variant = v;
        auto s = make_solver(cfg);
        HostGrid f(s->dims());
        s->download(f);
        double prev = peak(f);
        for (int i = 0; i < 10; ++i) {
            s->step();
            s->download(f);
            EXPECT_LE(peak(f), prev + 1e-5);
            prev = peak(f);
        }
    }
}
TEST(HeatSolver3DTest, TotalHeatNeverIncreases) {
    for (auto v : {KernelVariant::Naive, KernelVariant::Tiled}) {
        Config cfg = heat3d_cfg();
        cfg.variant = v;
        auto s = make_solver(cfg);
        HostGrid f(s->dims());
        s->download(f);
        double prev = total(f);
        for (int i = 0; i < cfg.steps; ++i) {
            s->step();
            s->download(f);
            EXPECT_LE(total(f), prev + 1e-3);
            prev = total(f);
        }
    }
}
TEST(HeatSolver3DTest, SolutionNonNegative) {
    Config cfg = heat3d_cfg();
    cfg.steps = 30;
    auto s = make_solver(cfg);
    s->run(cfg.steps);
    HostGrid f(s->dims());
    s->download(f);
    const auto& d = s->dims();
    for (std::size_t iz = STENCIL_RADIUS; iz < d.nz - STENCIL_RADIUS; ++iz)
        for (std::size_t iy = STENCIL_RADIUS; iy < d.ny - STENCIL_RADIUS; ++iy)
            for (std::size_t ix = STENCIL_RADIUS; ix < d.nx - STENCIL_RADIUS; ++ix)
                EXPECT_GE(double(f.at(ix, iy, iz)), -1e-6);
}
TEST(HeatSolver3DTest, GaussianWidthGrowsAnalytically) {
    Config cfg = heat3d_cfg(64);
    cfg.steps = 40;
    cfg.diffusivity = 1.0e-2f;
    cfg.pulse_sigma = 6.0f;
    auto s = make_solver(cfg);
    const GridDims& d = s->dims();
    auto sigma2 = [&](const HostGrid& g) {
        double tot = 0, w = 0;
        const double cx = d.nx * 0.5, cy = d.ny * 0.5, cz = d.nz * 0.5;
        for (std::size_t iz = 0; iz < d.nz; ++iz)
            for (std::size_t iy = 0; iy < d.ny; ++iy)
                for (std::size_t ix = 0; ix < d.nx; ++ix) {
                    const double v = double(g.at(ix, iy, iz));
                    const double r2 = (ix - cx) * (ix - cx) + (iy - cy) * (iy - cy) + (iz - cz) * (iz - cz);
                    tot += v;
                    w += v * r2;
                }
        return tot > 1e-12 ? w / tot : 0.0;
    };
    HostGrid f(d);
    s->download(f);
    const double s0 = sigma2(f);
    s->run(cfg.steps);
    s->download(f);
    const double st = sigma2(f);
    const double t = double(cfg.steps) * double(d.dt);
    const double expected = s0 + 6.0 * double(cfg.diffusivity) * t;
    EXPECT_NEAR(st, expected, expected * 0.20);
}
TEST(HeatSolver3DTest, SymmetryPreserved) {
    Config cfg = heat3d_cfg(48);
    cfg.steps = 10;
    auto s = make_solver(cfg);
    s->run(cfg.steps);
    HostGrid f(s->dims());
    s->download(f);
    const auto& d = s->dims();
    double m = 0;
    for (std::size_t iz = STENCIL_RADIUS; iz < d.nz - STENCIL_RADIUS; ++iz)
        for (std::size_t iy = STENCIL_RADIUS; iy < d.ny - STENCIL_RADIUS; ++iy)
            for (std::size_t ix = STENCIL_RADIUS; ix < d.nx / 2; ++ix)
                m = std::max(m, std::abs(double(f.at(ix, iy, iz)) - double(f.at(d.nx - 1 - ix, iy, iz))));
    EXPECT_LT(m, 1e-5);
}

MT_MAIN()
