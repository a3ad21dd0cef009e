// oracle/ref_driver.cpp -- extern "C" shim over the reference's own CPU path.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
// reference sources where they lie (src/kernels/stencil_naive.cu in its
// STENCIL_CPU_FALLBACK form, src/solver/*.cpp, src/utils/*.cpp) into
// oracle/_ref/.  Nothing here restates the algorithm: every call goes into the
// reference code, so the outputs are the reference's outputs.  real_t is the
// reference's (float, or double with -DSTENCIL_FP64).
#include <omp.h>
#include <xmmintrin.h>

#include <chrono>
#include <cstring>
#include <memory>

#include "stencil/config.hpp"
#include "stencil/grid.hpp"
#include "stencil/kernels.cuh"
#include "stencil/solver.hpp"
#include "stencil/types.hpp"

using namespace stencil;

extern "C" {

int ref_sizeof_real() { return static_cast<int>(sizeof(real_t)); }

// Threads the reference's OpenMP loop (stencil_naive.cu:151) runs with.
int ref_omp_max_threads() { return omp_get_max_threads(); }

// Flush-to-zero and denormals-are-zero (MXCSR bits 15 and 6) on or off in
// every thread of the OpenMP pool, which later parallel regions reuse; the
// reference's CPU build leaves both off (its flags: CMakeLists.txt:109).
// Returns how many pool threads report the requested state.
int ref_set_ftz_daz(int on) {
    int ok = 0;
#pragma omp parallel reduction(+ : ok)
    {
        const unsigned bits = 0x8040u;
        const unsigned csr = _mm_getcsr();
        _mm_setcsr(on ? (csr | bits) : (csr & ~bits));
        ok += ((_mm_getcsr() & bits) == (on ? bits : 0u)) ? 1 : 0;
    }
    return ok;
}

// One launch_naive call (include/stencil/kernels.cuh:42) on host buffers.
void ref_wave_step(const real_t* cur, const real_t* prev, const real_t* vel2, real_t* next,
                   size_t nx, size_t ny, size_t nz, real_t dx, real_t dy, real_t dz, real_t dt) {
    upload_fd_coefficients();
    GridDims g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.dx = dx;
    g.dy = dy;
    g.dz = dz;
    g.dt = dt;
    launch_naive(cur, prev, vel2, next, g, nullptr);
}

// make_solver (src/solver/solver_base.cpp:119) -> n_steps of step() (or
// run(n_steps, warmup) when use_run != 0) -> download into out.  Returns the
// ms/step that run() reports (0 for the step() form).
double ref_solver(int physics, int variant, size_t nx, size_t ny, size_t nz, real_t dx,
                  real_t velocity, real_t diffusivity, real_t cx, real_t cy, real_t cz,
                  real_t sigma, int n_steps, int warmup, int use_run, real_t* out) {
    Config cfg;
    cfg.nx = nx;
    cfg.ny = ny;
    cfg.nz = nz;
    cfg.dx = cfg.dy = cfg.dz = dx;
    cfg.velocity = velocity;
    cfg.diffusivity = diffusivity;
    cfg.pulse_cx = cx;
    cfg.pulse_cy = cy;
    cfg.pulse_cz = cz;
    cfg.pulse_sigma = sigma;
    cfg.physics = physics == 0 ? Physics::WavePropagation : Physics::HeatDiffusion;
    cfg.variant = variant == 0 ? KernelVariant::Naive : KernelVariant::Tiled;
    auto s = make_solver(cfg);
    double ms = 0.0;
    if (use_run) {
        ms = s->run(n_steps, warmup).time_ms;
    } else {
        for (int i = 0; i < n_steps; ++i) s->step();
    }
    HostGrid h(s->dims());
    s->download(h);
    std::memcpy(out, h.ptr(), h.bytes());
    return ms;
}

// init_gaussian (include/stencil/grid.hpp:57).
void ref_init_gaussian(real_t* out, size_t nx, size_t ny, size_t nz, real_t cx, real_t cy,
                       real_t cz, real_t sigma) {
    GridDims g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    HostGrid h(g);
    init_gaussian(h, cx, cy, cz, sigma);
    std::memcpy(out, h.ptr(), h.bytes());
}

// init_layered_velocity (include/stencil/grid.hpp:91-99).
void ref_init_layered_velocity(real_t* out, size_t nx, size_t ny, size_t nz, real_t v_top,
                               real_t v_bot) {
    GridDims g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    HostGrid h(g);
    init_layered_velocity(h, v_top, v_bot);
    std::memcpy(out, h.ptr(), h.bytes());
}

// Config::make_grid().dt (include/stencil/config.hpp:75).
real_t ref_make_grid_dt(real_t dx, real_t dy, real_t dz, real_t velocity) {
    Config cfg;
    cfg.dx = dx;
    cfg.dy = dy;
    cfg.dz = dz;
    cfg.velocity = velocity;
    return cfg.make_grid().dt;
}

// l2_error / max_abs_error (src/utils/grid.cpp:86-115).
double ref_l2_error(const real_t* a, const real_t* b, size_t nx, size_t ny, size_t nz) {
    GridDims g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    HostGrid ha(g), hb(g);
    std::memcpy(ha.ptr(), a, ha.bytes());
    std::memcpy(hb.ptr(), b, hb.bytes());
    return l2_error(ha, hb);
}

double ref_max_abs_error(const real_t* a, const real_t* b, size_t nx, size_t ny, size_t nz) {
    GridDims g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    HostGrid ha(g), hb(g);
    std::memcpy(ha.ptr(), a, ha.bytes());
    std::memcpy(hb.ptr(), b, hb.bytes());
    return max_abs_error(ha, hb);
}

// Wall-clock ms/step of n_steps launch_naive calls over caller buffers after
// `warmup` untimed steps, ring-rotated as WavePropSolver does
// (src/solver/wave_prop.cpp:67-75); timed like SolverBase::run in CPU mode
// (src/solver/solver_base.cpp:65-108).
double ref_time_wave_steps(real_t* b0, real_t* b1, real_t* b2, const real_t* vel2, size_t nx,
                           size_t ny, size_t nz, real_t dx, real_t dt, int warmup, int n_steps) {
    upload_fd_coefficients();
    GridDims g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.dx = g.dy = g.dz = dx;
    g.dt = dt;
    real_t* buf[3] = {b0, b1, b2};
    int idx = 0;
    for (int i = 0; i < warmup; ++i) {
        launch_naive(buf[(idx + 1) % 3], buf[idx], vel2, buf[(idx + 2) % 3], g, nullptr);
        idx = (idx + 1) % 3;
    }
    CpuTimer t;
    t.start();
    for (int i = 0; i < n_steps; ++i) {
        launch_naive(buf[(idx + 1) % 3], buf[idx], vel2, buf[(idx + 2) % 3], g, nullptr);
        idx = (idx + 1) % 3;
    }
    t.stop();
    return t.elapsed_ms() / static_cast<double>(n_steps);
}

}  // extern "C"
