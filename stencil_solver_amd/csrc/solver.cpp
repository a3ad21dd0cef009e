// Solvers: SolverBase / WavePropSolver / HeatDiffusionSolver and make_solver,
// the reference's src/solver/{solver_base,wave_prop,heat_diffusion}.cpp on HIP,
// plus the C-ABI solver handles of include/stencil_capi.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "stencil/config.hpp"
#include "stencil/grid.hpp"
#include "stencil/kernels.hpp"
#include "stencil/solver.hpp"
#include "stencil/timer.hpp"
#include "stencil_capi.h"

namespace stencil {

namespace {

const char* variant_suffix(KernelVariant v) {
    switch (v) {
        case KernelVariant::Naive: return "naive";
        case KernelVariant::Tiled: return "tiled";
        case KernelVariant::MultiGPU: return "multigpu";
        default: return "unknown";
    }
}

/// Reference src/solver/solver_base.hpp:35-66 + solver_base.cpp:28-116.
template <class T>
class SolverBase : public BasicISolver<T> {
public:
    ~SolverBase() override {
        grid_.reset();
        if (stream_) (void)hipStreamDestroy(stream_);
    }

    /// Device and stream, coefficients, zeroed ring, IC into cur AND prev
    /// (quiescent start), vel2 (reference solver_base.cpp:28-62).
    void init(const BasicConfig<T>& cfg) override {
        cfg_ = cfg;
        HIP_CHECK(hipSetDevice(cfg.device));
        HIP_CHECK(hipStreamCreate(&stream_));
        upload_fd_coefficients();
        grid_ = std::make_unique<BasicDeviceGrid<T>>(cfg.make_grid());

        BasicHostGrid<T> h_field(grid_->dims());
        if (cfg.ic == InitialCondition::NoiseCube) {
            fill_noise_cube(h_field);
        } else {
            init_gaussian(h_field, cfg.pulse_cx, cfg.pulse_cy, cfg.pulse_cz, cfg.pulse_sigma);
        }
        grid_->upload_cur(h_field, stream_);
        grid_->upload_prev(h_field, stream_);

        BasicHostGrid<T> h_vel(grid_->dims());
        if (cfg.vmodel == VelocityModel::LayeredNoise) {
            fill_layered_noise(h_vel);
        } else {
            init_constant_velocity(h_vel, cfg.velocity);
        }
        grid_->upload_vel2(h_vel, stream_);
        HIP_CHECK(hipStreamSynchronize(stream_));
        if (cfg.verbose) print_grid_info(grid_->dims(), this->name().c_str());
    }

    void step() override { step_impl(); }

    /// Warmup steps (they advance the state), then n timed steps between two
    /// hipEvents on the solver stream (reference solver_base.cpp:65-108).
    PerfResult run(int n_steps, int warmup = 0) override {
        STENCIL_PUSH(this->name().c_str());
        for (int i = 0; i < warmup; ++i) step_impl();
        HIP_CHECK(hipStreamSynchronize(stream_));
        GpuTimer timer;
        timer.start(stream_);
        for (int i = 0; i < n_steps; ++i) step_impl();
        timer.stop(stream_);
        const double total_ms = static_cast<double>(timer.elapsed_ms());
        const double ms_per_step = total_ms / static_cast<double>(n_steps);
        STENCIL_POP();

        PerfResult r;
        r.label = this->name();
        r.time_ms = ms_per_step;
        r.bandwidth_gbs = bytes_per_step() / (ms_per_step * 1e-3) / 1e9;
        r.model_bandwidth_gbs =
            bandwidth_bytes_per_step(grid_->dims()) / (ms_per_step * 1e-3) / 1e9;
        r.gflops = flops_per_step(grid_->dims()) / (ms_per_step * 1e-3) / 1e9;
        r.grid_size = grid_->dims().nx;
        return r;
    }

    /// Newest time level, synchronised (reference solver_base.cpp:111-116).
    void download(BasicHostGrid<T>& out) const override {
        grid_->download_cur(out, stream_);
        HIP_CHECK(hipStreamSynchronize(stream_));
    }

    const BasicGridDims<T>& dims() const override { return grid_->dims(); }

    /// One plane of the newest level: the plane is contiguous in the
    /// [nz][ny][nx] layout, so this is a single nx*ny copy.
    void download_plane_z(std::size_t z, T* out) const override {
        const auto& d = grid_->dims();
        if (z >= d.nz) throw std::out_of_range("download_plane_z: plane " + std::to_string(z) + " outside the grid");
        const std::size_t plane = d.nx * d.ny;
        HIP_CHECK(hipMemcpyAsync(out, grid_->cur() + z * plane, plane * sizeof(T), hipMemcpyDeviceToHost, stream_));
        HIP_CHECK(hipStreamSynchronize(stream_));
    }

    hipStream_t stream() const { return stream_; }

protected:
    virtual void step_impl() = 0;
    virtual double bytes_per_step() const { return compulsory_bytes_per_step(grid_->dims()); }

    void fill_noise_cube(BasicHostGrid<T>& h) const {
        const auto& d = h.dims;
        std::vector<float> tmp(d.size());
        init_noise_cube(tmp.data(), d.nx, d.ny, d.nz, cfg_.cube_side,
                        static_cast<float>(cfg_.cube_value), cfg_.seed);
        for (std::size_t i = 0; i < tmp.size(); ++i) h.data[i] = static_cast<T>(tmp[i]);
    }

    // Computed in float and widened, so FP32 and FP64 runs see the same model.
    void fill_layered_noise(BasicHostGrid<T>& h) const {
        const auto& d = h.dims;
        std::vector<float> tmp(d.size());
        init_layered_noise_vel2(tmp.data(), d.nx, d.ny, d.nz, cfg_.n_layers,
                                static_cast<float>(cfg_.vmin), static_cast<float>(cfg_.vmax),
                                static_cast<float>(cfg_.vnoise), cfg_.seed);
        for (std::size_t i = 0; i < tmp.size(); ++i) h.data[i] = static_cast<T>(tmp[i]);
    }

    BasicConfig<T> cfg_;
    std::unique_ptr<BasicDeviceGrid<T>> grid_;
    hipStream_t stream_ = nullptr;
};

/// Reference src/solver/wave_prop.cpp:35-75: launch(variant), then rotate.
template <class T>
class WavePropSolver final : public SolverBase<T> {
public:
    std::string name() const override {
        return std::string("wave/") + variant_suffix(this->cfg_.variant);
    }

protected:
    void step_impl() override {
        auto& g = *this->grid_;
        launch(this->cfg_.variant, static_cast<const T*>(g.cur()), static_cast<const T*>(g.prev()),
               static_cast<const T*>(g.vel2()), g.next(), g.dims(), this->stream_);
        g.rotate();
    }
};

/**
 * Reference src/solver/heat_diffusion.cpp:55-122.  The reference runs the wave
 * kernel with vel2 := alpha, dims.dt := (real_t)sqrt((double)dt) and copies
 * cur -> prev after every step so that the leapfrog collapses to forward Euler.
 * Here a single-level kernel computes the same T + (dt_eff^2 alpha) Lap(T)
 * directly: 8 B/point instead of 16 B + an 8 B/point copy.  Only cur is
 * observable (download), so the ring still rotates.
 */
template <class T>
class HeatDiffusionSolver final : public SolverBase<T> {
public:
    std::string name() const override {
        return std::string("heat/") + variant_suffix(this->cfg_.variant);
    }

    void init(const BasicConfig<T>& cfg) override {
        SolverBase<T>::init(cfg);
        heat_dims_ = this->grid_->dims();
        heat_dims_.dt = static_cast<T>(std::sqrt(static_cast<double>(this->cfg_.make_grid().dt)));
        alpha_ = cfg.diffusivity;
    }

protected:
    void step_impl() override {
        auto& g = *this->grid_;
        const stencil_grid cg = detail::to_capi(heat_dims_);
        const int v = this->cfg_.variant == KernelVariant::Naive ? STENCIL_VARIANT_NAIVE
                                                                 : STENCIL_VARIANT_TILED;
        int st;
        if constexpr (std::is_same_v<T, float>) {
            st = stencil_heat_step_f32(g.cur(), g.next(), alpha_, &cg, v, this->stream_);
        } else {
            st = stencil_heat_step_f64(g.cur(), g.next(), alpha_, &cg, v, this->stream_);
        }
        detail::check(st, "heat step");
        g.rotate();
    }
    double bytes_per_step() const override {
        return 2.0 * sizeof(T) * static_cast<double>(this->grid_->dims().interior_size());
    }

private:
    BasicGridDims<T> heat_dims_{};
    T alpha_ = 0;
};

}  // namespace

template <class T>
std::unique_ptr<BasicISolver<T>> make_basic_solver(const BasicConfig<T>& cfg) {
    std::unique_ptr<BasicISolver<T>> solver;
    switch (cfg.physics) {
        case Physics::WavePropagation:
            solver = std::make_unique<WavePropSolver<T>>();
            break;
        case Physics::HeatDiffusion:
            solver = std::make_unique<HeatDiffusionSolver<T>>();
            break;
        default:
            throw std::invalid_argument("make_solver: cfg.physics is neither wave propagation nor heat diffusion");
    }
    solver->init(cfg);
    return solver;
}

template std::unique_ptr<BasicISolver<float>> make_basic_solver<float>(const BasicConfig<float>&);
template std::unique_ptr<BasicISolver<double>> make_basic_solver<double>(
    const BasicConfig<double>&);

}  // namespace stencil

// ── C-ABI solver handles ─────────────────────────────────────────────────────
struct stencil_solver {
    std::unique_ptr<stencil::BasicISolver<float>> f32;
    std::unique_ptr<stencil::BasicISolver<double>> f64;
};

namespace {

template <class T>
stencil::BasicConfig<T> to_config(const stencil_config* c) {
    stencil::BasicConfig<T> cfg;
    cfg.nx = c->nx;
    cfg.ny = c->ny;
    cfg.nz = c->nz;
    cfg.dx = static_cast<T>(c->dx);
    cfg.dy = static_cast<T>(c->dy);
    cfg.dz = static_cast<T>(c->dz);
    cfg.dt = static_cast<T>(c->dt);
    cfg.steps = c->steps;
    cfg.warmup_steps = c->warmup_steps;
    cfg.velocity = static_cast<T>(c->velocity);
    cfg.diffusivity = static_cast<T>(c->diffusivity);
    cfg.pulse_cx = static_cast<T>(c->pulse_cx);
    cfg.pulse_cy = static_cast<T>(c->pulse_cy);
    cfg.pulse_cz = static_cast<T>(c->pulse_cz);
    cfg.pulse_sigma = static_cast<T>(c->pulse_sigma);
    cfg.physics = c->physics == 1 ? stencil::Physics::HeatDiffusion
                                  : stencil::Physics::WavePropagation;
    cfg.variant = c->variant == 0   ? stencil::KernelVariant::Naive
                  : c->variant == 1 ? stencil::KernelVariant::Tiled
                                    : stencil::KernelVariant::MultiGPU;
    cfg.device = c->device;
    cfg.vmodel = c->vmodel == 1 ? stencil::VelocityModel::LayeredNoise
                                : stencil::VelocityModel::Constant;
    cfg.n_layers = c->n_layers;
    cfg.vmin = static_cast<T>(c->vmin);
    cfg.vmax = static_cast<T>(c->vmax);
    cfg.vnoise = static_cast<T>(c->vnoise);
    cfg.ic = c->ic == 1 ? stencil::InitialCondition::NoiseCube : stencil::InitialCondition::Gaussian;
    cfg.cube_side = c->cube_side;
    cfg.cube_value = static_cast<T>(c->cube_value);
    cfg.seed = c->seed;
    return cfg;
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return STENCIL_ERR_NOMEM;
    } catch (...) {
        return STENCIL_ERR_INVALID;
    }
}

}  // namespace

extern "C" {

void stencil_config_default(stencil_config* c) {
    if (!c) return;
    const stencil::BasicConfig<double> d;
    std::memset(c, 0, sizeof(*c));
    c->nx = d.nx;
    c->ny = d.ny;
    c->nz = d.nz;
    c->dx = d.dx;
    c->dy = d.dy;
    c->dz = d.dz;
    c->dt = 0.0;
    c->steps = d.steps;
    c->warmup_steps = d.warmup_steps;
    c->velocity = d.velocity;
    c->diffusivity = static_cast<double>(static_cast<float>(1.0e-4));
    c->pulse_cx = c->pulse_cy = c->pulse_cz = 0.5;
    c->pulse_sigma = 5.0;
    c->physics = 0;
    c->variant = 0;
    c->device = 0;
    c->vmodel = 0;
    c->n_layers = d.n_layers;
    c->vmin = d.vmin;
    c->vmax = d.vmax;
    c->vnoise = d.vnoise;
    c->ic = 0;
    c->cube_side = d.cube_side;
    c->cube_value = d.cube_value;
    c->seed = d.seed;
    c->fp64 = 0;
}

int stencil_solver_create(const stencil_config* cfg, stencil_solver** out) {
    if (!cfg || !out) return STENCIL_ERR_INVALID;
    if (cfg->nx <= 2 * STENCIL_RADIUS || cfg->ny <= 2 * STENCIL_RADIUS ||
        cfg->nz <= 2 * STENCIL_RADIUS)
        return STENCIL_ERR_INVALID;
    if (cfg->physics < 0 || cfg->physics > 1 || cfg->variant < 0 || cfg->variant > 2)
        return STENCIL_ERR_INVALID;
    return guarded([&] {
        auto s = std::make_unique<stencil_solver>();
        if (cfg->fp64)
            s->f64 = stencil::make_basic_solver<double>(to_config<double>(cfg));
        else
            s->f32 = stencil::make_basic_solver<float>(to_config<float>(cfg));
        *out = s.release();
        return STENCIL_OK;
    });
}

int stencil_solver_step(stencil_solver* s) {
    if (!s) return STENCIL_ERR_INVALID;
    return guarded([&] {
        if (s->f64) s->f64->step();
        else s->f32->step();
        return STENCIL_OK;
    });
}

int stencil_solver_run(stencil_solver* s, int n_steps, int warmup, stencil_perf* out) {
    if (!s || n_steps <= 0 || warmup < 0) return STENCIL_ERR_INVALID;
    return guarded([&] {
        const stencil::PerfResult r = s->f64 ? s->f64->run(n_steps, warmup) : s->f32->run(n_steps, warmup);
        if (out) {
            out->time_ms = r.time_ms;
            out->bandwidth_gbs = r.bandwidth_gbs;
            out->gflops = r.gflops;
            out->model_bandwidth_gbs = r.model_bandwidth_gbs;
            out->grid_size = r.grid_size;
            std::memset(out->label, 0, sizeof(out->label));
            std::strncpy(out->label, r.label.c_str(), sizeof(out->label) - 1);
        }
        return STENCIL_OK;
    });
}

int stencil_solver_download(stencil_solver* s, void* host, uint64_t bytes) {
    if (!s || !host) return STENCIL_ERR_INVALID;
    return guarded([&] {
        if (s->f64) {
            stencil::BasicHostGrid<double> h(s->f64->dims());
            if (bytes != h.bytes()) return static_cast<int>(STENCIL_ERR_INVALID);
            s->f64->download(h);
            std::memcpy(host, h.ptr(), h.bytes());
        } else {
            stencil::BasicHostGrid<float> h(s->f32->dims());
            if (bytes != h.bytes()) return static_cast<int>(STENCIL_ERR_INVALID);
            s->f32->download(h);
            std::memcpy(host, h.ptr(), h.bytes());
        }
        return static_cast<int>(STENCIL_OK);
    });
}

int stencil_solver_dims(const stencil_solver* s, stencil_grid* out) {
    if (!s || !out) return STENCIL_ERR_INVALID;
    *out = s->f64 ? stencil::detail::to_capi(s->f64->dims()) : stencil::detail::to_capi(s->f32->dims());
    return STENCIL_OK;
}

int stencil_solver_name(const stencil_solver* s, char* buf, uint64_t len) {
    if (!s || !buf || len == 0) return STENCIL_ERR_INVALID;
    const std::string n = s->f64 ? s->f64->name() : s->f32->name();
    std::strncpy(buf, n.c_str(), len - 1);
    buf[len - 1] = '\0';
    return STENCIL_OK;
}

void stencil_solver_destroy(stencil_solver* s) { delete s; }

}  // extern "C"
