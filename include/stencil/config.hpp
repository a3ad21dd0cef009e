#pragma once

/**
 * include/stencil/config.hpp -- runtime configuration and CLI parser.
 *
 * Same fields, defaults, flags and errors as the reference's
 * include/stencil/config.hpp:19-191.  Additions (none change the reference
 * behaviour when unused): an explicit time step `dt` (0 = derive it from the
 * CFL limit as the reference does), seeded velocity models / initial
 * conditions for the MI355X workloads, and `--dt`, `--vmodel`, `--ic`,
 * `--seed` flags.
 */

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "stencil/types.hpp"

namespace stencil {

enum class Physics {
    WavePropagation,  ///< Acoustic wave equation (default)
    HeatDiffusion,    ///< Parabolic heat equation
};

enum class KernelVariant {
    Naive,     ///< Coalesced global-memory kernel
    Tiled,     ///< LDS plane + z register queue
    MultiGPU,  ///< Z-slab decomposition; one GPU here, runs as Tiled
};

/// Velocity model for vel2 (additions; the reference fills a constant v^2).
enum class VelocityModel {
    Constant,      ///< v everywhere (reference SolverBase::init)
    LayeredNoise,  ///< n_layers z-bands, v ~ U[vmin, vmax] per band, x (1 +- noise) per cell
};

/// Initial field (additions; the reference uses the Gaussian pulse).
enum class InitialCondition {
    Gaussian,   ///< init_gaussian at pulse_c*, pulse_sigma (reference)
    NoiseCube,  ///< interior U[0,1] + centred cube of side cube_side at cube_value
};

template <class T>
struct BasicConfig {
    // Grid
    std::size_t nx = 256;
    std::size_t ny = 256;
    std::size_t nz = 256;
    T dx = static_cast<T>(10.0);
    T dy = static_cast<T>(10.0);
    T dz = static_cast<T>(10.0);
    bool ny_explicit = false;
    bool nz_explicit = false;

    // Time stepping
    int steps = 100;
    int warmup_steps = 10;
    T velocity = static_cast<T>(1500.0);
    T diffusivity = static_cast<T>(1.0e-4);
    T dt = 0;  ///< explicit time step; 0 = CFL-derived (reference make_grid)

    // Source pulse
    T pulse_cx = static_cast<T>(0.5);
    T pulse_cy = static_cast<T>(0.5);
    T pulse_cz = static_cast<T>(0.5);
    T pulse_sigma = static_cast<T>(5.0);

    // Seeded synthetic inputs
    VelocityModel vmodel = VelocityModel::Constant;
    int n_layers = 16;
    T vmin = static_cast<T>(1500.0);
    T vmax = static_cast<T>(4000.0);
    T vnoise = static_cast<T>(0.05);
    InitialCondition ic = InitialCondition::Gaussian;
    std::size_t cube_side = 64;
    T cube_value = static_cast<T>(10.0);
    std::uint64_t seed = 2;

    // Execution
    Physics physics = Physics::WavePropagation;
    KernelVariant variant = KernelVariant::Naive;
    int device = 0;
    bool verbose = false;

    std::string output_path;

    /// Reference config.hpp:75-87; an explicit dt (> 0) replaces the CFL one.
    BasicGridDims<T> make_grid() const {
        BasicGridDims<T> g;
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        g.dx = dx;
        g.dy = dy;
        g.dz = dz;
        const T min_h = std::min({dx, dy, dz});
        g.dt = dt > T{0} ? dt : BasicGridDims<T>::CFL_LIMIT * min_h / velocity;
        return g;
    }
};
using Config = BasicConfig<real_t>;

/**
 * parse_args -- reference config.hpp:108-191: same flags, same exceptions
 * (std::runtime_error on unknown flags / values and on non-cubic --ny/--nz).
 */
template <class T>
inline BasicConfig<T> parse_args_t(int argc, char** argv) {
    BasicConfig<T> cfg;
    for (int i = 1; i < argc; ++i) {
        auto eq = [&](const char* flag) { return std::strcmp(argv[i], flag) == 0; };
        auto next_int = [&]() -> std::size_t {
            if (++i >= argc) throw std::runtime_error("missing value for flag");
            return static_cast<std::size_t>(std::stoul(argv[i]));
        };
        auto next_float = [&]() -> T {
            if (++i >= argc) throw std::runtime_error("missing value for flag");
            return static_cast<T>(std::stod(argv[i]));
        };
        auto next_str = [&]() -> std::string {
            if (++i >= argc) throw std::runtime_error("missing value for flag");
            return argv[i];
        };

        if (eq("--nx"))
            cfg.nx = next_int();
        else if (eq("--ny")) {
            cfg.ny = next_int();
            cfg.ny_explicit = true;
        } else if (eq("--nz")) {
            cfg.nz = next_int();
            cfg.nz_explicit = true;
        } else if (eq("--dx"))
            cfg.dx = next_float();
        else if (eq("--dy"))
            cfg.dy = next_float();
        else if (eq("--dz"))
            cfg.dz = next_float();
        else if (eq("--steps"))
            cfg.steps = static_cast<int>(next_int());
        else if (eq("--warmup"))
            cfg.warmup_steps = static_cast<int>(next_int());
        else if (eq("--velocity"))
            cfg.velocity = next_float();
        else if (eq("--diffusivity"))
            cfg.diffusivity = next_float();
        else if (eq("--dt"))
            cfg.dt = next_float();
        else if (eq("--seed"))
            cfg.seed = static_cast<std::uint64_t>(next_int());
        else if (eq("--device"))
            cfg.device = static_cast<int>(next_int());
        else if (eq("--output"))
            cfg.output_path = next_str();
        else if (eq("--verbose"))
            cfg.verbose = true;
        else if (eq("--physics")) {
            const std::string v = next_str();
            if (v == "wave")
                cfg.physics = Physics::WavePropagation;
            else if (v == "heat")
                cfg.physics = Physics::HeatDiffusion;
            else
                throw std::runtime_error("unknown physics: " + v);
        } else if (eq("--kernel")) {
            const std::string v = next_str();
            if (v == "naive")
                cfg.variant = KernelVariant::Naive;
            else if (v == "tiled")
                cfg.variant = KernelVariant::Tiled;
            else if (v == "multigpu")
                cfg.variant = KernelVariant::MultiGPU;
            else
                throw std::runtime_error("unknown kernel: " + v);
        } else if (eq("--vmodel")) {
            const std::string v = next_str();
            if (v == "constant")
                cfg.vmodel = VelocityModel::Constant;
            else if (v == "layered-noise")
                cfg.vmodel = VelocityModel::LayeredNoise;
            else
                throw std::runtime_error("unknown velocity model: " + v);
        } else if (eq("--ic")) {
            const std::string v = next_str();
            if (v == "gaussian")
                cfg.ic = InitialCondition::Gaussian;
            else if (v == "noise-cube")
                cfg.ic = InitialCondition::NoiseCube;
            else
                throw std::runtime_error("unknown initial condition: " + v);
        } else {
            throw std::runtime_error(std::string("unknown flag: ") + argv[i]);
        }
    }

    if ((cfg.ny_explicit || cfg.nz_explicit) && (cfg.nx != cfg.ny || cfg.ny != cfg.nz)) {
        throw std::runtime_error("Non-cubic grids aren't supported yet (got --nx " +
                                 std::to_string(cfg.nx) + " --ny " + std::to_string(cfg.ny) +
                                 " --nz " + std::to_string(cfg.nz) +
                                 ") -- pass --nx alone for an N×N×N cube, or make --ny/--nz "
                                 "match it.");
    }
    return cfg;
}

inline Config parse_args(int argc, char** argv) { return parse_args_t<real_t>(argc, argv); }

}  // namespace stencil
