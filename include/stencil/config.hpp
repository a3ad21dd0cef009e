#pragma once

/**
 * include/stencil/config.hpp -- runtime configuration and CLI parser.
 *
 * The Config fields, their defaults and the command-line flags are the
 * reference's interface (include/stencil/config.hpp:19-191), so callers and
 * scripts written against it work unchanged: the grid (--nx/--ny/--nz,
 * --dx/--dy/--dz; --ny/--nz must then equal --nx), --steps, --warmup,
 * --velocity, --diffusivity, --physics wave|heat, --kernel
 * naive|tiled|multigpu, --device, --output, --verbose.  parse_args throws
 * std::runtime_error for an unknown flag, a missing or malformed value, an
 * unknown physics / kernel name and a non-cubic --ny/--nz.  Numbers are
 * parsed strictly (the whole argument, non-negative integers for counts).
 * Additions (none change the reference behaviour when unused): an explicit
 * time step `dt` (0 = derive it from the CFL limit as the reference does),
 * seeded velocity models / initial conditions for the MI355X workloads, and
 * the --dt, --vmodel, --ic, --seed flags.
 */

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <utility>

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

namespace detail {

inline std::runtime_error bad_value(const std::string& flag, const std::string& v, const char* what) {
    return std::runtime_error("parse_args: " + flag + " expects " + what + ", got '" + v + "'");
}

/// A count: decimal digits only, fits in 64 bits.
inline std::uint64_t parse_count(const std::string& flag, const std::string& v) {
    if (v.empty() || v.find_first_not_of("0123456789") != std::string::npos)
        throw bad_value(flag, v, "a non-negative integer");
    errno = 0;
    char* end = nullptr;
    const unsigned long long x = std::strtoull(v.c_str(), &end, 10);
    if (errno == ERANGE || end != v.c_str() + v.size()) throw bad_value(flag, v, "a non-negative integer");
    return static_cast<std::uint64_t>(x);
}

inline int parse_small(const std::string& flag, const std::string& v) {
    const std::uint64_t x = parse_count(flag, v);
    if (x > 0x7fffffffu) throw bad_value(flag, v, "an integer below 2^31");
    return static_cast<int>(x);
}

/// A finite real number, the whole argument.
inline double parse_real(const std::string& flag, const std::string& v) {
    errno = 0;
    char* end = nullptr;
    const double x = std::strtod(v.c_str(), &end);
    if (v.empty() || errno == ERANGE || end != v.c_str() + v.size() || !std::isfinite(x))
        throw bad_value(flag, v, "a finite number");
    return x;
}

/// One of a fixed set of names.
template <class E>
E parse_choice(const std::string& flag, const std::string& v,
               std::initializer_list<std::pair<const char*, E>> names) {
    std::string known;
    for (const auto& n : names) {
        if (v == n.first) return n.second;
        known += known.empty() ? n.first : std::string("|") + n.first;
    }
    throw std::runtime_error("parse_args: " + flag + " '" + v + "' is not one of " + known);
}

/// A flag and what its value does to the config (value-less flags get "").
template <class T>
struct FlagRule {
    const char* name;
    bool has_value;
    void (*apply)(BasicConfig<T>&, const std::string&);
};

template <class T>
const FlagRule<T>* find_rule(const std::string& flag) {
    using C = BasicConfig<T>;
    static const FlagRule<T> rules[] = {
        {"--nx", true, [](C& c, const std::string& v) { c.nx = parse_count("--nx", v); }},
        {"--ny", true, [](C& c, const std::string& v) { c.ny = parse_count("--ny", v); c.ny_explicit = true; }},
        {"--nz", true, [](C& c, const std::string& v) { c.nz = parse_count("--nz", v); c.nz_explicit = true; }},
        {"--dx", true, [](C& c, const std::string& v) { c.dx = static_cast<T>(parse_real("--dx", v)); }},
        {"--dy", true, [](C& c, const std::string& v) { c.dy = static_cast<T>(parse_real("--dy", v)); }},
        {"--dz", true, [](C& c, const std::string& v) { c.dz = static_cast<T>(parse_real("--dz", v)); }},
        {"--dt", true, [](C& c, const std::string& v) { c.dt = static_cast<T>(parse_real("--dt", v)); }},
        {"--steps", true, [](C& c, const std::string& v) { c.steps = parse_small("--steps", v); }},
        {"--warmup", true, [](C& c, const std::string& v) { c.warmup_steps = parse_small("--warmup", v); }},
        {"--velocity", true, [](C& c, const std::string& v) { c.velocity = static_cast<T>(parse_real("--velocity", v)); }},
        {"--diffusivity", true,
         [](C& c, const std::string& v) { c.diffusivity = static_cast<T>(parse_real("--diffusivity", v)); }},
        {"--seed", true, [](C& c, const std::string& v) { c.seed = parse_count("--seed", v); }},
        {"--device", true, [](C& c, const std::string& v) { c.device = parse_small("--device", v); }},
        {"--output", true, [](C& c, const std::string& v) { c.output_path = v; }},
        {"--verbose", false, [](C& c, const std::string&) { c.verbose = true; }},
        {"--physics", true,
         [](C& c, const std::string& v) {
             c.physics = parse_choice<Physics>("--physics", v,
                                               {{"wave", Physics::WavePropagation}, {"heat", Physics::HeatDiffusion}});
         }},
        {"--kernel", true,
         [](C& c, const std::string& v) {
             c.variant = parse_choice<KernelVariant>(
                 "--kernel", v,
                 {{"naive", KernelVariant::Naive}, {"tiled", KernelVariant::Tiled}, {"multigpu", KernelVariant::MultiGPU}});
         }},
        {"--vmodel", true,
         [](C& c, const std::string& v) {
             c.vmodel = parse_choice<VelocityModel>(
                 "--vmodel", v, {{"constant", VelocityModel::Constant}, {"layered-noise", VelocityModel::LayeredNoise}});
         }},
        {"--ic", true,
         [](C& c, const std::string& v) {
             c.ic = parse_choice<InitialCondition>(
                 "--ic", v, {{"gaussian", InitialCondition::Gaussian}, {"noise-cube", InitialCondition::NoiseCube}});
         }},
    };
    for (const auto& r : rules)
        if (flag == r.name) return &r;
    return nullptr;
}

}  // namespace detail

/// Fills a config from argv (argv[0] is the program name).
template <class T>
inline BasicConfig<T> parse_args_t(int argc, char** argv) {
    BasicConfig<T> cfg;
    int i = 1;
    while (i < argc) {
        const std::string flag = argv[i++];
        const detail::FlagRule<T>* rule = detail::find_rule<T>(flag);
        if (!rule) throw std::runtime_error("parse_args: unrecognised option '" + flag + "'");
        std::string value;
        if (rule->has_value) {
            if (i >= argc) throw std::runtime_error("parse_args: option " + flag + " needs a value");
            value = argv[i++];
        }
        rule->apply(cfg, value);
    }
    // The solvers take N^3 grids: --nx alone selects one; explicit --ny/--nz must agree.
    const bool cubic = cfg.nx == cfg.ny && cfg.ny == cfg.nz;
    if ((cfg.ny_explicit || cfg.nz_explicit) && !cubic)
        throw std::runtime_error("parse_args: grids must be cubic, got nx " + std::to_string(cfg.nx) + ", ny " +
                                 std::to_string(cfg.ny) + ", nz " + std::to_string(cfg.nz) +
                                 " (give --nx alone, or --ny / --nz equal to it)");
    return cfg;
}

inline Config parse_args(int argc, char** argv) { return parse_args_t<real_t>(argc, argv); }

}  // namespace stencil
