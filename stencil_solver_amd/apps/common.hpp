// Helpers shared by the command-line programs stencil_naive, stencil_tiled
// and stencil_slice (flag-compatible with the reference's programs of the
// same names, src/main_{naive,tiled,slice}.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "stencil/config.hpp"
#include "stencil/io.hpp"
#include "stencil/solver.hpp"
#include "stencil/timer.hpp"

namespace stencil::app {

/// One line per fact about the device the run uses.
inline void describe_device(int device) {
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, device) != hipSuccess) {
        std::printf("device %d: not available\n", device);
        return;
    }
    std::printf("device %d: %s [%s], %d CUs, %.1f GB HBM\n", device, device_name(device).c_str(), p.gcnArchName,
                p.multiProcessorCount, static_cast<double>(p.totalGlobalMem) / 1e9);
}

/// Grid sizes a run covers: the one given, or -- when --nx/--ny/--nz stay at
/// the default 256 -- the 128^3, 256^3, 512^3 sweep the reference programs run.
inline std::vector<std::size_t> grid_sizes(const Config& cfg) {
    const bool defaults = cfg.nx == 256 && cfg.ny == 256 && cfg.nz == 256;
    return defaults ? std::vector<std::size_t>{128, 256, 512} : std::vector<std::size_t>{cfg.nx};
}

/// path with `tag` inserted before its extension ("a/b.json" -> "a/b_naive.json").
inline std::string tagged_path(const std::string& path, const std::string& tag) {
    if (path.empty()) return path;
    const std::size_t slash = path.find_last_of('/');
    const std::size_t dot = path.find_last_of('.');
    const bool has_ext = dot != std::string::npos && (slash == std::string::npos || dot > slash);
    return has_ext ? path.substr(0, dot) + tag + path.substr(dot) : path + tag;
}

/// Takes this program's own value options (names in `own`) out of argv:
/// their values land in `values`, everything else stays for parse_args.
inline std::vector<char*> split_own_options(int argc, char** argv, const std::vector<std::string>& own,
                                            std::map<std::string, std::string>& values) {
    std::vector<char*> rest;
    rest.reserve(static_cast<std::size_t>(argc));
    rest.push_back(argv[0]);
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        bool mine = false;
        for (const auto& o : own) mine = mine || a == o;
        if (!mine) {
            rest.push_back(argv[i]);
            continue;
        }
        if (i + 1 >= argc) throw std::runtime_error("option " + a + " needs a value");
        values[a] = argv[++i];
    }
    return rest;
}

/// Prints why the arguments were refused and how to call the program.
inline int refuse(const char* prog, const std::string& why, const char* options) {
    std::fprintf(stderr, "%s: %s\nUsage: %s %s\n", prog, why.c_str(), prog, options);
    return 2;
}

/// One timed run (reference protocol: `warmup` untimed steps, then `steps`
/// timed ones) of the configured solver on an n^3 grid.
inline PerfResult timed_run(Config cfg, std::size_t n, KernelVariant variant) {
    cfg.nx = cfg.ny = cfg.nz = n;
    cfg.variant = variant;
    auto solver = make_solver(cfg);
    return solver->run(cfg.steps, cfg.warmup_steps);
}

}  // namespace stencil::app
