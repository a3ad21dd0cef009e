// stencil_naive: the coalesced global-memory kernel over one grid or the
// 128/256/512 sweep (same flags, output and JSON as the reference's
// src/main_naive.cpp).
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "common.hpp"
#include "stencil/config.hpp"
#include "stencil/io.hpp"
#include "stencil/solver.hpp"
#include "stencil/timer.hpp"

int main(int argc, char** argv) {
    std::printf("=== stencil-solver: naive kernel ===\n\n");
    stencil::Config cfg;
    try {
        cfg = stencil::parse_args(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr,
                     "Error: %s\n\nUsage: %s [--nx N] [--ny N] [--nz N] [--steps N] [--warmup N] "
                     "[--physics wave|heat] [--velocity F] [--device N] [--output path] [--verbose]\n",
                     e.what(), argv[0]);
        return EXIT_FAILURE;
    }
    cfg.variant = stencil::KernelVariant::Naive;  // this program always runs the global-memory kernel
    stencil::app::print_device_info(cfg.device);
    std::printf("\n");

    std::vector<stencil::PerfResult> results;
    for (std::size_t n : stencil::app::sweep_sizes(cfg.nx, cfg.ny, cfg.nz)) {
        stencil::Config run = cfg;
        run.nx = run.ny = run.nz = n;
        std::printf("Grid: %zu^3   steps: %d   warmup: %d\n", n, run.steps, run.warmup_steps);
        try {
            auto solver = stencil::make_solver(run);
            stencil::PerfResult r = solver->run(run.steps, run.warmup_steps);
            stencil::print_perf(r);
            results.push_back(r);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "  FAILED: %s\n", e.what());
            return EXIT_FAILURE;
        }
    }
    if (results.size() > 1) {
        std::printf("\n-- Summary --\n");
        stencil::format_perf_table(results);
    }
    stencil::write_results_json(cfg.output_path, results, cfg.device, cfg.steps, cfg.warmup_steps);
    return EXIT_SUCCESS;
}
