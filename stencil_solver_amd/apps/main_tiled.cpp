// stencil_tiled: naive vs tiled on one grid or the 128/256/512 sweep (same
// flags, table and JSON files as the reference's src/main_tiled.cpp; the
// tiled launch shape is the autotuner's for each grid).
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "common.hpp"
#include "stencil/config.hpp"
#include "stencil/io.hpp"
#include "stencil/solver.hpp"
#include "stencil/timer.hpp"
#include "stencil_capi.h"

int main(int argc, char** argv) {
    std::printf("=== stencil-solver: naive vs tiled comparison ===\n\n");
    stencil::Config cfg;
    try {
        cfg = stencil::parse_args(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr,
                     "Error: %s\n\nUsage: %s [--nx N] [--ny N] [--nz N] [--steps N] [--warmup N]\n"
                     "          [--physics wave|heat] [--velocity F] [--device N]\n"
                     "          [--output path.json] [--verbose]\n",
                     e.what(), argv[0]);
        return EXIT_FAILURE;
    }
    stencil::app::print_device_info(cfg.device);
    std::printf("\nStencil radius : %d  (%d-th order)\n", STENCIL_RADIUS, 2 * STENCIL_RADIUS);
    std::printf("Tiled kernel   : LDS plane + z register queue, float4 per thread; launch shape\n"
                "                 from the autotuner per grid (stencil_tune_tiled_f32)\n\n");
    std::printf("%-8s  %10s  %10s  %10s  %10s  %8s  %s\n", "Grid", "Naive ms", "BW (GB/s)", "Tiled ms",
                "BW (GB/s)", "Speedup", "Tile (x, y, z-chunk)");

    std::vector<stencil::PerfResult> naive, tiled;
    for (std::size_t n : stencil::app::sweep_sizes(cfg.nx, cfg.ny, cfg.nz)) {
        stencil::Config run = cfg;
        run.nx = run.ny = run.nz = n;
        stencil::PerfResult rn, rt;
        try {
            run.variant = stencil::KernelVariant::Naive;
            auto s = stencil::make_solver(run);
            rn = s->run(run.steps, run.warmup_steps);
            rn.label = "naive";
            naive.push_back(rn);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "  naive %zu^3 FAILED: %s\n", n, e.what());
            continue;
        }
        try {
            run.variant = stencil::KernelVariant::Tiled;
            auto s = stencil::make_solver(run);
            rt = s->run(run.steps, run.warmup_steps);
            rt.label = "tiled";
            tiled.push_back(rt);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "  tiled %zu^3 FAILED: %s\n", n, e.what());
            continue;
        }
        stencil_tile_shape sh{};
        const stencil_grid g{n, n, n, 1.0, 1.0, 1.0, 1.0};
        (void)stencil_tuned_shape(&g, run.physics == stencil::Physics::HeatDiffusion ? 1 : 0, &sh);
        std::printf("%-8zu  %8.3f ms  %8.1f     %8.3f ms  %8.1f     %6.2fx  %d x %d, %d chunks\n", n, rn.time_ms,
                    rn.bandwidth_gbs, rt.time_ms, rt.bandwidth_gbs, rt.time_ms > 0 ? rn.time_ms / rt.time_ms : 0.0,
                    4 * sh.tile_x4, sh.tile_y, sh.n_chunks);
        std::fflush(stdout);
    }
    if (!naive.empty()) {
        std::printf("\n-- Naive --\n");
        stencil::format_perf_table(naive);
        std::printf("-- Tiled --\n");
        stencil::format_perf_table(tiled);
    }
    // tiled results -> --output, naive results -> the "_naive" sibling
    stencil::write_results_json(cfg.output_path, tiled, cfg.device, cfg.steps, cfg.warmup_steps);
    stencil::write_results_json(stencil::app::with_suffix(cfg.output_path, "_naive"), naive, cfg.device, cfg.steps,
                                cfg.warmup_steps);
    return EXIT_SUCCESS;
}
