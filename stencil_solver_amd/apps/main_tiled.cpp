// stencil_tiled: naive and tiled kernels side by side on one grid or on the
// 128/256/512 sweep; tiled results go to --output and naive ones to the
// "_naive" file beside it (flags and JSON schema of the reference's
// stencil_tiled, src/main_tiled.cpp).  The tiled launch shape of each grid is
// the autotuner's cached pick (stencil_tuned_shape), reported per row.
#include <cstdlib>
#include <exception>
#include <vector>

#include "common.hpp"
#include "stencil_capi.h"

int main(int argc, char** argv) {
    using namespace stencil;
    Config cfg;
    try {
        cfg = parse_args(argc, argv);
    } catch (const std::exception& e) {
        return app::refuse(argv[0], e.what(),
                           "[--nx N] [--steps N] [--warmup N] [--physics wave|heat] [--velocity V] "
                           "[--diffusivity D] [--device N] [--output FILE.json] [--verbose]");
    }
    app::describe_device(cfg.device);
    const bool heat = cfg.physics == Physics::HeatDiffusion;
    std::vector<PerfResult> naive, tiled;
    for (std::size_t n : app::grid_sizes(cfg)) {
        PerfResult a, b;
        try {
            a = app::timed_run(cfg, n, KernelVariant::Naive);
            b = app::timed_run(cfg, n, KernelVariant::Tiled);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "stencil_tiled: %zu^3 run failed: %s\n", n, e.what());
            continue;
        }
        a.label = "naive";
        b.label = "tiled";
        naive.push_back(a);
        tiled.push_back(b);
        stencil_tile_shape shape{};
        const stencil_grid g{n, n, n, 1.0, 1.0, 1.0, 1.0};
        const bool tuned = stencil_tuned_shape(&g, heat ? 1 : 0, &shape) == 1;
        std::printf("%zu^3: tiled tile %d x %d cells, %d z-chunks (%s)\n", n, 4 * shape.tile_x4, shape.tile_y,
                    shape.n_chunks, tuned ? "tuned" : "default");
        print_perf(a);
        print_perf(b);
        print_speedup(a, b);
        std::fflush(stdout);
    }
    format_perf_table(naive);
    format_perf_table(tiled);
    try {
        write_results_json(cfg.output_path, tiled, cfg.device, cfg.steps, cfg.warmup_steps);
        write_results_json(app::tagged_path(cfg.output_path, "_naive"), naive, cfg.device, cfg.steps,
                           cfg.warmup_steps);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "stencil_tiled: %s\n", e.what());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
