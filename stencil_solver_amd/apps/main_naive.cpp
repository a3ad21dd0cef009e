// stencil_naive: times the coalesced global-memory kernel on one grid or on
// the 128/256/512 sweep and writes the results file (flags and JSON schema of
// the reference's stencil_naive, src/main_naive.cpp).
#include <cstdlib>
#include <exception>
#include <vector>

#include "common.hpp"

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
    std::vector<PerfResult> rows;
    for (std::size_t n : app::grid_sizes(cfg)) {
        try {
            rows.push_back(app::timed_run(cfg, n, KernelVariant::Naive));
        } catch (const std::exception& e) {
            std::fprintf(stderr, "stencil_naive: %zu^3 run failed: %s\n", n, e.what());
            return EXIT_FAILURE;
        }
        std::printf("%zu^3, %d steps after %d warm-up steps:\n", n, cfg.steps, cfg.warmup_steps);
        print_perf(rows.back());
    }
    format_perf_table(rows);
    try {
        write_results_json(cfg.output_path, rows, cfg.device, cfg.steps, cfg.warmup_steps);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "stencil_naive: %s\n", e.what());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
