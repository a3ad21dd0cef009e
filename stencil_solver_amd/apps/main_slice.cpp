// stencil_slice: records the middle z-plane of a run every N steps in the
// slice format of include/stencil/slice.hpp (the reference stencil_slice's
// files, src/main_slice.cpp:106-152).  Its own options are --every N
// (default 5) and --slice-out PREFIX (default "slices"); every other flag is
// parse_args's, and --nx N selects the N^3 grid.
#include <cstdlib>
#include <exception>
#include <map>
#include <string>

#include "common.hpp"
#include "stencil/slice.hpp"

int main(int argc, char** argv) {
    using namespace stencil;
    const char* options =
        "[--nx N] [--steps N] [--every N] [--slice-out PREFIX] [--physics wave|heat] [--kernel naive|tiled] "
        "[--velocity V] [--device N]";
    Config cfg;
    int every = 5;
    std::string prefix = "slices";
    try {
        std::map<std::string, std::string> own;
        std::vector<char*> rest = app::split_own_options(argc, argv, {"--every", "--slice-out"}, own);
        cfg = parse_args(static_cast<int>(rest.size()), rest.data());
        if (own.count("--every")) {
            const std::uint64_t e = detail::parse_count("--every", own["--every"]);
            if (e < 1 || e > 0x7fffffffu) throw std::runtime_error("--every takes a step count of at least 1");
            every = static_cast<int>(e);
        }
        if (own.count("--slice-out")) prefix = own["--slice-out"];
    } catch (const std::exception& e) {
        return app::refuse(argv[0], e.what(), options);
    }
    cfg.ny = cfg.nz = cfg.nx;
    try {
        auto solver = make_solver(cfg);
        const SliceMeta m = record_middle_plane(*solver, cfg, cfg.steps, every, prefix);
        std::printf("stencil_slice: %d frames of plane z = %zu (%zu x %zu) in %s.f32, metadata in %s.json\n",
                    (cfg.steps / every) + 1, m.plane_index, m.nx, m.ny, prefix.c_str(), prefix.c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "stencil_slice: %s\n", e.what());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
