// stencil_slice: runs the solver and saves the middle z-plane every N steps
// (the reference's src/main_slice.cpp and its on-disk format):
//   PREFIX.f32   float32 frames, frame-major, each ny rows of nx values (x fastest)
//   PREFIX.json  nx, ny, nz, plane, plane_index, frames, every, steps, dt, dx,
//                velocity, physics, kernel, dtype
// Frame 0 is the initial state.  Extra flags: --every N (default 5) and
// --slice-out PREFIX (default "slices"); the rest is parse_args's.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "stencil/config.hpp"
#include "stencil/grid.hpp"
#include "stencil/solver.hpp"

namespace {

struct SliceArgs {
    int every = 5;
    std::string prefix = "slices";
};

std::vector<char*> take_slice_args(int argc, char** argv, SliceArgs& sa) {
    std::vector<char*> rest{argv[0]};
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--every" || a == "--slice-out") {
            if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
            const std::string v = argv[++i];
            if (a == "--every") sa.every = std::stoi(v);
            else sa.prefix = v;
        } else {
            rest.push_back(argv[i]);
        }
    }
    return rest;
}

}  // namespace

int main(int argc, char** argv) {
    SliceArgs sa;
    stencil::Config cfg;
    try {
        std::vector<char*> rest = take_slice_args(argc, argv, sa);
        cfg = stencil::parse_args(static_cast<int>(rest.size()), rest.data());
        if (sa.every < 1) throw std::runtime_error("--every must be >= 1");
    } catch (const std::exception& e) {
        std::fprintf(stderr,
                     "Error: %s\n\nUsage: %s [--nx N] [--steps N] [--every N] [--slice-out PREFIX]\n"
                     "          [--physics wave|heat] [--kernel naive|tiled] [--velocity F] [--device N]\n",
                     e.what(), argv[0]);
        return EXIT_FAILURE;
    }
    cfg.ny = cfg.nz = cfg.nx;  // the solvers take cubic grids; --nx selects N^3
    try {
        auto solver = stencil::make_solver(cfg);
        const stencil::GridDims& d = solver->dims();
        const std::size_t zmid = d.nz / 2;
        std::ofstream bin(sa.prefix + ".f32", std::ios::binary);
        if (!bin) throw std::runtime_error("cannot open " + sa.prefix + ".f32");
        stencil::HostGrid field(d);
        std::vector<float> frame(d.nx * d.ny);
        int frames = 0;
        auto save = [&] {
            solver->download(field);
            for (std::size_t y = 0; y < d.ny; ++y)
                for (std::size_t x = 0; x < d.nx; ++x) frame[y * d.nx + x] = static_cast<float>(field.at(x, y, zmid));
            bin.write(reinterpret_cast<const char*>(frame.data()), static_cast<std::streamsize>(frame.size() * 4));
            ++frames;
        };
        save();
        for (int s = 1; s <= cfg.steps; ++s) {
            solver->step();
            if (s % sa.every == 0) save();
        }
        bin.close();
        std::ofstream js(sa.prefix + ".json");
        if (!js) throw std::runtime_error("cannot open " + sa.prefix + ".json");
        const bool heat = cfg.physics == stencil::Physics::HeatDiffusion;
        const char* kernel = cfg.variant == stencil::KernelVariant::Naive ? "naive"
                             : cfg.variant == stencil::KernelVariant::Tiled ? "tiled" : "multigpu";
        js << "{\n  \"nx\": " << d.nx << ",\n  \"ny\": " << d.ny << ",\n  \"nz\": " << d.nz
           << ",\n  \"plane\": \"xy\",\n  \"plane_index\": " << zmid << ",\n  \"frames\": " << frames
           << ",\n  \"every\": " << sa.every << ",\n  \"steps\": " << cfg.steps
           << ",\n  \"dt\": " << static_cast<double>(d.dt) << ",\n  \"dx\": " << static_cast<double>(d.dx)
           << ",\n  \"velocity\": " << static_cast<double>(cfg.velocity) << ",\n  \"physics\": \""
           << (heat ? "heat" : "wave") << "\",\n  \"kernel\": \"" << kernel << "\",\n  \"dtype\": \"float32\"\n}\n";
        std::printf("Wrote %d frames (%zu x %zu, z = %zu) to %s.f32 and %s.json\n", frames, d.nx, d.ny, zmid,
                    sa.prefix.c_str(), sa.prefix.c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "FAILED: %s\n", e.what());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
