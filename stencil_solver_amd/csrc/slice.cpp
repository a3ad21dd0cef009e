// Slice recordings (include/stencil/slice.hpp): frames go straight from a
// one-plane device copy (ISolver::download_plane_z) to the .f32 file.
#include "stencil/slice.hpp"

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace stencil {
namespace {

std::string json_number(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.9g", v);
    return buf;
}

const char* physics_key(Physics p) { return p == Physics::HeatDiffusion ? "heat" : "wave"; }

const char* kernel_key(KernelVariant v) {
    if (v == KernelVariant::Naive) return "naive";
    if (v == KernelVariant::Tiled) return "tiled";
    return "multigpu";
}

}  // namespace

SliceWriter::SliceWriter(std::string prefix, std::size_t nx, std::size_t ny)
    : prefix_(std::move(prefix)), values_(nx * ny) {
    const std::string path = prefix_ + ".f32";
    bin_ = std::fopen(path.c_str(), "wb");
    if (!bin_) throw std::runtime_error("SliceWriter: cannot create " + path);
}

SliceWriter::~SliceWriter() {
    if (bin_) std::fclose(bin_);
}

void SliceWriter::append(const float* frame) {
    if (!bin_) throw std::runtime_error("SliceWriter: " + prefix_ + " is already finished");
    if (std::fwrite(frame, sizeof(float), values_, bin_) != values_)
        throw std::runtime_error("SliceWriter: short write to " + prefix_ + ".f32");
    ++frames_;
}

void SliceWriter::finish(SliceMeta m) {
    if (bin_) {
        const bool ok = std::fclose(bin_) == 0;
        bin_ = nullptr;
        if (!ok) throw std::runtime_error("SliceWriter: cannot close " + prefix_ + ".f32");
    }
    // key order as the format lists it; one key per line
    const std::vector<std::pair<std::string, std::string>> kv = {
        {"nx", std::to_string(m.nx)},
        {"ny", std::to_string(m.ny)},
        {"nz", std::to_string(m.nz)},
        {"plane", "\"xy\""},
        {"plane_index", std::to_string(m.plane_index)},
        {"frames", std::to_string(frames_)},
        {"every", std::to_string(m.every)},
        {"steps", std::to_string(m.steps)},
        {"dt", json_number(m.dt)},
        {"dx", json_number(m.dx)},
        {"velocity", json_number(m.velocity)},
        {"physics", "\"" + m.physics + "\""},
        {"kernel", "\"" + m.kernel + "\""},
        {"dtype", "\"float32\""},
    };
    std::string text = "{\n";
    for (std::size_t i = 0; i < kv.size(); ++i)
        text += "  \"" + kv[i].first + "\": " + kv[i].second + (i + 1 < kv.size() ? ",\n" : "\n");
    text += "}\n";
    const std::string path = prefix_ + ".json";
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("SliceWriter: cannot create " + path);
    const bool ok = std::fputs(text.c_str(), f) >= 0;
    if (std::fclose(f) != 0 || !ok) throw std::runtime_error("SliceWriter: cannot write " + path);
}

template <class T>
SliceMeta record_middle_plane(BasicISolver<T>& solver, const BasicConfig<T>& cfg, int steps, int every,
                              const std::string& prefix) {
    if (every < 1) throw std::invalid_argument("record_middle_plane: every must be at least 1");
    const BasicGridDims<T>& d = solver.dims();
    SliceMeta m;
    m.nx = d.nx;
    m.ny = d.ny;
    m.nz = d.nz;
    m.plane_index = d.nz / 2;
    m.every = every;
    m.steps = steps;
    m.dt = static_cast<double>(d.dt);
    m.dx = static_cast<double>(d.dx);
    m.velocity = static_cast<double>(cfg.velocity);
    m.physics = physics_key(cfg.physics);
    m.kernel = kernel_key(cfg.variant);
    SliceWriter w(prefix, d.nx, d.ny);
    std::vector<T> plane(d.nx * d.ny);
    std::vector<float> frame(plane.size());
    auto grab = [&] {
        solver.download_plane_z(m.plane_index, plane.data());
        for (std::size_t i = 0; i < plane.size(); ++i) frame[i] = static_cast<float>(plane[i]);
        w.append(frame.data());
    };
    grab();
    for (int s = 1; s <= steps; ++s) {
        solver.step();
        if (s % every == 0) grab();
    }
    w.finish(m);
    return m;
}

template SliceMeta record_middle_plane<float>(BasicISolver<float>&, const BasicConfig<float>&, int, int,
                                              const std::string&);
template SliceMeta record_middle_plane<double>(BasicISolver<double>&, const BasicConfig<double>&, int, int,
                                               const std::string&);

}  // namespace stencil
