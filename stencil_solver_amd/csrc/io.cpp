// Benchmark result files (reference src/utils/io.cpp:91-162): the same JSON
// schema, written by hand; the device fields come from HIP.
#include "stencil/io.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "stencil/types.hpp"
#include "launchers.hpp"
#include "stencil_capi.h"

namespace stencil {
namespace {

std::string json_escape(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

std::string runtime_version() {
    int v = 0;
    if (hipRuntimeGetVersion(&v) != hipSuccess) return "unknown";
    // HIP encodes major * 10^7 + minor * 10^5 + patch
    return std::to_string(v / 10000000) + "." + std::to_string((v / 100000) % 100);
}

void write_file(const std::string& path, const std::string& text, const char* what) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error(std::string("io: cannot open ") + what + ": " + path);
    f << text;
}

}  // namespace

std::string device_name(int device) {
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return "unknown";
    if (p.name[0] != '\0') return p.name;
    std::string arch = p.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    return arch.empty() ? "unknown" : "AMD " + arch;
}

void write_results_json(const std::string& path, const std::vector<PerfResult>& results, int device,
                        int steps, int warmup) {
    if (path.empty()) return;
    // launch shape of the tiled kernel on the largest grid of the run
    std::size_t n = 0;
    for (const auto& r : results) n = std::max(n, r.grid_size);
    stencil_tile_shape shape{};
    int tile_x = 0, tile_y = 0, pencil_z = 0;
    if (n > 8) {
        const stencil_grid g{n, n, n, 1.0, 1.0, 1.0, 1.0};
        if (stencil_tuned_shape(&g, 0, &shape) >= 0) {
            tile_x = 4 * shape.tile_x4;
            tile_y = shape.tile_y;
            pencil_z = std::min(int(n - 8), hipk::zq_chunk_planes(int64_t(n - 8), shape.n_chunks));
        }
    }
    std::ostringstream js;
    js << "{\n";
    js << "  \"gpu_name\": \"" << json_escape(device_name(device)) << "\",\n";
    js << "  \"cuda_version\": \"" << json_escape(runtime_version()) << "\",\n";
    js << "  \"runtime\": \"hip\",\n";
    js << "  \"stencil_radius\": " << STENCIL_RADIUS << ",\n";
    js << "  \"precision\": \"" << (sizeof(real_t) == 8 ? "fp64" : "fp32") << "\",\n";
    js << "  \"tile_x\": " << tile_x << ",\n";
    js << "  \"tile_y\": " << tile_y << ",\n";
    js << "  \"pencil_z\": " << pencil_z << ",\n";
    js << "  \"steps\": " << steps << ",\n";
    js << "  \"warmup\": " << warmup << ",\n";
    js << "  \"results\": [\n";
    for (std::size_t i = 0; i < results.size(); ++i) {
        const PerfResult& r = results[i];
        js << "    {\n";
        js << "      \"label\": \"" << json_escape(r.label) << "\",\n";
        js << "      \"grid_size\": " << r.grid_size << ",\n";
        js << "      \"time_ms\": " << r.time_ms << ",\n";
        js << "      \"bandwidth_gbs\": " << r.bandwidth_gbs << ",\n";
        js << "      \"model_bandwidth_gbs\": " << r.model_bandwidth_gbs << ",\n";
        js << "      \"gflops\": " << r.gflops << "\n";
        js << "    }" << (i + 1 < results.size() ? "," : "") << "\n";
    }
    js << "  ]\n}\n";
    write_file(path, js.str(), "output file");
    std::printf("Results written to: %s\n", path.c_str());
}

void write_scaling_json(const std::string& path, const std::vector<ScalingResult>& results) {
    if (path.empty()) return;
    std::ostringstream js;
    js << "{\n  \"results\": [\n";
    for (std::size_t i = 0; i < results.size(); ++i) {
        const ScalingResult& r = results[i];
        js << "    {\n";
        js << "      \"mode\": \"" << json_escape(r.mode) << "\",\n";
        js << "      \"ngpu\": " << r.ngpu << ",\n";
        js << "      \"time_ms\": " << r.time_ms << ",\n";
        js << "      \"efficiency\": " << r.efficiency << "\n";
        js << "    }" << (i + 1 < results.size() ? "," : "") << "\n";
    }
    js << "  ]\n}\n";
    write_file(path, js.str(), "scaling output");
    std::printf("Scaling results written to: %s\n", path.c_str());
}

}  // namespace stencil
