#pragma once
/*
 * include/stencil/io.hpp -- benchmark result files (reference
 * include/stencil/io.hpp:14-34).  Same structs, functions and JSON schema, so
 * the reference's plot_results.py / summarize_results.py / repeat_bench.py
 * read these files unchanged.
 */
#include <string>
#include <vector>

#include "stencil/timer.hpp"

namespace stencil {

/// One multi-GPU scaling point (schema of the reference's scaling files).
struct ScalingResult {
    std::string mode;  ///< "weak" or "strong"
    int ngpu = 1;
    double time_ms = 0.0;
    double efficiency = 0.0;
};

/// The device's marketing name, or "AMD <gfx arch>" where the runtime reports
/// none (some ROCm installs leave hipDeviceProp_t::name empty).
std::string device_name(int device);

/// Writes {gpu_name, cuda_version, stencil_radius, precision, tile_x, tile_y,
/// pencil_z, steps, warmup, results[]} to `path` (nothing if the path is
/// empty).  gpu_name / cuda_version report the HIP device and runtime (the
/// key names are the schema's); tile_x / tile_y / pencil_z are the tiled
/// kernel's launch shape for the largest grid in `results` (cells, rows,
/// z-chunk length).  Throws std::runtime_error if the file cannot be opened.
void write_results_json(const std::string& path, const std::vector<PerfResult>& results, int device,
                        int steps, int warmup);

/// Writes {results: [{mode, ngpu, time_ms, efficiency}]}.
void write_scaling_json(const std::string& path, const std::vector<ScalingResult>& results);

}  // namespace stencil
