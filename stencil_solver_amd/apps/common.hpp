// Shared pieces of the command-line programs (the reference's
// src/main_naive.cpp, main_tiled.cpp, main_slice.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "stencil/io.hpp"

namespace stencil::app {

// Device banner: name, compute units, memory and the nominal HBM bandwidth.
inline void print_device_info(int device) {
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, device) != hipSuccess) {
        std::printf("GPU [%d]: unavailable\n", device);
        return;
    }
    std::printf("GPU [%d]: %s (%s)\n", device, device_name(device).c_str(), p.gcnArchName);
    std::printf("  Compute units: %d   Memory: %.1f GB\n", p.multiProcessorCount,
                static_cast<double>(p.totalGlobalMem) / 1e9);
    if (p.memoryClockRate > 0 && p.memoryBusWidth > 0)
        std::printf("  Peak bandwidth: %.1f GB/s\n",
                    2.0 * static_cast<double>(p.memoryClockRate) * 1e3 *
                        static_cast<double>(p.memoryBusWidth) / 8.0 / 1e9);
}

// Without an explicit grid (--nx left at its default 256) the programs sweep
// 128^3, 256^3 and 512^3 for the report table.
inline std::vector<std::size_t> sweep_sizes(std::size_t nx, std::size_t ny, std::size_t nz) {
    if (nx != 256 || ny != 256 || nz != 256) return {nx};
    return {128, 256, 512};
}

// "results/tiled.json" -> "results/tiled_naive.json"
inline std::string with_suffix(const std::string& path, const std::string& suffix) {
    if (path.empty()) return {};
    const auto dot = path.rfind('.');
    const auto slash = path.rfind('/');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return path + suffix;
    return path.substr(0, dot) + suffix + path.substr(dot);
}

}  // namespace stencil::app
