#pragma once

/**
 * include/stencil/timer.hpp -- CpuTimer, hipEvent GpuTimer, PerfResult and
 * console helpers; the reference's include/stencil/timer.hpp:25-122 on HIP.
 */

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "stencil/grid.hpp"
#include "stencil/types.hpp"

namespace stencil {

class CpuTimer {
public:
    void start() noexcept { t0_ = clock_t::now(); }
    void stop() noexcept { t1_ = clock_t::now(); }
    double elapsed_ms() const noexcept {
        return std::chrono::duration<double, std::milli>(t1_ - t0_).count();
    }

private:
    using clock_t = std::chrono::steady_clock;
    clock_t::time_point t0_, t1_;
};

/// hipEvent timer on a stream (reference timer.hpp:43-71).
class GpuTimer {
public:
    GpuTimer() {
        HIP_CHECK(hipEventCreate(&start_));
        HIP_CHECK(hipEventCreate(&stop_));
    }
    ~GpuTimer() {
        (void)hipEventDestroy(start_);
        (void)hipEventDestroy(stop_);
    }
    GpuTimer(const GpuTimer&) = delete;
    GpuTimer& operator=(const GpuTimer&) = delete;

    void start(hipStream_t s = nullptr) { HIP_CHECK(hipEventRecord(start_, s)); }
    void stop(hipStream_t s = nullptr) { HIP_CHECK(hipEventRecord(stop_, s)); }
    /// Blocks until the stop event completes, then returns elapsed ms.
    float elapsed_ms() {
        float ms = 0.0f;
        HIP_CHECK(hipEventSynchronize(stop_));
        HIP_CHECK(hipEventElapsedTime(&ms, start_, stop_));
        return ms;
    }

private:
    hipEvent_t start_{};
    hipEvent_t stop_{};
};

/// Reference timer.hpp:88-96.
struct PerfResult {
    std::string label;
    double time_ms = 0.0;        ///< mean time per step [ms]
    double bandwidth_gbs = 0.0;  ///< compulsory bytes / time [GB/s]
    double gflops = 0.0;         ///< [GFLOP/s]
    double model_bandwidth_gbs = 0.0;  ///< no-reuse model bytes / time [GB/s]
    std::size_t grid_size = 0;   ///< N of an N^3 grid
};

inline void print_perf(const PerfResult& r) {
    std::printf("  %-24s  %8.3f ms/step  %7.1f GB/s  %7.1f GFLOP/s\n", r.label.c_str(), r.time_ms,
                r.bandwidth_gbs, r.gflops);
}

inline void print_speedup(const PerfResult& base, const PerfResult& opt) {
    if (base.time_ms <= 0.0) return;
    std::printf("\n  Speedup  (%s → %s) : %.2f×\n", base.label.c_str(), opt.label.c_str(),
                base.time_ms / opt.time_ms);
    std::printf("  Bandwidth gain          : %+.1f%%\n",
                (opt.bandwidth_gbs / base.bandwidth_gbs - 1.0) * 100.0);
}

/// One row per result, labelled by grid size; no cross-row speedup column
/// (reference src/utils/timer.cpp:36-49).
void format_perf_table(const std::vector<PerfResult>& results);

}  // namespace stencil
