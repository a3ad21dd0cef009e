#pragma once

/**
 * include/stencil/timer.hpp -- step timers and performance records.
 *
 * Interface kept from the reference's include/stencil/timer.hpp:25-122
 * (CpuTimer, GpuTimer, PerfResult, print_perf, print_speedup,
 * format_perf_table): callers written against it compile unchanged.
 *
 *  - CpuTimer: host wall clock (steady), milliseconds between the last
 *    start() and stop().
 *  - GpuTimer: a pair of HIP events recorded on a stream; elapsed_ms() waits
 *    for the stop event.  A HIP failure ends the process through HIP_CHECK,
 *    the reference's error behaviour for device calls.
 *  - PerfResult: one measured configuration (per-step time and the rates
 *    derived from the byte / flop models of types.hpp).
 */

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "stencil/grid.hpp"
#include "stencil/types.hpp"

namespace stencil {

class CpuTimer {
public:
    void start() noexcept { begin_ = now_ns(); }
    void stop() noexcept { end_ = now_ns(); }
    double elapsed_ms() const noexcept { return static_cast<double>(end_ - begin_) * 1e-6; }

private:
    static long long now_ns() noexcept {
        return std::chrono::duration_cast<std::chrono::nanoseconds>(
                   std::chrono::steady_clock::now().time_since_epoch())
            .count();
    }
    long long begin_ = 0, end_ = 0;
};

class GpuTimer {
public:
    GpuTimer() : start_(make_event()), stop_(make_event()) {}
    GpuTimer(const GpuTimer&) = delete;
    GpuTimer& operator=(const GpuTimer&) = delete;

    /// Records the start / stop event on stream s (nullptr: the default stream).
    void start(hipStream_t s = nullptr) { HIP_CHECK(hipEventRecord(start_.get(), s)); }
    void stop(hipStream_t s = nullptr) { HIP_CHECK(hipEventRecord(stop_.get(), s)); }
    /// Waits for the stop event; milliseconds from start to stop.
    float elapsed_ms() {
        HIP_CHECK(hipEventSynchronize(stop_.get()));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, start_.get(), stop_.get()));
        return ms;
    }

private:
    struct EventDeleter {
        void operator()(hipEvent_t e) const noexcept { (void)hipEventDestroy(e); }
    };
    using Event = std::unique_ptr<std::remove_pointer_t<hipEvent_t>, EventDeleter>;
    static Event make_event() {
        hipEvent_t e = nullptr;
        HIP_CHECK(hipEventCreate(&e));
        return Event(e);
    }
    Event start_, stop_;
};

/// One measured configuration (field names as in the reference).
struct PerfResult {
    std::string label;                 ///< solver name, e.g. "wave/tiled"
    double time_ms = 0.0;              ///< mean time of one step [ms]
    double bandwidth_gbs = 0.0;        ///< compulsory bytes per step / time [GB/s]
    double gflops = 0.0;               ///< model flops per step / time [GFLOP/s]
    double model_bandwidth_gbs = 0.0;  ///< no-reuse model bytes per step / time [GB/s]
    std::size_t grid_size = 0;         ///< N of an N^3 grid
};

/// One line: label, time per step, bandwidth and flop rate.
inline void print_perf(const PerfResult& r) {
    std::printf("  [%s] %.3f ms per step | %.1f GB/s | %.1f GFLOP/s\n", r.label.c_str(), r.time_ms,
                r.bandwidth_gbs, r.gflops);
}

/// How much faster `opt` is than `base` (time ratio) and its bandwidth gain;
/// prints nothing when base has no time.
inline void print_speedup(const PerfResult& base, const PerfResult& opt) {
    if (base.time_ms <= 0.0 || opt.time_ms <= 0.0) return;
    const double gain = base.bandwidth_gbs > 0.0 ? 100.0 * (opt.bandwidth_gbs / base.bandwidth_gbs - 1.0) : 0.0;
    std::printf("  %s runs %.2fx as fast as %s (bandwidth %+.1f %%)\n", opt.label.c_str(),
                base.time_ms / opt.time_ms, base.label.c_str(), gain);
}

/// A table with one row per result (label, grid N^3, ms per step, GB/s,
/// GFLOP/s); rows are independent measurements, so there is no ratio column.
void format_perf_table(const std::vector<PerfResult>& results);

}  // namespace stencil
