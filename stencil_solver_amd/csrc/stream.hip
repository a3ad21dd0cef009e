// HBM calibration kernels (the STREAM-style copy / read / write of the
// reference's tests/perf/bench_stream.cu:53-101), dwordx4 per lane.
// They give the achievable-bandwidth line next to the stencil's roofline
// fraction and the known-size copy that scales rocprof's FETCH/WRITE counters.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "stencil_capi.h"

namespace {

constexpr int kThreads = 256;

// Copy, write and triad run one float4 per thread over a grid that covers the
// array once, with non-temporal accesses: on MI355X that reaches ~6.6 TB/s
// for a copy, well above a grid-stride loop over a capped grid.
__global__ __launch_bounds__(kThreads) void copy_kernel(const float4* __restrict__ src,
                                                        float4* __restrict__ dst, size_t n4) {
    const size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n4) stencil::dev::store_stream(dst + i, stencil::dev::load_stream(src + i));
}

// Grid-stride over at most 2048 workgroups (one partial sum each), four
// independent loads in flight per lane.
__global__ __launch_bounds__(kThreads) void read_kernel(const float4* __restrict__ src,
                                                        float* __restrict__ sink, size_t n4) {
    const size_t stride = size_t(gridDim.x) * kThreads;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float4 v = stencil::dev::load_stream(src + i + u * stride);
            acc[u] += (v.x + v.y) + (v.z + v.w);
        }
    }
    for (; i < n4; i += stride) {
        const float4 v = stencil::dev::load_stream(src + i);
        acc[0] += (v.x + v.y) + (v.z + v.w);
    }
    // One store per workgroup keeps the loads live without measurable writes.
    __shared__ float red[kThreads];
    red[threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) sink[blockIdx.x] = red[0];
}

// dst = a + b + c over three n-float arrays laid out back to back at src: the
// stencil's 3-read / 1-write byte mix with no reuse.
__global__ __launch_bounds__(kThreads) void triad3_kernel(const float4* __restrict__ src,
                                                          float4* __restrict__ dst, size_t n4) {
    const size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n4) return;
    const float4 a = stencil::dev::load_stream(src + i);
    const float4 b = stencil::dev::load_stream(src + n4 + i);
    const float4 c = stencil::dev::load_stream(src + 2 * n4 + i);
    stencil::dev::store_stream(dst + i, make_float4(a.x + b.x + c.x, a.y + b.y + c.y, a.z + b.z + c.z,
                                                   a.w + b.w + c.w));
}

__global__ __launch_bounds__(kThreads) void write_kernel(float4* __restrict__ dst, float value,
                                                         size_t n4) {
    const size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n4) stencil::dev::store_stream(dst + i, make_float4(value, value, value, value));
}

unsigned one_shot_grid(size_t n4) { return static_cast<unsigned>((n4 + kThreads - 1) / kThreads); }

unsigned capped_grid(size_t n4) {
    const size_t want = 2048;
    const size_t need = (n4 + kThreads - 1) / kThreads;
    return static_cast<unsigned>(need < want ? (need ? need : 1) : want);
}

}  // namespace

extern "C" {

int stencil_stream_f32(int op, float* dst, const float* src, uint64_t n, void* stream) {
    if (n == 0 || n % 4 != 0) return STENCIL_ERR_INVALID;
    if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16 != 0)
        return STENCIL_ERR_INVALID;
    const size_t n4 = n / 4;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // one float4 per thread: grid.x must stay below 2^31
    if ((n4 + kThreads - 1) / kThreads > 0x7fffffffu) return STENCIL_ERR_INVALID;
    const unsigned g = one_shot_grid(n4);
    switch (op) {
        case 0:  // copy: src -> dst
            if (!dst || !src || dst == src) return STENCIL_ERR_INVALID;
            hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(kThreads), 0, st,
                               reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n4);
            break;
        case 1:  // read: src, one partial sum per workgroup into dst[0..grid)
            if (!dst || !src) return STENCIL_ERR_INVALID;
            hipLaunchKernelGGL(read_kernel, dim3(capped_grid(n4)), dim3(kThreads), 0, st,
                               reinterpret_cast<const float4*>(src), dst, n4);
            break;
        case 3:  // triad3: dst = src[0:n] + src[n:2n] + src[2n:3n]
            if (!dst || !src) return STENCIL_ERR_INVALID;
            hipLaunchKernelGGL(triad3_kernel, dim3(g), dim3(kThreads), 0, st,
                               reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n4);
            break;
        case 2:  // write: dst := 1.0
            if (!dst) return STENCIL_ERR_INVALID;
            hipLaunchKernelGGL(write_kernel, dim3(g), dim3(kThreads), 0, st,
                               reinterpret_cast<float4*>(dst), 1.0f, n4);
            break;
        default:
            return STENCIL_ERR_INVALID;
    }
    return hipGetLastError() == hipSuccess ? STENCIL_OK : STENCIL_ERR_HIP;
}

}  // extern "C"
