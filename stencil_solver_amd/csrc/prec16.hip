// 16-bit field conversion kernels and the FP64 reductions used to measure the errors of the
// reduced-precision and adjoint paths (energy norm, relative L2 against an
// FP64 run, dot products).
//
// Reductions are deterministic: a fixed grid of kRedBlocks workgroups, each
// thread sums its grid-strided interior points in order in double, a tree
// reduction per workgroup writes one partial per workgroup into the caller's
// workspace, and one more workgroup sums the partials in a fixed order.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "device_common.hpp"
#include "launchers.hpp"
#include "stencil_capi.h"

namespace stencil::hipk {
namespace {

constexpr int kThreads = 256;

template <int FMT>
__global__ __launch_bounds__(kThreads) void encode16_kernel(const float* __restrict__ src, void* dst,
                                                            int8_t* dst_e, size_t n4, float scale) {
    const size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    // Every lane of a quad must reach the block-exponent reduction.
    const bool in = i < n4;
    const float4 v = in ? reinterpret_cast<const float4*>(src)[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    uint2 u;
    if constexpr (FMT == dev::kBf16) {
        u = dev::float4_to_bf16x4(v);
    } else if constexpr (FMT == dev::kFp16s) {
        u = dev::float4_to_fp16x4(v, scale);
    } else {
        const int e = dev::bfp16_exponent(dev::quad_max(dev::max_abs4(v)));
        u = dev::float4_to_bfp16x4(v, e);
        if (in && (threadIdx.x & 3) == 0) dst_e[i >> 2] = static_cast<int8_t>(e);
    }
    if (in) reinterpret_cast<uint2*>(dst)[i] = u;
}

template <int FMT>
__global__ __launch_bounds__(kThreads) void decode16_kernel(const void* src, const int8_t* src_e,
                                                            float* __restrict__ dst, size_t n4,
                                                            float inv) {
    const size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n4) return;
    const uint2 u = reinterpret_cast<const uint2*>(src)[i];
    float4 v;
    if constexpr (FMT == dev::kBf16) v = dev::bf16x4_to_float4(u);
    else if constexpr (FMT == dev::kFp16s) v = dev::fp16x4_to_float4(u, inv);
    else v = dev::bfp16x4_to_float4(u, src_e[i >> 2]);
    reinterpret_cast<float4*>(dst)[i] = v;
}

// ── deterministic FP64 reductions over the interior ─────────────────────────
constexpr int kRedBlocks = 1024;
constexpr int kMaxSums = 3;

struct RedArgs {
    const void* a;
    const void* b;
    const void* c;
    int nx, ny, nz;
    double inv_dx2, inv_dy2, inv_dz2, inv_dt2;
};

template <class T>
__device__ __forceinline__ double at(const void* p, size_t i) {
    return static_cast<double>(static_cast<const T*>(p)[i]);
}

// Per-point terms.  kind 0: energy of the leapfrog pair (a = p^n, b = p^{n+1},
// c = vel2): sums[0] = (b-a)^2 / (v^2 dt^2), sums[1] = -b * L a.  kind 1:
// sums[0] = (a-b)^2, sums[1] = b^2 (a FP32, b FP64).  kind 2: sums[0] = a*b.
template <int KIND, class TA, class TB>
__device__ __forceinline__ void point_terms(const RedArgs& r, size_t i, size_t sy, size_t sz,
                                            double* acc) {
    if constexpr (KIND == 0) {
        const double p0 = at<TA>(r.a, i), p1 = at<TA>(r.b, i), v2 = at<TB>(r.c, i);
        const double d = p1 - p0;
        acc[0] += d * d * r.inv_dt2 / v2;
        // Lap p^n in double with the 8th-order weights.
        const double c[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
        double lx = c[0] * p0, ly = c[0] * p0, lz = c[0] * p0;
#pragma unroll
        for (int k = 1; k <= 4; ++k) {
            lx += c[k] * (at<TA>(r.a, i + k) + at<TA>(r.a, i - k));
            ly += c[k] * (at<TA>(r.a, i + k * sy) + at<TA>(r.a, i - k * sy));
            lz += c[k] * (at<TA>(r.a, i + k * sz) + at<TA>(r.a, i - k * sz));
        }
        acc[1] -= p1 * (lx * r.inv_dx2 + ly * r.inv_dy2 + lz * r.inv_dz2);
    } else if constexpr (KIND == 1) {
        const double x = at<TA>(r.a, i), y = at<TB>(r.b, i);
        acc[0] += (x - y) * (x - y);
        acc[1] += y * y;
    } else {
        acc[0] += at<TA>(r.a, i) * at<TB>(r.b, i);
    }
}

template <int KIND, class TA, class TB>
__global__ __launch_bounds__(kThreads) void reduce_partial(RedArgs r, double* part) {
    const int ix = r.nx - 8, iy = r.ny - 8, iz = r.nz - 8;
    const size_t n = size_t(ix) * iy * iz;
    const size_t sy = size_t(r.nx), sz = size_t(r.nx) * r.ny;
    double acc[kMaxSums] = {0.0, 0.0, 0.0};
    for (size_t k = size_t(blockIdx.x) * kThreads + threadIdx.x; k < n;
         k += size_t(kRedBlocks) * kThreads) {
        const size_t x = k % ix, y = (k / ix) % iy, z = k / (size_t(ix) * iy);
        const size_t i = (z + 4) * sz + (y + 4) * sy + (x + 4);
        point_terms<KIND, TA, TB>(r, i, sy, sz, acc);
    }
    __shared__ double red[kMaxSums][kThreads];
#pragma unroll
    for (int s = 0; s < kMaxSums; ++s) red[s][threadIdx.x] = acc[s];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
#pragma unroll
            for (int s = 0; s < kMaxSums; ++s) red[s][threadIdx.x] += red[s][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int s = 0; s < kMaxSums; ++s) part[s * kRedBlocks + blockIdx.x] = red[s][0];
}

__global__ __launch_bounds__(kThreads) void reduce_final(const double* part, double* out) {
    __shared__ double red[kMaxSums][kThreads];
#pragma unroll
    for (int s = 0; s < kMaxSums; ++s) {
        double v = 0.0;
        for (int k = threadIdx.x; k < kRedBlocks; k += kThreads) v += part[s * kRedBlocks + k];
        red[s][threadIdx.x] = v;
    }
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
#pragma unroll
            for (int s = 0; s < kMaxSums; ++s) red[s][threadIdx.x] += red[s][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int s = 0; s < kMaxSums; ++s) out[s] = red[s][0];
}

template <int KIND, class TA, class TB>
hipError_t reduce(const RedArgs& r, double* work, double* out, hipStream_t st) {
    hipLaunchKernelGGL((reduce_partial<KIND, TA, TB>), dim3(kRedBlocks), dim3(kThreads), 0, st, r, work);
    hipLaunchKernelGGL(reduce_final, dim3(1), dim3(kThreads), 0, st, work, out);
    return hipGetLastError();
}

RedArgs red_args(const stencil_grid* g, const void* a, const void* b, const void* c) {
    RedArgs r;
    r.a = a;
    r.b = b;
    r.c = c;
    r.nx = int(g->nx);
    r.ny = int(g->ny);
    r.nz = int(g->nz);
    r.inv_dx2 = 1.0 / (g->dx * g->dx);
    r.inv_dy2 = 1.0 / (g->dy * g->dy);
    r.inv_dz2 = 1.0 / (g->dz * g->dz);
    r.inv_dt2 = 1.0 / (g->dt * g->dt);
    return r;
}

bool red_grid_ok(const stencil_grid* g) {
    return g && g->nx > 8 && g->ny > 8 && g->nz > 8 && g->nx < (1u << 20) && g->ny < (1u << 20) &&
           g->nz < (1u << 20);
}

}  // namespace

hipError_t encode16(int fmt, const float* src, Field16 dst, size_t n, float scale, hipStream_t st) {
    const size_t n4 = n / 4;
    const dim3 grid(static_cast<unsigned>((n4 + kThreads - 1) / kThreads));
    switch (fmt) {
        case dev::kBf16:
            hipLaunchKernelGGL(encode16_kernel<dev::kBf16>, grid, dim3(kThreads), 0, st, src, dst.data, dst.exp, n4, scale);
            break;
        case dev::kFp16s:
            hipLaunchKernelGGL(encode16_kernel<dev::kFp16s>, grid, dim3(kThreads), 0, st, src, dst.data, dst.exp, n4, scale);
            break;
        case dev::kBfp16:
            hipLaunchKernelGGL(encode16_kernel<dev::kBfp16>, grid, dim3(kThreads), 0, st, src, dst.data, dst.exp, n4, scale);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t decode16(int fmt, Field16 src, float* dst, size_t n, float scale, hipStream_t st) {
    const size_t n4 = n / 4;
    const dim3 grid(static_cast<unsigned>((n4 + kThreads - 1) / kThreads));
    const float inv = 1.0f / scale;
    switch (fmt) {
        case dev::kBf16:
            hipLaunchKernelGGL(decode16_kernel<dev::kBf16>, grid, dim3(kThreads), 0, st, src.data, src.exp, dst, n4, inv);
            break;
        case dev::kFp16s:
            hipLaunchKernelGGL(decode16_kernel<dev::kFp16s>, grid, dim3(kThreads), 0, st, src.data, src.exp, dst, n4, inv);
            break;
        case dev::kBfp16:
            hipLaunchKernelGGL(decode16_kernel<dev::kBfp16>, grid, dim3(kThreads), 0, st, src.data, src.exp, dst, n4, inv);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace stencil::hipk

using namespace stencil::hipk;

// Partial sums of stencil_energy_f32_noalloc: kMaxSums x kRedBlocks doubles
// (24 KiB), one copy per device in the library's code object, held until the
// process ends.  In host code this name is a stand-in, not a device address:
// kernels get the per-device address that hipGetSymbolAddress returns.
static __device__ __align__(256) double scratch[kMaxSums * kRedBlocks];

namespace {
int status_of(hipError_t e) { return e == hipSuccess ? STENCIL_OK : STENCIL_ERR_HIP; }

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxScratchDevices] = {};

// The current device's copy of `scratch`, resolved on the first call on that
// device (with it current) and kept; no allocation is involved.
hipError_t scratch_of_current_device(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}
}  // namespace

extern "C" {

uint64_t stencil_reduce_workspace_bytes(void) { return sizeof(double) * kMaxSums * kRedBlocks; }

int stencil_energy_f32(const float* p_n, const float* p_np1, const float* vel2, const stencil_grid* grid,
                       double* out2, double* work, void* stream) {
    if (!p_n || !p_np1 || !vel2 || !out2 || !work || !red_grid_ok(grid)) return STENCIL_ERR_INVALID;
    return status_of(reduce<0, float, float>(red_args(grid, p_n, p_np1, vel2), work, out2,
                                             static_cast<hipStream_t>(stream)));
}

int stencil_energy_f32_noalloc(const float* p_n, const float* p_np1, const float* vel2,
                               const stencil_grid* grid, double* out2, void* stream) {
    if (!p_n || !p_np1 || !vel2 || !out2 || !red_grid_ok(grid)) return STENCIL_ERR_INVALID;
    static_assert(sizeof(scratch) == sizeof(double) * kMaxSums * kRedBlocks, "partials fill the scratch");
    double* part = nullptr;
    const hipError_t e = scratch_of_current_device(&part);
    if (e != hipSuccess) return hip_status(e);
    // Every workgroup of reduce_partial overwrites its own partials before
    // reduce_final reads them (stream order), so nothing needs clearing.
    return hip_status(reduce<0, float, float>(red_args(grid, p_n, p_np1, vel2), part, out2,
                                              static_cast<hipStream_t>(stream)));
}

int stencil_energy_f64(const double* p_n, const double* p_np1, const double* vel2,
                       const stencil_grid* grid, double* out2, double* work, void* stream) {
    if (!p_n || !p_np1 || !vel2 || !out2 || !work || !red_grid_ok(grid)) return STENCIL_ERR_INVALID;
    return status_of(reduce<0, double, double>(red_args(grid, p_n, p_np1, vel2), work, out2,
                                               static_cast<hipStream_t>(stream)));
}

int stencil_diff_norms_f32_f64(const float* a, const double* b, const stencil_grid* grid, double* out2,
                               double* work, void* stream) {
    if (!a || !b || !out2 || !work || !red_grid_ok(grid)) return STENCIL_ERR_INVALID;
    return status_of(reduce<1, float, double>(red_args(grid, a, b, nullptr), work, out2,
                                              static_cast<hipStream_t>(stream)));
}

int stencil_dot_f64(const double* a, const double* b, const stencil_grid* grid, double* out1,
                    double* work, void* stream) {
    if (!a || !b || !out1 || !work || !red_grid_ok(grid)) return STENCIL_ERR_INVALID;
    return status_of(reduce<2, double, double>(red_args(grid, a, b, nullptr), work, out1,
                                               static_cast<hipStream_t>(stream)));
}

int stencil_dot_f32(const float* a, const float* b, const stencil_grid* grid, double* out1,
                    double* work, void* stream) {
    if (!a || !b || !out1 || !work || !red_grid_ok(grid)) return STENCIL_ERR_INVALID;
    return status_of(reduce<2, float, float>(red_args(grid, a, b, nullptr), work, out1,
                                             static_cast<hipStream_t>(stream)));
}

}  // extern "C"
