// Coalesced global-memory stencil kernels for gfx950 (the "naive" variant).
//
// One thread per interior point, wave64 rows along x.  Threads are mapped onto
// the interior only (the reference maps the full grid and returns early in the
// halo, src/kernels/stencil_naive.cu:105-111).  Arithmetic is the reference's
// association, op for op (src/kernels/stencil_naive.cu:75-88), compiled with
// -ffp-contract=off, so the FP32 result is bit-identical to the reference CPU
// path built without FMA contraction.  Neighbour re-reads are left to L1/L2.
#include "device_common.hpp"
#include "launchers.hpp"

namespace stencil::hipk {
namespace {

using dev::Fd;
using dev::kR;

constexpr int kNX = 64;  // one wave64 per x-row segment
constexpr int kNY = 4;

template <class T>
__device__ __forceinline__ T lap_axis_exact(const T* __restrict__ p, size_t c, size_t s, T ctr,
                                            T inv) {
    T lap = Fd<T>::c0 * ctr * inv;
#pragma unroll
    for (int r = 1; r <= kR; ++r)
        lap += Fd<T>::c(r) * (p[c + r * s] + p[c - r * s]) * inv;
    return lap;
}

template <class T>
__global__ __launch_bounds__(kNX* kNY) void wave_naive_kernel(
    const T* __restrict__ p_cur, const T* __restrict__ p_prev, const T* __restrict__ vel2,
    T* __restrict__ p_next, int nx, int ny, Scalars<T> s) {
    const int ix = kR + static_cast<int>(blockIdx.x * kNX + threadIdx.x);
    const int iy = kR + static_cast<int>(blockIdx.y * kNY + threadIdx.y);
    const int iz = kR + static_cast<int>(blockIdx.z);
    if (ix >= nx - kR || iy >= ny - kR) return;

    const size_t sy = static_cast<size_t>(nx);
    const size_t sz = sy * static_cast<size_t>(ny);
    const size_t c = static_cast<size_t>(iz) * sz + static_cast<size_t>(iy) * sy + ix;
    const T p = p_cur[c];
    const T lap_x = lap_axis_exact(p_cur, c, 1, p, s.inv_dx2);
    const T lap_y = lap_axis_exact(p_cur, c, sy, p, s.inv_dy2);
    const T lap_z = lap_axis_exact(p_cur, c, sz, p, s.inv_dz2);
    p_next[c] = T(2) * p - p_prev[c] + s.dt2 * vel2[c] * (lap_x + lap_y + lap_z);
}

// Heat: the reference runs the wave kernel with p_prev == p_cur and
// vel2 == alpha (src/solver/heat_diffusion.cpp:79-81,103-121), i.e.
// (2T - T) + (dt2*alpha) * lap; 2T - T == T exactly, so this is that value.
template <class T>
__global__ __launch_bounds__(kNX* kNY) void heat_naive_kernel(const T* __restrict__ t,
                                                              T* __restrict__ tn, int nx, int ny,
                                                              Scalars<T> s, T coef) {
    const int ix = kR + static_cast<int>(blockIdx.x * kNX + threadIdx.x);
    const int iy = kR + static_cast<int>(blockIdx.y * kNY + threadIdx.y);
    const int iz = kR + static_cast<int>(blockIdx.z);
    if (ix >= nx - kR || iy >= ny - kR) return;

    const size_t sy = static_cast<size_t>(nx);
    const size_t sz = sy * static_cast<size_t>(ny);
    const size_t c = static_cast<size_t>(iz) * sz + static_cast<size_t>(iy) * sy + ix;
    const T p = t[c];
    const T lap_x = lap_axis_exact(t, c, 1, p, s.inv_dx2);
    const T lap_y = lap_axis_exact(t, c, sy, p, s.inv_dy2);
    const T lap_z = lap_axis_exact(t, c, sz, p, s.inv_dz2);
    tn[c] = p + coef * (lap_x + lap_y + lap_z);
}

inline dim3 naive_grid(const Dims& d) {
    const int64_t ix = d.nx - 2 * kR, iy = d.ny - 2 * kR, iz = d.nz - 2 * kR;
    return dim3(static_cast<unsigned>((ix + kNX - 1) / kNX),
                static_cast<unsigned>((iy + kNY - 1) / kNY), static_cast<unsigned>(iz));
}

inline bool has_interior(const Dims& d) {
    return d.nx > 2 * kR && d.ny > 2 * kR && d.nz > 2 * kR;
}

template <class T>
hipError_t wave_naive_t(const T* cur, const T* prev, const T* vel2, T* next, const Dims& d,
                        const Scalars<T>& s, hipStream_t st) {
    if (!has_interior(d)) return hipSuccess;
    hipLaunchKernelGGL(wave_naive_kernel<T>, naive_grid(d), dim3(kNX, kNY, 1), 0, st, cur, prev,
                       vel2, next, static_cast<int>(d.nx), static_cast<int>(d.ny), s);
    return hipGetLastError();
}

}  // namespace

hipError_t wave_naive_f32(const float* cur, const float* prev, const float* vel2, float* next,
                          const Dims& d, const Scalars<float>& s, hipStream_t st) {
    return wave_naive_t(cur, prev, vel2, next, d, s, st);
}

hipError_t wave_naive_f64(const double* cur, const double* prev, const double* vel2,
                          double* next, const Dims& d, const Scalars<double>& s,
                          hipStream_t st) {
    return wave_naive_t(cur, prev, vel2, next, d, s, st);
}

hipError_t heat_naive_f32(const float* t, float* tn, const Dims& d, const Scalars<float>& s,
                          float coef, hipStream_t st) {
    if (!has_interior(d)) return hipSuccess;
    hipLaunchKernelGGL(heat_naive_kernel<float>, naive_grid(d), dim3(kNX, kNY, 1), 0, st, t, tn,
                       static_cast<int>(d.nx), static_cast<int>(d.ny), s, coef);
    return hipGetLastError();
}

hipError_t heat_naive_f64(const double* t, double* tn, const Dims& d, const Scalars<double>& s,
                          double coef, hipStream_t st) {
    if (!has_interior(d)) return hipSuccess;
    hipLaunchKernelGGL(heat_naive_kernel<double>, naive_grid(d), dim3(kNX, kNY, 1), 0, st, t,
                       tn, static_cast<int>(d.nx), static_cast<int>(d.ny), s, coef);
    return hipGetLastError();
}

}  // namespace stencil::hipk
