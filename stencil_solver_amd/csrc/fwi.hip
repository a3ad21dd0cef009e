// Kernels of the adjoint-state FWI gradient (no reference counterpart: the
// reference has forward modelling only).
//
// Discrete problem (m = v^2 per cell, L the 8th-order Laplacian with zero
// walls, b^n a point source, R the receiver-plane sampling):
//   p^{n+1} = 2 p^n - p^{n-1} + dt^2 m (L p^n) + b^n,   p^0 = p^-1 = 0
//   J = 1/2 sum_{k=1..N} |R p^k - d^k|^2
// Its exact discrete adjoint (L is symmetric: interior-to-interior with zero
// walls), lambda^{N+1} = lambda^{N+2} = 0:
//   lambda^k = 2 lambda^{k+1} - lambda^{k+2} + dt^2 L(m lambda^{k+1}) + R^T (R p^k - d^k)
//   dJ/dm    = sum_k dt^2 (L p^k) lambda^{k+1}
// is run in the variable psi = m lambda, which turns it into a step of the
// forward's shape (no m inside the Laplacian, so no m at the halo):
//   psi^k = 2 psi^{k+1} - psi^{k+2} + dt^2 m L psi^{k+1} + m R^T (R p^k - d^k)
//   dJ/dm = (sum_k dt^2 (L p^k) psi^{k+1}) / m
// The forward step here is the reference acoustic step (same association as
// src/kernels/stencil_naive.cu:75-88) plus the source term, and can also write
// L p^n for the gradient.  One thread per interior point; neighbour re-reads
// are L1/L2 hits.
#include <algorithm>

#include "device_common.hpp"
#include "launchers.hpp"

namespace stencil::hipk {
namespace {

using dev::Fd;
using dev::kR;

constexpr int kNX = 64;
constexpr int kNY = 4;

template <class T>
__device__ __forceinline__ T lap_axis(const T* __restrict__ p, size_t c, size_t s, T ctr, T inv) {
    T lap = Fd<T>::c0 * ctr * inv;
#pragma unroll
    for (int r = 1; r <= kR; ++r) lap += Fd<T>::c(r) * (p[c + r * s] + p[c - r * s]) * inv;
    return lap;
}

struct Cell {
    int ix, iy, iz;
    size_t c, sy, sz;
};

__device__ __forceinline__ bool interior_cell(int nx, int ny, Cell& k) {
    k.ix = kR + static_cast<int>(blockIdx.x * kNX + threadIdx.x);
    k.iy = kR + static_cast<int>(blockIdx.y * kNY + threadIdx.y);
    k.iz = kR + static_cast<int>(blockIdx.z);
    if (k.ix >= nx - kR || k.iy >= ny - kR) return false;
    k.sy = static_cast<size_t>(nx);
    k.sz = k.sy * static_cast<size_t>(ny);
    k.c = static_cast<size_t>(k.iz) * k.sz + static_cast<size_t>(k.iy) * k.sy + k.ix;
    return true;
}

// p_next = 2 p - p_prev + dt^2 m L p  (+ wavelet[n] at cell src);  lap_out = L p.
template <class T>
__global__ __launch_bounds__(kNX* kNY) void fwd_step_kernel(const T* __restrict__ p, const T* __restrict__ pp,
                                                            const T* __restrict__ m, T* __restrict__ pn,
                                                            T* __restrict__ lap_out, int nx, int ny,
                                                            Scalars<T> s, long long src,
                                                            const T* __restrict__ wavelet, int n) {
    Cell k;
    if (!interior_cell(nx, ny, k)) return;
    const T pc = p[k.c];
    const T lx = lap_axis(p, k.c, 1, pc, s.inv_dx2);
    const T ly = lap_axis(p, k.c, k.sy, pc, s.inv_dy2);
    const T lz = lap_axis(p, k.c, k.sz, pc, s.inv_dz2);
    const T lap = lx + ly + lz;
    T v = T(2) * pc - pp[k.c] + s.dt2 * m[k.c] * lap;
    if (static_cast<long long>(k.c) == src) v += wavelet[n];
    pn[k.c] = v;
    if (lap_out) lap_out[k.c] = lap;
}

// psi_out = 2 psi1 - psi2 + dt^2 m L psi1  (+ m res[y][x] on plane rec_z);
// grad += dt^2 lap * psi1 (lap = L p^k of the same k; the sum is divided by m
// once at the end), preceded, when lap2 = L p^{k+1} is given, by step k+1's
// term dt^2 lap2 * psi2 (psi2 = psi^{k+2}): the two additions in the order and
// rounding of the two steps applying one each, with one gradient read and
// write; trace[0] = psi_out / m at src (the adjoint field there).
template <class T>
__global__ __launch_bounds__(kNX* kNY) void adj_step_kernel(const T* __restrict__ psi1, const T* __restrict__ psi2,
                                                            const T* __restrict__ m, const T* __restrict__ lap,
                                                            const T* __restrict__ lap2, T* __restrict__ grad,
                                                            T* __restrict__ psi_out, const T* __restrict__ res,
                                                            int rec_z, int nx, int ny, Scalars<T> s, long long src,
                                                            T* __restrict__ trace) {
    Cell k;
    if (!interior_cell(nx, ny, k)) return;
    const T pc = psi1[k.c];
    const T lx = lap_axis(psi1, k.c, 1, pc, s.inv_dx2);
    const T ly = lap_axis(psi1, k.c, k.sy, pc, s.inv_dy2);
    const T lz = lap_axis(psi1, k.c, k.sz, pc, s.inv_dz2);
    const T lp = lx + ly + lz;
    const T mc = m[k.c];
    const T p2 = psi2[k.c];
    T v = T(2) * pc - p2 + s.dt2 * mc * lp;
    if (res && k.iz == rec_z) v += mc * res[static_cast<size_t>(k.iy) * k.sy + k.ix];
    psi_out[k.c] = v;
    if (lap) {
        T g = grad[k.c];
        if (lap2) g += s.dt2 * lap2[k.c] * p2;
        g += s.dt2 * lap[k.c] * pc;
        grad[k.c] = g;
    }
    if (trace && static_cast<long long>(k.c) == src) trace[0] = v / mc;
}

// grad = acc / m on interior cells (the halo keeps its zeros).
template <class T>
__global__ __launch_bounds__(kNX* kNY) void divide_kernel(T* __restrict__ g, const T* __restrict__ m, int nx, int ny) {
    Cell k;
    if (!interior_cell(nx, ny, k)) return;
    g[k.c] = g[k.c] / m[k.c];
}

inline dim3 grid_of(const Dims& d) {
    const int64_t ix = d.nx - 2 * kR, iy = d.ny - 2 * kR, iz = d.nz - 2 * kR;
    return dim3(static_cast<unsigned>((ix + kNX - 1) / kNX), static_cast<unsigned>((iy + kNY - 1) / kNY),
                static_cast<unsigned>(iz));
}

// res = syn - obs over n values; partial[b] = sum of res^2 (FP64) over block b's range.
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

template <class T>
__global__ __launch_bounds__(kRedThreads) void residual_kernel(const T* __restrict__ syn, const T* __restrict__ obs,
                                                               T* __restrict__ res, size_t n,
                                                               double* __restrict__ partial) {
    __shared__ double sh[kRedThreads];
    double acc = 0.0;
    for (size_t i = blockIdx.x * size_t(kRedThreads) + threadIdx.x; i < n; i += size_t(gridDim.x) * kRedThreads) {
        const T r = syn[i] - obs[i];
        res[i] = r;
        acc += static_cast<double>(r) * static_cast<double>(r);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out[0] = scale * sum(partial[0..nb)), fixed order.
__global__ __launch_bounds__(kRedThreads) void sum_partials_kernel(const double* __restrict__ partial, int nb,
                                                                   double scale, double* __restrict__ out) {
    __shared__ double sh[kRedThreads];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += kRedThreads) acc += partial[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = scale * sh[0];
}

template <class T>
__global__ void fill_kernel(T* __restrict__ p, size_t n, T v) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

template <class T>
hipError_t fwd_step_t(const T* p, const T* pp, const T* m, T* pn, T* lap_out, const Dims& d, const Scalars<T>& s,
                      long long src, const T* wavelet, int n, hipStream_t st) {
    if (d.nx <= 2 * kR || d.ny <= 2 * kR || d.nz <= 2 * kR) return hipSuccess;
    hipLaunchKernelGGL(fwd_step_kernel<T>, grid_of(d), dim3(kNX, kNY, 1), 0, st, p, pp, m, pn, lap_out,
                       static_cast<int>(d.nx), static_cast<int>(d.ny), s, src, wavelet, n);
    return hipGetLastError();
}

template <class T>
hipError_t adj_step_t(const T* psi1, const T* psi2, const T* m, const T* lap, const T* lap2, T* grad, T* psi_out,
                      const T* res, int rec_z, const Dims& d, const Scalars<T>& s, long long src, T* trace,
                      hipStream_t st) {
    if (d.nx <= 2 * kR || d.ny <= 2 * kR || d.nz <= 2 * kR) return hipSuccess;
    if (lap2 && !lap) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adj_step_kernel<T>, grid_of(d), dim3(kNX, kNY, 1), 0, st, psi1, psi2, m, lap, lap2, grad,
                       psi_out, res, rec_z, static_cast<int>(d.nx), static_cast<int>(d.ny), s, src, trace);
    return hipGetLastError();
}

template <class T>
hipError_t divide_t(T* g, const T* m, const Dims& d, hipStream_t st) {
    if (d.nx <= 2 * kR || d.ny <= 2 * kR || d.nz <= 2 * kR) return hipSuccess;
    hipLaunchKernelGGL(divide_kernel<T>, grid_of(d), dim3(kNX, kNY, 1), 0, st, g, m, static_cast<int>(d.nx),
                       static_cast<int>(d.ny));
    return hipGetLastError();
}

template <class T>
hipError_t residual_t(const T* syn, const T* obs, T* res, size_t n, double* partial, double* out, hipStream_t st) {
    hipLaunchKernelGGL(residual_kernel<T>, dim3(kRedBlocks), dim3(kRedThreads), 0, st, syn, obs, res, n, partial);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kRedThreads), 0, st, partial, kRedBlocks, 0.5, out);
    return hipGetLastError();
}

template <class T>
hipError_t fill_t(T* p, size_t n, T v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(fill_kernel<T>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

}  // namespace

hipError_t fwi_fwd_step_f32(const float* p, const float* pp, const float* m, float* pn, float* lap_out,
                            const Dims& d, const Scalars<float>& s, long long src, const float* wavelet, int n,
                            hipStream_t st) {
    return fwd_step_t(p, pp, m, pn, lap_out, d, s, src, wavelet, n, st);
}
hipError_t fwi_fwd_step_f64(const double* p, const double* pp, const double* m, double* pn, double* lap_out,
                            const Dims& d, const Scalars<double>& s, long long src, const double* wavelet, int n,
                            hipStream_t st) {
    return fwd_step_t(p, pp, m, pn, lap_out, d, s, src, wavelet, n, st);
}
hipError_t fwi_adj_step_f32(const float* psi1, const float* psi2, const float* m, const float* lap, const float* lap2,
                            float* grad, float* psi_out, const float* res, int rec_z, const Dims& d,
                            const Scalars<float>& s, long long src, float* trace, hipStream_t st) {
    return adj_step_t(psi1, psi2, m, lap, lap2, grad, psi_out, res, rec_z, d, s, src, trace, st);
}
hipError_t fwi_adj_step_f64(const double* psi1, const double* psi2, const double* m, const double* lap,
                            const double* lap2, double* grad, double* psi_out, const double* res, int rec_z,
                            const Dims& d, const Scalars<double>& s, long long src, double* trace, hipStream_t st) {
    return adj_step_t(psi1, psi2, m, lap, lap2, grad, psi_out, res, rec_z, d, s, src, trace, st);
}
hipError_t fwi_divide_f32(float* g, const float* m, const Dims& d, hipStream_t st) { return divide_t(g, m, d, st); }
hipError_t fwi_divide_f64(double* g, const double* m, const Dims& d, hipStream_t st) { return divide_t(g, m, d, st); }
hipError_t fwi_residual_f32(const float* syn, const float* obs, float* res, size_t n, double* partial, double* out,
                            hipStream_t st) {
    return residual_t(syn, obs, res, n, partial, out, st);
}
hipError_t fwi_residual_f64(const double* syn, const double* obs, double* res, size_t n, double* partial,
                            double* out, hipStream_t st) {
    return residual_t(syn, obs, res, n, partial, out, st);
}
hipError_t fill_f32(float* p, size_t n, float v, hipStream_t st) { return fill_t(p, n, v, st); }
hipError_t fill_f64(double* p, size_t n, double v, hipStream_t st) { return fill_t(p, n, v, st); }

}  // namespace stencil::hipk
