// Device-side constants and helpers shared by the gfx950 stencil kernels.
//
// All kernels are compiled with -ffp-contract=off: the "exact" forms restate
// the reference association op for op (src/kernels/stencil_naive.cu:75-88),
// and the "fused" forms spell out every fma explicitly.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace stencil::dev {

constexpr int kR = 4;  // stencil radius (8th order), include/stencil/types.hpp:42-44

// FDCoeffs<4> (reference include/stencil/types.hpp:175-181): the double
// rationals narrowed to T.
template <class T>
struct Fd {
    static constexpr T c0 = static_cast<T>(-205.0 / 72.0);
    static constexpr T c1 = static_cast<T>(8.0 / 5.0);
    static constexpr T c2 = static_cast<T>(-1.0 / 5.0);
    static constexpr T c3 = static_cast<T>(8.0 / 315.0);
    static constexpr T c4 = static_cast<T>(-1.0 / 560.0);
    static __device__ __forceinline__ constexpr T c(int r) {
        return r == 0 ? c0 : r == 1 ? c1 : r == 2 ? c2 : r == 3 ? c3 : c4;
    }
};

// Streaming accesses that are touched once per step (p_prev, vel2, p_next):
// non-temporal, so they do not evict the p_cur planes whose halos are re-read.
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 load_stream(const float4* p) {
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store_stream(float4* p, float4 v) {
    const v4f w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<v4f*>(p));
}

__device__ __forceinline__ float comp(const float4& v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}
__device__ __forceinline__ void set_comp(float4& v, int i, float f) {
    if (i == 0) v.x = f;
    else if (i == 1) v.y = f;
    else if (i == 2) v.z = f;
    else v.w = f;
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// (workgroup b runs on XCD b % 8), and each XCD has its own L2.  Remapping
// b -> (b % 8) * (n / 8) + b / 8 hands every XCD one contiguous range of
// logical tiles, so tiles that share halo rows share an L2.  Needs n % 8 == 0.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned n, bool on) {
    return on ? (b & 7u) * (n >> 3) + (b >> 3) : b;
}

}  // namespace stencil::dev
