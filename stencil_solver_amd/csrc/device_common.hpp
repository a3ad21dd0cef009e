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

typedef unsigned v2u __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint2 load_stream_u2(const uint2* p) {
    const v2u v = __builtin_nontemporal_load(reinterpret_cast<const v2u*>(p));
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void store_stream_u2(uint2* p, uint2 v) {
    const v2u w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<v2u*>(p));
}

// ── 16-bit field storage (arithmetic stays FP32; decoding is exact) ─────────
// kBf16:  bfloat16, round to nearest even on store.
// kFp16s: IEEE half of value * scale, scale a power of two per field.
// kBfp16: block floating point: int16 mantissas, one shared power-of-two
//         exponent e (int8) per 16 x-consecutive values: v = m * 2^(e-15),
//         e = frexp exponent of the block's max |v|, |m| <= 32767.
enum Fmt : int { kF32 = 0, kBf16 = 1, kFp16s = 2, kBfp16 = 3 };

__device__ __forceinline__ float bf16_to_float(unsigned h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ unsigned float_to_bf16(float f) {
    const unsigned b = __float_as_uint(f);
    return (b + 0x7FFFu + ((b >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float4 bf16x4_to_float4(uint2 u) {
    return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xFFFF0000u),
                       __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xFFFF0000u));
}
__device__ __forceinline__ uint2 float4_to_bf16x4(float4 v) {
    return make_uint2(float_to_bf16(v.x) | (float_to_bf16(v.y) << 16),
                      float_to_bf16(v.z) | (float_to_bf16(v.w) << 16));
}

__device__ __forceinline__ float half_bits_to_float(unsigned h) {
    const unsigned short s = static_cast<unsigned short>(h);
    _Float16 f;
    __builtin_memcpy(&f, &s, 2);
    return static_cast<float>(f);
}
__device__ __forceinline__ unsigned float_to_half_bits(float f) {
    const _Float16 h = static_cast<_Float16>(f);
    unsigned short s;
    __builtin_memcpy(&s, &h, 2);
    return s;
}
__device__ __forceinline__ float4 fp16x4_to_float4(uint2 u, float inv) {
    return make_float4(half_bits_to_float(u.x & 0xFFFFu) * inv, half_bits_to_float(u.x >> 16) * inv,
                       half_bits_to_float(u.y & 0xFFFFu) * inv, half_bits_to_float(u.y >> 16) * inv);
}
__device__ __forceinline__ uint2 float4_to_fp16x4(float4 v, float scale) {
    return make_uint2(float_to_half_bits(v.x * scale) | (float_to_half_bits(v.y * scale) << 16),
                      float_to_half_bits(v.z * scale) | (float_to_half_bits(v.w * scale) << 16));
}

__device__ __forceinline__ float bfp16_value(unsigned m16, int e) {
    return __builtin_ldexpf(static_cast<float>(static_cast<short>(m16)), e - 15);
}
__device__ __forceinline__ float4 bfp16x4_to_float4(uint2 u, int e) {
    return make_float4(bfp16_value(u.x & 0xFFFFu, e), bfp16_value(u.x >> 16, e),
                       bfp16_value(u.y & 0xFFFFu, e), bfp16_value(u.y >> 16, e));
}
__device__ __forceinline__ unsigned bfp16_mantissa(float v, int e) {
    float q = __builtin_rintf(__builtin_ldexpf(v, 15 - e));
    q = q > 32767.f ? 32767.f : (q < -32767.f ? -32767.f : q);
    return static_cast<unsigned>(static_cast<unsigned short>(static_cast<short>(q)));
}
__device__ __forceinline__ uint2 float4_to_bfp16x4(float4 v, int e) {
    return make_uint2(bfp16_mantissa(v.x, e) | (bfp16_mantissa(v.y, e) << 16),
                      bfp16_mantissa(v.z, e) | (bfp16_mantissa(v.w, e) << 16));
}
// Shared exponent of a block whose max |v| is m: frexp exponent (m < 2^e),
// clamped to int8; 0 for an all-zero block.
__device__ __forceinline__ int bfp16_exponent(float m) {
    if (!(m > 0.f)) return 0;
    int e = __builtin_amdgcn_frexp_expf(m);
    return e < -127 ? -127 : (e > 127 ? 127 : e);
}
__device__ __forceinline__ float max_abs4(float4 v) {
    return fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w)));
}
// Max over the 4 lanes of a quad (DPP quad_perm [1,0,3,2] then [2,3,0,1]);
// m >= 0, so the int ordering of the bits is the float ordering.
__device__ __forceinline__ float quad_max(float m) {
    int v = __float_as_int(m);
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));
    return __int_as_float(v);
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
