// Device-side constants and helpers shared by the gfx950 stencil kernels.
//
// All kernels are compiled with -ffp-contract=off: the "exact" forms restate
// the reference association op for op (src/kernels/stencil_naive.cu:75-88),
// and the "fused" forms spell out every fma explicitly.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

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
// kBfp16: block floating point: int16 mantissas m, one int8 block scale k per
//         16 x-consecutive values: v = m * 2^k; k = e - 15 with e the frexp
//         exponent of fl(max|v| * (1 + 2^-15)) clamped to [-112, 127] (0 for
//         an all-zero block), so |m| <= 32767 and 2^-k is a normal float.
// Stores use gfx950's packed converts (v_cvt_pk_bf16_f32, v_cvt_pk_f16_f32:
// two values per instruction, round to nearest even); loads are one shift,
// mask or SDWA convert per value.
enum Fmt : int { kF32 = 0, kBf16 = 1, kFp16s = 2, kBfp16 = 3 };

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
    const f32x2 v = {a, b};
    const bf16x2 h = __builtin_convertvector(v, bf16x2);
    unsigned r;
    __builtin_memcpy(&r, &h, 4);
    return r;
}
__device__ __forceinline__ unsigned pack_fp16(float a, float b) {
    const f32x2 v = {a, b};
    const f16x2 h = __builtin_convertvector(v, f16x2);
    unsigned r;
    __builtin_memcpy(&r, &h, 4);
    return r;
}
__device__ __forceinline__ float half_lo(unsigned u) {
    const unsigned short s = static_cast<unsigned short>(u);
    _Float16 f;
    __builtin_memcpy(&f, &s, 2);
    return static_cast<float>(f);
}
__device__ __forceinline__ float half_hi(unsigned u) { return half_lo(u >> 16); }

__device__ __forceinline__ float4 bf16x4_to_float4(uint2 u) {
    return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xFFFF0000u),
                       __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xFFFF0000u));
}
__device__ __forceinline__ uint2 float4_to_bf16x4(float4 v) {
    return make_uint2(pack_bf16(v.x, v.y), pack_bf16(v.z, v.w));
}

// Half values as stored (still multiplied by the field's scale).
__device__ __forceinline__ float4 fp16x4_raw(uint2 u) {
    return make_float4(half_lo(u.x), half_hi(u.x), half_lo(u.y), half_hi(u.y));
}
__device__ __forceinline__ float4 fp16x4_to_float4(uint2 u, float inv) {
    const float4 r = fp16x4_raw(u);
    return make_float4(r.x * inv, r.y * inv, r.z * inv, r.w * inv);
}
__device__ __forceinline__ uint2 fp16x4_pack(float4 v) {
    return make_uint2(pack_fp16(v.x, v.y), pack_fp16(v.z, v.w));
}
__device__ __forceinline__ uint2 float4_to_fp16x4(float4 v, float scale) {
    return make_uint2(pack_fp16(v.x * scale, v.y * scale), pack_fp16(v.z * scale, v.w * scale));
}

// The 4 mantissas as floats (exact); the block scale 2^k is applied by the
// caller.
__device__ __forceinline__ float4 bfp16x4_mantissas(uint2 u) {
    return make_float4(static_cast<float>(static_cast<short>(u.x)), static_cast<float>(static_cast<short>(u.x >> 16)),
                       static_cast<float>(static_cast<short>(u.y)), static_cast<float>(static_cast<short>(u.y >> 16)));
}
__device__ __forceinline__ float4 bfp16x4_to_float4(uint2 u, int k) {
    const float4 m = bfp16x4_mantissas(u);
    return make_float4(__builtin_ldexpf(m.x, k), __builtin_ldexpf(m.y, k), __builtin_ldexpf(m.z, k),
                       __builtin_ldexpf(m.w, k));
}
// Mantissas of a block with scale k: rint(v * 2^-k), |v * 2^-k| < 32767.5 by
// the choice of k.  Adding 1.5 * 2^23 in the fma rounds v * 2^-k (exact, a
// power-of-two scaling) to an integer (ulp 1, ties to even) whose
// two's-complement low 16 bits are the mantissa.
__device__ __forceinline__ uint2 float4_to_bfp16x4(float4 v, int k) {
    const float s = __uint_as_float(static_cast<unsigned>(127 - k) << 23);  // 2^-k
    auto q = [s](float x) { return __float_as_uint(__builtin_fmaf(x, s, 12582912.0f)); };
    return make_uint2(__builtin_amdgcn_perm(q(v.y), q(v.x), 0x05040100u),
                      __builtin_amdgcn_perm(q(v.w), q(v.z), 0x05040100u));
}
// Block scale of a block whose max |v| is m >= 0 (v_frexp_exp gives 0 for 0).
__device__ __forceinline__ int bfp16_exponent(float m) {
    int e = __builtin_amdgcn_frexp_expf(m * (1.0f + 0x1p-15f));
    e = e < -112 ? -112 : (e > 127 ? 127 : e);
    return e - 15;
}
__device__ __forceinline__ float max_abs4(float4 v) {
    return fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w)));
}
// Max over the 4 lanes of a quad (DPP quad_perm [1,0,3,2] then [2,3,0,1]);
// m >= 0, so the int ordering of the bits is the float ordering.
__device__ __forceinline__ float quad_max(float m) {
    int v = __float_as_int(m);
    // quad_perm moves always have a source lane, so no old value is needed
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));
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
// (workgroup b runs on XCD b % 8), and each XCD has its own L2.  XCD k holds
// blocks k, k+8, ...: floor(n/8) of them, plus one when k < n % 8.  Mapping
// block b to k * floor(n/8) + min(k, n % 8) + b / 8 hands every XCD one
// contiguous range of logical tiles (a bijection of [0, n) for any n), so
// tiles that share halo rows share an L2.
// Whole-wave lane shifts (DPP wave_shr:1 / wave_shl:1, gfx9): lane i gets
// lane i-1's (shr) or lane i+1's (shl) value of v; the lane with no source
// (0 for shr, 63 for shl) gets `edge`.
__device__ __forceinline__ float4 wave_shr1(float4 v, float4 edge) {
    auto s = [](float e, float x) {
        return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(e), __float_as_int(x), 0x138, 0xF, 0xF, false));
    };
    return make_float4(s(edge.x, v.x), s(edge.y, v.y), s(edge.z, v.z), s(edge.w, v.w));
}
__device__ __forceinline__ float4 wave_shl1(float4 v, float4 edge) {
    auto s = [](float e, float x) {
        return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(e), __float_as_int(x), 0x130, 0xF, 0xF, false));
    };
    return make_float4(s(edge.x, v.x), s(edge.y, v.y), s(edge.z, v.z), s(edge.w, v.w));
}

// f(std::integral_constant<int, 0>{}), ..., f(std::integral_constant<int, N-1>{}):
// a loop whose index is a compile-time constant in every iteration.
template <typename F, int... J>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, J...>) {
    (f(std::integral_constant<int, J>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned n, bool on) {
    if (!on) return b;
    const unsigned k = b & 7u;
    return k * (n >> 3) + min(k, n & 7u) + (b >> 3);
}

}  // namespace stencil::dev
