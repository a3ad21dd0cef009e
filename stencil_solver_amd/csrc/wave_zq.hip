// LDS-staged z-pencil register-queue stencil kernel for gfx950 (the "tiled"
// variant; replaces the reference's shared-memory tile + register pencil,
// src/kernels/stencil_tiled.cu:37-145).
//
// Work decomposition.  A workgroup owns a tile of (x, y) columns, 4*TXV wide
// and TY rows tall, and marches through one z-chunk of planes.  Each thread
// owns 4 consecutive x (one 16-byte float4) of one row, so every global access
// is a dwordx4 and a wave64 covers 64 float4 of consecutive rows.  Tiles start
// on x = 0 (not on the first interior column), so a tile row is whole 128-byte
// lines and no line is written by two workgroups.
//   * z neighbours: a 9-deep register queue of float4 per thread (planes
//     z-4..z+4), refilled with one load per plane; the z loop is unrolled by 9
//     so queue slots are renamed, never moved.
//   * x / y neighbours: plane z is staged in LDS as (TY+8) rows x (TXV+2)
//     float4, written from the queue centre plus the halo float4 (top/bottom 4
//     rows, left/right 4 columns; corners are never read by the star stencil).
//     LDS is double-buffered, so one barrier per plane suffices.
//   * p_prev, vel2 and p_next are touched once per step; their loads and the
//     halo loads for plane z+1 are issued before plane z is computed (gfx950
//     barriers do not drain outstanding global loads).  With NT they use
//     non-temporal accesses so they do not displace p_cur lines in L2.
// HBM traffic per interior point and step: read p_cur, p_prev, vel2 once and
// write p_next once = 16 B (halo re-reads are L2 hits; the tile order keeps
// neighbouring tiles on one XCD).
#include "device_common.hpp"
#include "launchers.hpp"

namespace stencil::hipk {
namespace {

using dev::comp;
using dev::Fd;
using dev::kR;

enum Mode : int { kWaveFused = 0, kWaveExact = 1, kHeatFused = 2, kHeatExact = 3 };

struct ZqArgs {
    const void* cur;
    const void* prev;
    const void* vel2;
    void* next;
    // 16-bit storage: block exponents (kBfp16) and power-of-two scales (kFp16s).
    const int8_t* cur_e;
    const int8_t* prev_e;
    const int8_t* vel2_e;
    int8_t* next_e;
    float p_scale, p_inv, v_inv;
    int nx, ny, nz;
    int tiles_x, tiles_y, chunk;
    int xcd_remap;
    Scalars<float> s;
    float coef;                      // heat: dt2 * alpha
    float w0, wx[4], wy[4], wz[4];   // fused weights (heat: pre-multiplied by coef)
};

using dev::Fmt;

// Loads the 4 values at element index idx (a multiple of 4) of a field in
// storage format FMT and returns them as float (decoding is exact).
template <int FMT, bool NT>
__device__ __forceinline__ float4 load4(const void* base, const int8_t* ebase, size_t idx, float inv) {
    if constexpr (FMT == dev::kF32) {
        const float4* p = reinterpret_cast<const float4*>(static_cast<const float*>(base) + idx);
        if constexpr (NT) return dev::load_stream(p);
        return *p;
    } else {
        const uint2* p = reinterpret_cast<const uint2*>(static_cast<const uint16_t*>(base) + idx);
        const uint2 u = NT ? dev::load_stream_u2(p) : *p;
        if constexpr (FMT == dev::kBf16) return dev::bf16x4_to_float4(u);
        if constexpr (FMT == dev::kFp16s) return dev::fp16x4_to_float4(u, inv);
        return dev::bfp16x4_to_float4(u, ebase[idx >> 4]);
    }
}

// Stores 4 values at element index idx.  Every lane of a wave reaches this
// (kBfp16 shares one exponent across a quad of lanes); `vals` says whether this
// lane writes its values, `blk` whether its quad's block holds outputs.
template <int FMT, bool NT>
__device__ __forceinline__ void store4(void* base, int8_t* ebase, size_t idx, float4 v, float scale,
                                       bool vals, bool blk) {
    if constexpr (FMT == dev::kF32) {
        if (vals) {
            float4* p = reinterpret_cast<float4*>(static_cast<float*>(base) + idx);
            if constexpr (NT) dev::store_stream(p, v);
            else *p = v;
        }
    } else if constexpr (FMT == dev::kBfp16) {
        const float m = vals ? dev::max_abs4(v) : 0.f;
        const int e = dev::bfp16_exponent(dev::quad_max(m));
        uint2* p = reinterpret_cast<uint2*>(static_cast<uint16_t*>(base) + idx);
        const uint2 u = dev::float4_to_bfp16x4(v, e);
        if (vals) {
            if constexpr (NT) dev::store_stream_u2(p, u);
            else *p = u;
        }
        if (blk && (threadIdx.x & 3) == 0) ebase[idx >> 4] = static_cast<int8_t>(e);
    } else {
        if (vals) {
            const uint2 u = FMT == dev::kBf16 ? dev::float4_to_bf16x4(v) : dev::float4_to_fp16x4(v, scale);
            uint2* p = reinterpret_cast<uint2*>(static_cast<uint16_t*>(base) + idx);
            if constexpr (NT) dev::store_stream_u2(p, u);
            else *p = u;
        }
    }
}

template <int TXV, int TY, int MODE, bool NT, int FMT = dev::kF32>
__global__ __launch_bounds__(TXV* TY) void zq_kernel(ZqArgs a) {
    constexpr bool kHeat = (MODE == kHeatFused || MODE == kHeatExact);
    constexpr bool kExact = (MODE == kWaveExact || MODE == kHeatExact);
    constexpr int kThreads = TXV * TY;
    constexpr int kCols = TXV + 2;  // own float4 + left/right halo float4
    constexpr int kRows = TY + 2 * kR;
    constexpr int kHalo = 2 * kR * TXV + 2 * TY;  // halo float4 per plane
    constexpr int kHaloPer = (kHalo + kThreads - 1) / kThreads;
    __shared__ float4 lds[2][kRows][kCols];

    const unsigned logical = dev::xcd_remap(blockIdx.x, gridDim.x, a.xcd_remap != 0);
    const int tile_x = static_cast<int>(logical % static_cast<unsigned>(a.tiles_x));
    const unsigned rest = logical / static_cast<unsigned>(a.tiles_x);
    const int tile_y = static_cast<int>(rest % static_cast<unsigned>(a.tiles_y));
    const int chunk_id = static_cast<int>(rest / static_cast<unsigned>(a.tiles_y));

    const int tid = static_cast<int>(threadIdx.x);
    const int tx = tid % TXV;
    const int ty = tid / TXV;
    const int x0 = tile_x * 4 * TXV;
    const int y0 = kR + tile_y * TY;
    const int x = x0 + 4 * tx;
    const int y = y0 + ty;
    const int z_begin = kR + chunk_id * a.chunk;
    const int z_end = min(z_begin + a.chunk, a.nz - kR);
    if (z_begin >= z_end) return;  // uniform per workgroup

    const size_t plane = static_cast<size_t>(a.nx) * static_cast<size_t>(a.ny);
    const bool own_in = (x < a.nx) && (y < a.ny);
    const bool own_out = (x >= kR) && (x < a.nx - kR) && (y < a.ny - kR);
    const unsigned own_off = static_cast<unsigned>(y) * static_cast<unsigned>(a.nx) +
                             static_cast<unsigned>(x);

    // This thread's halo float4s: item i -> (row, col) of the LDS tile.
    int h_row[kHaloPer], h_col[kHaloPer];
    bool h_in[kHaloPer];
    unsigned h_off[kHaloPer];
#pragma unroll
    for (int k = 0; k < kHaloPer; ++k) {
        const int i = tid + k * kThreads;
        int row = 0, col = 0;
        if (i < kR * TXV) {
            row = i / TXV;
            col = 1 + i % TXV;
        } else if (i < 2 * kR * TXV) {
            row = TY + kR + (i - kR * TXV) / TXV;
            col = 1 + (i - kR * TXV) % TXV;
        } else if (i < 2 * kR * TXV + TY) {
            row = kR + (i - 2 * kR * TXV);
            col = 0;
        } else {
            row = kR + (i - 2 * kR * TXV - TY);
            col = kCols - 1;
        }
        const int hy = y0 - kR + row;
        const int hx = x0 - 4 + 4 * col;
        h_row[k] = row;
        h_col[k] = col;
        h_in[k] = (i < kHalo) && hx >= 0 && (hy < a.ny) && (hx < a.nx);
        h_off[k] = h_in[k] ? static_cast<unsigned>(hy) * static_cast<unsigned>(a.nx) +
                                 static_cast<unsigned>(hx)
                           : 0u;
    }

    // Predicated loads; the zero is built in registers so the compiler cannot
    // turn the predicate into a select between a global and a private address.
    const void* const cur_base = a.cur;
    const int8_t* const cur_e = a.cur_e;
    const float p_inv = a.p_inv;
    auto cur_at = [cur_base, cur_e, p_inv, plane](int z, unsigned off, bool ok) -> float4 {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok) v = load4<FMT, false>(cur_base, cur_e, static_cast<size_t>(z) * plane + off, p_inv);
        return v;
    };
    auto stream_at = [plane, own_off, own_out](const void* f, const int8_t* fe, float inv, int z) -> float4 {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (own_out) v = load4<FMT, NT>(f, fe, static_cast<size_t>(z) * plane + own_off, inv);
        return v;
    };
    // This lane's quad (16 x-consecutive cells) holds output cells: it owns
    // the block exponent of kBfp16 storage.
    const int xq = x & ~15;
    const bool blk_out = (y >= kR) && (y < a.ny - kR) && (xq < a.nx - kR);

    // Register queue: plane (zz - 4 + k) lives in q[(j + k) % 9] at unroll step j.
    float4 q[9];
#pragma unroll
    for (int k = 0; k < 8; ++k) q[k] = cur_at(z_begin - kR + k, own_off, own_in);
    float4 h_cur[kHaloPer];
#pragma unroll
    for (int k = 0; k < kHaloPer; ++k) h_cur[k] = cur_at(z_begin, h_off[k], h_in[k]);
    float4 pv_cur = make_float4(0.f, 0.f, 0.f, 0.f), vv_cur = pv_cur;
    if constexpr (!kHeat) {
        pv_cur = stream_at(a.prev, a.prev_e, a.p_inv, z_begin);
        vv_cur = stream_at(a.vel2, a.vel2_e, a.v_inv, z_begin);
    }

    for (int zb = z_begin; zb < z_end; zb += 9) {
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int zz = zb + j;
            if (zz < z_end) {  // uniform per workgroup; a constant trip count keeps the unroll
                // Refill the queue (plane zz+4) and prefetch plane zz+1's halo
                // and streams; all in bounds since zz+4 <= nz-1.
                q[(j + 8) % 9] = cur_at(zz + kR, own_off, own_in);
                float4 h_nxt[kHaloPer];
#pragma unroll
                for (int k = 0; k < kHaloPer; ++k) h_nxt[k] = cur_at(zz + 1, h_off[k], h_in[k]);
                float4 pv_nxt = make_float4(0.f, 0.f, 0.f, 0.f), vv_nxt = pv_nxt;
                if constexpr (!kHeat) {
                    pv_nxt = stream_at(a.prev, a.prev_e, a.p_inv, zz + 1);
                    vv_nxt = stream_at(a.vel2, a.vel2_e, a.v_inv, zz + 1);
                }

                const int b = (zz - z_begin) & 1;
                const float4 ctr4 = q[(j + 4) % 9];
                lds[b][ty + kR][tx + 1] = ctr4;
#pragma unroll
                for (int k = 0; k < kHaloPer; ++k)
                    if (tid + k * kThreads < kHalo) lds[b][h_row[k]][h_col[k]] = h_cur[k];
                __syncthreads();

                float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
                if (own_out) {
                    const float4 xl = lds[b][ty + kR][tx];
                    const float4 xr = lds[b][ty + kR][tx + 2];
                    const float w[12] = {xl.x,   xl.y,   xl.z,   xl.w,   ctr4.x, ctr4.y,
                                         ctr4.z, ctr4.w, xr.x,   xr.y,   xr.z,   xr.w};
                    float4 ym[kR], yp[kR];
#pragma unroll
                    for (int r = 1; r <= kR; ++r) {
                        ym[r - 1] = lds[b][ty + kR - r][tx + 1];
                        yp[r - 1] = lds[b][ty + kR + r][tx + 1];
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float p = w[4 + i];
                        float res;
                        if constexpr (kExact) {
                            // src/kernels/stencil_naive.cu:75-88, op for op.
                            float lx = Fd<float>::c0 * p * a.s.inv_dx2;
                            float ly = Fd<float>::c0 * p * a.s.inv_dy2;
                            float lz = Fd<float>::c0 * p * a.s.inv_dz2;
#pragma unroll
                            for (int r = 1; r <= kR; ++r) {
                                lx += Fd<float>::c(r) * (w[4 + i + r] + w[4 + i - r]) * a.s.inv_dx2;
                                ly += Fd<float>::c(r) * (comp(yp[r - 1], i) + comp(ym[r - 1], i)) *
                                      a.s.inv_dy2;
                                lz += Fd<float>::c(r) *
                                      (comp(q[(j + 4 + r) % 9], i) + comp(q[(j + 4 - r) % 9], i)) *
                                      a.s.inv_dz2;
                            }
                            if constexpr (kHeat) {
                                res = p + a.coef * (lx + ly + lz);
                            } else {
                                res = 2.f * p - comp(pv_cur, i) +
                                      a.s.dt2 * comp(vv_cur, i) * (lx + ly + lz);
                            }
                        } else {
                            // Combined weights w_r = c_r / h^2, one accumulator.
                            float acc = kHeat ? __builtin_fmaf(a.w0, p, p) : a.w0 * p;
#pragma unroll
                            for (int r = 1; r <= kR; ++r) {
                                acc = __builtin_fmaf(a.wx[r - 1], w[4 + i + r] + w[4 + i - r], acc);
                                acc = __builtin_fmaf(a.wy[r - 1],
                                                     comp(yp[r - 1], i) + comp(ym[r - 1], i), acc);
                                acc = __builtin_fmaf(
                                    a.wz[r - 1],
                                    comp(q[(j + 4 + r) % 9], i) + comp(q[(j + 4 - r) % 9], i), acc);
                            }
                            if constexpr (kHeat) {
                                res = acc;
                            } else {
                                const float base = __builtin_fmaf(2.f, p, -comp(pv_cur, i));
                                res = __builtin_fmaf(a.s.dt2 * comp(vv_cur, i), acc, base);
                            }
                        }
                        dev::set_comp(out, i, res);
                    }
                }
                store4<FMT, NT>(a.next, a.next_e, static_cast<size_t>(zz) * plane + own_off, out,
                                a.p_scale, own_out, blk_out);
#pragma unroll
                for (int k = 0; k < kHaloPer; ++k) h_cur[k] = h_nxt[k];
                pv_cur = pv_nxt;
                vv_cur = vv_nxt;
            }
        }
    }
}

using KernelFn = void (*)(ZqArgs);

// Tile shapes (float4 columns x rows) the library is built with.
struct TileShape {
    int txv, ty;
};
constexpr TileShape kShapes[] = {{16, 16}, {16, 32}, {32, 8}, {32, 16}, {64, 4}, {64, 8}};

template <int MODE, bool NT>
KernelFn pick_shape(int txv, int ty) {
    if (txv == 16 && ty == 16) return zq_kernel<16, 16, MODE, NT>;
    if (txv == 16 && ty == 32) return zq_kernel<16, 32, MODE, NT>;
    if (txv == 32 && ty == 8) return zq_kernel<32, 8, MODE, NT>;
    if (txv == 32 && ty == 16) return zq_kernel<32, 16, MODE, NT>;
    if (txv == 64 && ty == 4) return zq_kernel<64, 4, MODE, NT>;
    if (txv == 64 && ty == 8) return zq_kernel<64, 8, MODE, NT>;
    return nullptr;
}

KernelFn pick(int txv, int ty, int mode, bool nt) {
    switch (mode) {
        case kWaveFused: return nt ? pick_shape<kWaveFused, true>(txv, ty) : pick_shape<kWaveFused, false>(txv, ty);
        case kWaveExact: return nt ? pick_shape<kWaveExact, true>(txv, ty) : pick_shape<kWaveExact, false>(txv, ty);
        case kHeatFused: return nt ? pick_shape<kHeatFused, true>(txv, ty) : pick_shape<kHeatFused, false>(txv, ty);
        default: return nt ? pick_shape<kHeatExact, true>(txv, ty) : pick_shape<kHeatExact, false>(txv, ty);
    }
}

template <int FMT, bool NT>
KernelFn pick_shape16(int txv, int ty) {
    if (txv == 16 && ty == 16) return zq_kernel<16, 16, kWaveExact, NT, FMT>;
    if (txv == 16 && ty == 32) return zq_kernel<16, 32, kWaveExact, NT, FMT>;
    if (txv == 32 && ty == 8) return zq_kernel<32, 8, kWaveExact, NT, FMT>;
    if (txv == 32 && ty == 16) return zq_kernel<32, 16, kWaveExact, NT, FMT>;
    if (txv == 64 && ty == 4) return zq_kernel<64, 4, kWaveExact, NT, FMT>;
    if (txv == 64 && ty == 8) return zq_kernel<64, 8, kWaveExact, NT, FMT>;
    return nullptr;
}

KernelFn pick16(int txv, int ty, int fmt, bool nt) {
    switch (fmt) {
        case dev::kBf16: return nt ? pick_shape16<dev::kBf16, true>(txv, ty) : pick_shape16<dev::kBf16, false>(txv, ty);
        case dev::kFp16s: return nt ? pick_shape16<dev::kFp16s, true>(txv, ty) : pick_shape16<dev::kFp16s, false>(txv, ty);
        case dev::kBfp16: return nt ? pick_shape16<dev::kBfp16, true>(txv, ty) : pick_shape16<dev::kBfp16, false>(txv, ty);
        default: return nullptr;
    }
}

// Fused weights: c_r / h^2 formed in double from the exact rationals and the
// launcher's float 1/h^2, rounded once.
void fill_weights(ZqArgs& a, double scale) {
    const double c[5] = {-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0};
    const double ix = a.s.inv_dx2, iy = a.s.inv_dy2, iz = a.s.inv_dz2;
    a.w0 = static_cast<float>(scale * c[0] * (ix + iy + iz));
    for (int r = 1; r <= 4; ++r) {
        a.wx[r - 1] = static_cast<float>(scale * c[r] * ix);
        a.wy[r - 1] = static_cast<float>(scale * c[r] * iy);
        a.wz[r - 1] = static_cast<float>(scale * c[r] * iz);
    }
}

hipError_t launch_zq(ZqArgs a, const Dims& d, ZqShape shape, int mode, hipStream_t st,
                     int fmt = dev::kF32) {
    if (d.nx <= 2 * kR || d.ny <= 2 * kR || d.nz <= 2 * kR) return hipSuccess;
    shape = zq_resolve(d, shape);
    KernelFn fn = fmt == dev::kF32 ? pick(shape.txv, shape.ty, mode, shape.nt != 0)
                                   : pick16(shape.txv, shape.ty, fmt, shape.nt != 0);
    if (!fn) return hipErrorInvalidValue;
    const int64_t inz = d.nz - 2 * kR;
    a.nx = static_cast<int>(d.nx);
    a.ny = static_cast<int>(d.ny);
    a.nz = static_cast<int>(d.nz);
    a.tiles_x = static_cast<int>((d.nx + 4 * shape.txv - 1) / (4 * shape.txv));
    a.tiles_y = static_cast<int>((d.ny - 2 * kR + shape.ty - 1) / shape.ty);
    a.chunk = static_cast<int>((inz + shape.n_chunks - 1) / shape.n_chunks);
    const int64_t n_chunks = (inz + a.chunk - 1) / a.chunk;
    const int64_t nwg = int64_t(a.tiles_x) * a.tiles_y * n_chunks;
    a.xcd_remap = (shape.xcd_remap != 0 && nwg % 8 == 0) ? 1 : 0;
    hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(nwg)), dim3(shape.txv * shape.ty), 0, st, a);
    return hipGetLastError();
}

}  // namespace

bool zq_supported(const Dims& d, const void* const* ptrs, int n_ptrs) {
    if (d.nx % 4 != 0) return false;
    // 32-bit in-plane offsets.
    if (d.nx * d.ny >= (int64_t(1) << 32)) return false;
    for (int i = 0; i < n_ptrs; ++i)
        if (reinterpret_cast<uintptr_t>(ptrs[i]) % 16 != 0) return false;
    return true;
}

bool zq_shape_valid(int txv, int ty) {
    for (const auto& s : kShapes)
        if (s.txv == txv && s.ty == ty) return true;
    return false;
}

ZqShape zq_resolve(const Dims& d, ZqShape s) {
    if (!zq_shape_valid(s.txv, s.ty)) {
        // Widest tile whose row padding past nx is smallest (ties go wide):
        // 256-cell rows at nx = 512, 128 at nx = 128, ...; 512 threads each.
        int best = 16;
        int64_t best_waste = -1;
        for (int t : {64, 32, 16}) {
            const int64_t w = ((d.nx + 4 * t - 1) / (4 * t)) * (4 * t) - d.nx;
            if (best_waste < 0 || w < best_waste) {
                best = t;
                best_waste = w;
            }
        }
        s.txv = best;
        s.ty = best == 64 ? 8 : best == 32 ? 16 : 32;
    }
    if (s.nt < 0) s.nt = 1;
    const int64_t inz = d.nz - 2 * kR;
    if (s.n_chunks <= 0) {
        // About two workgroups per CU (256 CUs), chunks of at least 32 planes:
        // long z-marches stream best.
        const int64_t tiles =
            ((d.nx + 4 * s.txv - 1) / (4 * s.txv)) * ((d.ny - 2 * kR + s.ty - 1) / s.ty);
        int64_t want = (512 + tiles / 2) / tiles;
        const int64_t max_chunks = inz >= 32 ? inz / 32 : 1;
        if (want > max_chunks) want = max_chunks;
        if (want < 1) want = 1;
        s.n_chunks = static_cast<int>(want);
    }
    if (s.n_chunks > inz) s.n_chunks = static_cast<int>(inz);
    if (s.n_chunks < 1) s.n_chunks = 1;
    return s;
}

hipError_t wave_zq_f32(const float* cur, const float* prev, const float* vel2, float* next,
                       const Dims& d, const Scalars<float>& s, bool exact, ZqShape shape,
                       hipStream_t st) {
    ZqArgs a{};
    a.cur = cur;
    a.prev = prev;
    a.vel2 = vel2;
    a.next = next;
    a.s = s;
    fill_weights(a, 1.0);
    return launch_zq(a, d, shape, exact ? kWaveExact : kWaveFused, st);
}

hipError_t wave_zq_16(int fmt, const Field16& cur, const Field16& prev, const Field16& vel2,
                      const Field16& next, const Dims& d, const Scalars<float>& s, float p_scale,
                      float v_scale, ZqShape shape, hipStream_t st) {
    ZqArgs a{};
    a.cur = cur.data;
    a.prev = prev.data;
    a.vel2 = vel2.data;
    a.next = next.data;
    a.cur_e = cur.exp;
    a.prev_e = prev.exp;
    a.vel2_e = vel2.exp;
    a.next_e = next.exp;
    a.p_scale = p_scale;
    a.p_inv = 1.0f / p_scale;
    a.v_inv = 1.0f / v_scale;
    a.s = s;
    return launch_zq(a, d, shape, kWaveExact, st, fmt);
}

hipError_t heat_zq_f32(const float* t, float* tn, const Dims& d, const Scalars<float>& s,
                       float coef, bool exact, ZqShape shape, hipStream_t st) {
    ZqArgs a{};
    a.cur = t;
    a.next = tn;
    a.s = s;
    a.coef = coef;
    fill_weights(a, static_cast<double>(coef));
    return launch_zq(a, d, shape, exact ? kHeatExact : kHeatFused, st);
}

}  // namespace stencil::hipk
