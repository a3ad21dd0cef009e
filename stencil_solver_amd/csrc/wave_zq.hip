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
//     z-4..z+4), refilled with one load per plane; the 9 steps of a group are
//     expanded at compile time, so queue slots are renamed, never moved.
//   * x / y neighbours: plane z is staged in LDS as (TY+8) rows x (TXV+2)
//     float4, written from the queue centre plus the halo float4 (top/bottom 4
//     rows, left/right 4 columns; corners are never read by the star stencil).
//     LDS is double-buffered, so one barrier per plane suffices.  In FP32
//     tiles one wave wide (TXV = 64: a wave is one tile row) the x neighbours
//     come from the adjacent lanes' registers by DPP wave shifts instead, the
//     row's end lanes loading the float4 past the tile edge themselves; LDS
//     then holds rows only.
//   * p_prev, vel2 and p_next are touched once per step.  Every load is issued
//     ahead (one plane; three for the heat step): the queue refill for plane
//     z+5 and the stream and halo loads for plane z+1 go out before plane z is
//     computed (gfx950 barriers do not drain outstanding global loads).  With
//     NT they use non-temporal accesses so they do not displace p_cur lines
//     in L2; without, all four fields can stay in the 256 MB memory-side
//     cache on grids that fit it (256^3: 4 x 67 MB).  The FWI forward step's
//     L p output is non-temporal either way (written once, read much later).
// HBM traffic per interior point and step: read p_cur, p_prev, vel2 once and
// write p_next once = 16 B (measured 1.05x: halo re-reads are L2 hits, the
// tile order keeping neighbouring tiles on one XCD).  Chained steps alternate
// the z direction (ZqArgs::mirror).
#include <type_traits>
#include <cstdlib>

#include "device_common.hpp"
#include "launchers.hpp"

namespace stencil::hipk {
namespace {

using dev::comp;
using dev::Fd;
using dev::kR;

// kWaveFwi: the exact wave step plus a point source and an optional L p
// output (FWI forward / recompute); kAdjoint: the adjoint step in psi = m lambda,
// psi_out = 2 psi1 - psi2 + dt^2 m L psi1 with m times the residual on the
// receiver plane and the gradient update (FWI backward).
// kPair: one launch runs an adjoint step (the kAdjoint channel: cur, prev,
// next, aux, aux2, grad, res, trace) and a recompute step (the kWaveFwi
// channel: cur_b, prev_b, next_b, lap_out, src_val; no receiver plane) of the
// FWI backward pass on the same tiles, reading m (vel2) once for both.
enum Mode : int {
    kWaveFused = 0,
    kWaveExact = 1,
    kHeatFused = 2,
    kHeatExact = 3,
    kWaveFwi = 4,
    kAdjoint = 5,
    kPair = 6
};

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
    // FWI modes (FP32).  kWaveFwi: cur = p^n, prev = p^{n-1}, vel2 = m,
    // next = p^{n+1}, lap_out = L p^n (nullable), src_val = &wavelet[n]
    // (nullable) added at cell (src_z, src_y, src_x), trace (nullable) = the
    // nx*ny receiver plane that receives p^{n+1} on plane rec_z.  kAdjoint, in
    // psi = m lambda: cur = psi^{k+1}, prev = psi^{k+2}, vel2 = m, aux = L p^k
    // (nullable: no gradient term), grad += dt^2 (L p^k) psi^{k+1} in place,
    // preceded when aux2 = L p^{k+1} is non-null by step k+1's term
    // dt^2 (L p^{k+1}) psi^{k+2} (the two additions as the two steps would
    // make them, one gradient pass for both); next = psi^k = 2 psi^{k+1} -
    // psi^{k+2} + dt^2 m L psi^{k+1}, plus m times the residual plane res
    // (nullable) on plane rec_z, trace (nullable) = psi^k / m at the source cell.
    const void* aux;
    const void* aux2;
    void* lap_out;
    void* grad;
    const float* res;
    const float* src_val;
    float* trace;
    int rec_z, src_z, src_y, src_x;
    // Batched kWaveFwi: workgroup row blockIdx.y steps shot batch[blockIdx.y]
    // (all shots of a launch share the dims): p^{n-1}, p^n, p^{n+1} are its
    // ring[rot], ring[rot+1], ring[rot+2] (mod 3), the source adds
    // wavelet[step].
    const BatchShot* batch;
    int rot, step;
    // Ragged batched kWaveFwi: workgroup b steps work item items[b] of shot
    // rshots[items[b].shot], with that shot's dims, tiling and scalars (same
    // rotation and wavelet index as above).
    const RaggedItem* items;
    const RaggedShot* rshots;
    // Sweep direction: 1 = the workgroups walk z downwards (the kernel's plane
    // z is memory plane nz-1-z; src_z / rec_z arrive already mirrored).  The
    // step is symmetric in z, so both directions give bit-identical results;
    // alternating them between chained steps starts each step on the planes
    // the previous one touched last, which are still in the memory-side cache.
    int mirror;
    // kPair's recompute channel: p^j, p^{j-1}, p^{j+1} (lap_out = L p^j,
    // src_val = &wavelet[j]).
    const void* cur_b;
    const void* prev_b;
    void* next_b;
};

using dev::Fmt;

// Raw storage of the 4 values one thread owns in one plane.
template <int FMT>
using Raw = typename std::conditional<FMT == dev::kF32, float4, uint2>::type;

// Every field access goes through a raw buffer resource (built once per
// workgroup in SGPRs) whose base is the first z-plane the workgroup touches and
// whose range covers its planes; an access names its plane with the
// instruction's scalar offset, (plane - first plane) * plane bytes, and its
// place in the plane with a 32-bit per-lane byte offset.  Plane addressing is
// one scalar multiply per plane and field.  0x00020000 is the gfx9 word 3
// (32-bit data format).
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc plane_rsrc(const void* base, int z, size_t bytes) {
    char* p = const_cast<char*>(static_cast<const char*>(base)) + static_cast<size_t>(z) * bytes;
    return __builtin_amdgcn_make_buffer_rsrc(p, 0, static_cast<int>(static_cast<unsigned>(bytes)), 0x00020000);
}
// planes first .. first + n_planes - 1 of a field with plane_bytes per plane
// (n_planes * plane_bytes < 2^32: the launcher bounds the z-chunk length).
__device__ __forceinline__ Rsrc span_rsrc(const void* base, int first, size_t plane_bytes, unsigned n_planes) {
    char* p = const_cast<char*>(static_cast<const char*>(base)) + static_cast<size_t>(first) * plane_bytes;
    return __builtin_amdgcn_make_buffer_rsrc(p, 0, static_cast<int>(static_cast<unsigned>(plane_bytes) * n_planes),
                                             0x00020000);
}
// Cache policy bits of gfx950 buffer accesses: 2 = non-temporal.
constexpr int kNtAux = 2;

// Loads the 4 stored values at byte offset boff of the plane at scalar
// offset soff.
template <int FMT, bool NT>
__device__ __forceinline__ Raw<FMT> load_raw(Rsrc r, unsigned boff, int soff) {
    constexpr int aux = NT ? kNtAux : 0;
    if constexpr (FMT == dev::kF32) {
        typedef unsigned v4u __attribute__((ext_vector_type(4)));
        const v4u v = __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(boff), soff, aux);
        return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                           __uint_as_float(v.w));
    } else {
        const dev::v2u v = __builtin_amdgcn_raw_buffer_load_b64(r, static_cast<int>(boff), soff, aux);
        return make_uint2(v.x, v.y);
    }
}

// Exact decode of 4 stored values (e: the block exponent, kBfp16 only).
// With SCALED false a kFp16s field is returned as stored (value * scale): the
// fused forms run in the scaled domain, which moves every intermediate by the
// same power of two.
template <int FMT, bool SCALED>
__device__ __forceinline__ float4 decode(Raw<FMT> u, int e, float inv) {
    if constexpr (FMT == dev::kF32) return u;
    else if constexpr (FMT == dev::kBf16) return dev::bf16x4_to_float4(u);
    else if constexpr (FMT == dev::kFp16s) return SCALED ? dev::fp16x4_to_float4(u, inv) : dev::fp16x4_raw(u);
    else return dev::bfp16x4_to_float4(u, e);
}

// Stores 4 values at byte offset boff of the plane at scalar offset soff.
// Every lane of a wave reaches this (kBfp16 shares one block scale across a
// quad of lanes); `vals` says whether this lane writes its values, `blk`
// whether its quad's block holds outputs (then lane 0 of the quad writes the
// scale at byte eoff of the scale plane at esoff of `er`).
template <int FMT, bool NT, bool SCALED>
__device__ __forceinline__ void store4(Rsrc r, int soff, Rsrc er, int esoff, unsigned boff, unsigned eoff,
                                       float4 v, float scale, bool vals, bool blk) {
    constexpr int aux = NT ? kNtAux : 0;
    if constexpr (FMT == dev::kF32) {
        if (vals) {
            typedef unsigned v4u __attribute__((ext_vector_type(4)));
            const v4u w = {__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z),
                           __float_as_uint(v.w)};
            __builtin_amdgcn_raw_buffer_store_b128(w, r, static_cast<int>(boff), soff, aux);
        }
    } else {
        uint2 u;
        int e = 0;
        if constexpr (FMT == dev::kBfp16) {
            // lanes that store nothing do not raise the block maximum
            e = dev::bfp16_exponent(dev::quad_max(vals ? dev::max_abs4(v) : 0.f));
            u = dev::float4_to_bfp16x4(v, e);
        } else if constexpr (FMT == dev::kBf16) {
            u = dev::float4_to_bf16x4(v);
        } else {
            u = SCALED ? dev::float4_to_fp16x4(v, scale) : dev::fp16x4_pack(v);
        }
        if (vals) {
            const dev::v2u w = {u.x, u.y};
            __builtin_amdgcn_raw_buffer_store_b64(w, r, static_cast<int>(boff), soff, aux);
        }
        if constexpr (FMT == dev::kBfp16)
            if (blk && (threadIdx.x & 3) == 0)
                __builtin_amdgcn_raw_buffer_store_b8(static_cast<unsigned char>(e), er, static_cast<int>(eoff), esoff,
                                                     0);
    }
}

// FP32 prefetch depth (planes in flight per stream).  The heat step has one
// stream and registers to spare: three planes in flight (+11 % at 512^3).
// The wave step keeps one: its three streams already have enough bytes in
// flight, and a second plane costs the 128-VGPR budget (measured slower);
// so do 1024-thread workgroups (capped at 128 VGPRs) and the FWI modes.
constexpr int prefetch_depth(int txv, int ty, int mode) {
    if (txv * ty == 1024) return 1;
    if (mode == kHeatFused || mode == kHeatExact) return 3;
    return 1;
}

// Waves per SIMD the register budget is sized for: 4 (128 VGPRs: two
// 512-thread or four 256-thread workgroups per CU, one 1024-thread one)
// where the kernel fits without scratch; 3 (168 VGPRs) for a 768-thread
// workgroup, and where 128 would spill: three halo items per thread, two in
// a 256-thread workgroup (64x4) or with block-FP16 storage (exponent registers), the heat step with three
// planes in flight and two halo items, and the adjoint.
constexpr int min_waves(int txv, int ty, int fmt, int mode) {
    const int threads = txv * ty;
    // the step pair holds two register queues and two stencils' streams:
    // 256 VGPRs, one 512-thread or two 256-thread workgroups per CU
    if (mode == kPair) return 2;
    const bool shuffle_x = txv == 64 && fmt == dev::kF32;  // x neighbours by lane shifts: no column items
    const int halo_per = (2 * dev::kR * txv + (shuffle_x ? 0 : 2 * ty) + threads - 1) / threads;
    if (threads == 768 || halo_per >= 3 || (halo_per == 2 && (fmt == dev::kBfp16 || threads == 256))) return 3;
    if (mode == kAdjoint && threads != 1024) return 3;  // two more streams (L p, gradient)
    if (fmt == dev::kF32 && prefetch_depth(txv, ty, mode) > 2 && halo_per >= 2) return 3;
    return 4;
}

// A value the whole wave shares, read from the first lane, so the compiler
// keeps it (and the buffer resources built from it) in scalar registers: the
// work-item fields of a ragged launch are loaded by vector loads, and in the
// persistent kernel they hang off an index that came through LDS.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float wave_uniform(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
template <class T>
__device__ __forceinline__ T* wave_uniform(T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v >> 32)));
    return reinterpret_cast<T*>((static_cast<uint64_t>(hi) << 32) | lo);
}

// One workgroup's tile (work item `block`: the workgroup index of a plain
// launch, or an index into a.items) of one step.  Returns early only with a
// value uniform across the workgroup.
template <int TXV, int TY, int MODE, bool NT, int FMT = dev::kF32>
__device__ __forceinline__ void zq_tile(ZqArgs a, unsigned block) {
    constexpr bool kHeat = (MODE == kHeatFused || MODE == kHeatExact);
    constexpr bool kFwd = (MODE == kWaveFwi);
    constexpr bool kAdj = (MODE == kAdjoint);
    constexpr bool kPairM = (MODE == kPair);
    constexpr bool kExact = (MODE == kWaveExact || MODE == kHeatExact || kFwd || kAdj || kPairM);
    static_assert(!(kFwd || kAdj || kPairM) || FMT == dev::kF32, "FWI modes are FP32");
    // channels: the step pair stages two fields per plane (LDS set 1 is the
    // recompute's p, set 0 the adjoint's psi); every other mode one
    constexpr int kNch = kPairM ? 2 : 1;
    int item_tx = -1, item_ty = 0, item_ch = 0;  // ragged work item (tile, chunk)
    if constexpr (kFwd) {
        if (a.items) {  // uniform per workgroup
            const RaggedItem& it = a.items[block];
            const int shot = wave_uniform(it.shot);
            if (shot < 0) return;
            const RaggedShot* sh = a.rshots + shot;
            const int r = a.rot % 3;
            a.prev = wave_uniform(sh->ring[r]);
            a.cur = wave_uniform(sh->ring[r == 2 ? 0 : r + 1]);
            a.next = wave_uniform(sh->ring[r == 0 ? 2 : r - 1]);
            a.vel2 = wave_uniform(sh->vel2);
            a.src_z = wave_uniform(sh->src_z);
            a.src_y = wave_uniform(sh->src_y);
            a.src_x = wave_uniform(sh->src_x);
            const float* wavelet = wave_uniform(sh->wavelet);
            a.src_val = wavelet ? wavelet + a.step : nullptr;
            a.lap_out = nullptr;
            a.nx = wave_uniform(sh->nx);
            a.ny = wave_uniform(sh->ny);
            a.nz = wave_uniform(sh->nz);
            a.chunk = wave_uniform(sh->chunk);
            a.s.inv_dx2 = wave_uniform(sh->s.inv_dx2);
            a.s.inv_dy2 = wave_uniform(sh->s.inv_dy2);
            a.s.inv_dz2 = wave_uniform(sh->s.inv_dz2);
            a.s.dt2 = wave_uniform(sh->s.dt2);
            if (a.mirror) a.src_z = a.nz - 1 - a.src_z;
            item_tx = wave_uniform(it.tile_x);
            item_ty = wave_uniform(it.tile_y);
            item_ch = wave_uniform(it.chunk_id);
        } else if (a.batch) {  // uniform per workgroup
            // Fields read straight from the table (a private copy of the entry
            // indexed by rot would live in scratch memory).
            const BatchShot* sh = a.batch + blockIdx.y;
            const int r = a.rot % 3;
            a.prev = sh->ring[r];
            a.cur = sh->ring[r == 2 ? 0 : r + 1];
            a.next = sh->ring[r == 0 ? 2 : r - 1];
            a.vel2 = sh->vel2;
            a.src_z = sh->src_z;
            a.src_y = sh->src_y;
            a.src_x = sh->src_x;
            a.src_val = sh->wavelet ? sh->wavelet + a.step : nullptr;
            a.lap_out = nullptr;
            if (a.mirror) a.src_z = a.nz - 1 - a.src_z;
        }
    }
    constexpr bool kBlk = (FMT == dev::kBfp16);
    // FP32 tiles one wave wide take their x neighbours by lane shifts instead
    // of LDS reads (16-bit storage is VALU-bound: it keeps the LDS reads).
    constexpr bool kShuffleX = (TXV == 64 && FMT == dev::kF32);
    constexpr unsigned kEsz = FMT == dev::kF32 ? 4u : 2u;  // bytes per stored value
    constexpr int kThreads = TXV * TY;
    constexpr int kCols = TXV + 2;  // own float4 + left/right halo float4
    constexpr int kRows = TY + 2 * kR;
    // halo float4 per plane staged through LDS: top/bottom rows, plus the
    // left/right columns unless the x neighbours come by lane shifts
    constexpr int kHalo = 2 * kR * TXV + (kShuffleX ? 0 : 2 * TY);
    constexpr int kHaloPer = (kHalo + kThreads - 1) / kThreads;
    __shared__ float4 lds[kNch][2][kRows][kCols];

    int tile_x = item_tx, tile_y = item_ty, chunk_id = item_ch;
    if (item_tx < 0) {
        const unsigned logical = dev::xcd_remap(block, gridDim.x, a.xcd_remap != 0);
        tile_x = static_cast<int>(logical % static_cast<unsigned>(a.tiles_x));
        const unsigned rest = logical / static_cast<unsigned>(a.tiles_x);
        tile_y = static_cast<int>(rest % static_cast<unsigned>(a.tiles_y));
        chunk_id = static_cast<int>(rest / static_cast<unsigned>(a.tiles_y));
    }

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
    const size_t plane_bytes = plane * kEsz;
    const size_t eplane = plane >> 4;  // block exponents per plane (nx % 16 == 0)
    const bool own_in = (x < a.nx) && (y < a.ny);
    const bool own_out = (x >= kR) && (x < a.nx - kR) && (y < a.ny - kR);
    const unsigned own_off = static_cast<unsigned>(y) * static_cast<unsigned>(a.nx) +
                             static_cast<unsigned>(x);
    const unsigned own_b = own_off * kEsz;

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

    const float p_inv = a.p_inv, v_inv = a.v_inv;

    // Loads are unconditional.  A lane whose cell lies outside the grid (x or y
    // past the last row / column, the x < 0 halo of the first tile column, the
    // unused halo items) reads offset 0 instead: the LDS cell or register it
    // fills is never read by a lane that stores an output (an output lane at
    // x reads x-4..x+7 and y-4..y+4, all inside the grid).
    const unsigned own_ld = own_in ? own_b : 0u;
    // kShuffleX: the row's end lanes load the float4 just past the tile on
    // their side (left of lane 0, right of the last lane); an edge outside the
    // grid is never read by an output lane and loads offset 0.
    unsigned xe_ld = own_ld;
    if constexpr (kShuffleX) {
        const int ex = tx == 0 ? x0 - 4 : x0 + 4 * TXV;
        const bool e_in = (tx == 0 || tx == TXV - 1) && ex >= 0 && ex < a.nx && y < a.ny;
        xe_ld = e_in ? (static_cast<unsigned>(y) * static_cast<unsigned>(a.nx) + static_cast<unsigned>(ex)) * kEsz
                     : own_ld;
    }
    const unsigned own_eld = own_in ? (own_off >> 4) : 0u;
    unsigned h_ld[kHaloPer], h_eld[kHaloPer];
#pragma unroll
    for (int k = 0; k < kHaloPer; ++k) {
        h_ld[k] = h_in[k] ? h_off[k] * kEsz : 0u;
        h_eld[k] = h_in[k] ? (h_off[k] >> 4) : 0u;
    }
    const int z_last = a.nz - 1;
    // Kernel planes this workgroup touches: z_begin-4 (queue fill) up to the
    // last prefetch of its last group of 9 steps (z_end+14 at most), clamped
    // to the grid; the kernel's plane z is memory plane z_org + z_sgn * z.
    const int z_sgn = a.mirror ? -1 : 1;
    const int z_org = a.mirror ? z_last : 0;
    const int zl_lo = z_begin - kR;
    const int zl_hi = min(z_end + 16, z_last);
    const int p_lo = a.mirror ? z_last - zl_hi : zl_lo;  // first memory plane
    const unsigned n_span = static_cast<unsigned>(zl_hi - zl_lo + 1);
    // scalar offset of kernel plane z in a field with pb bytes per plane
    auto soff = [z_sgn, z_org, p_lo](int z, size_t pb) -> int {
        return static_cast<int>(static_cast<unsigned>(z_org + z_sgn * z - p_lo) * static_cast<unsigned>(pb));
    };
    auto field = [p_lo, n_span](const void* base, size_t pb) { return span_rsrc(base, p_lo, pb, n_span); };
    auto raw_at = [plane_bytes, soff](Rsrc f, int z, unsigned boff) -> Raw<FMT> {
        return load_raw<FMT, false>(f, boff, soff(z, plane_bytes));
    };
    auto raw_stream = [plane_bytes, own_ld, soff](Rsrc f, int z) -> Raw<FMT> {
        return load_raw<FMT, NT>(f, own_ld, soff(z, plane_bytes));
    };
    // Fields read once per pass (the adjoint's L p inputs): non-temporal in
    // either mode, like the forward step's L p output.
    auto raw_once = [plane_bytes, own_ld, soff](Rsrc f, int z) -> Raw<FMT> {
        return load_raw<FMT, true>(f, own_ld, soff(z, plane_bytes));
    };
    auto exp_at = [eplane, soff](Rsrc e, int z, unsigned eoff) -> int {
        (void)eplane;
        (void)soff;
        if constexpr (kBlk)
            return static_cast<signed char>(
                __builtin_amdgcn_raw_buffer_load_b8(e, static_cast<int>(eoff), soff(z, eplane), 0));
        return 0;
    };
    // The fields this mode touches (unused ones are dropped by the compiler).
    const Rsrc cur_r = field(a.cur, plane_bytes);
    const Rsrc prev_r = field(a.prev, plane_bytes);
    const Rsrc vel_r = field(a.vel2, plane_bytes);
    const Rsrc next_r = field(a.next, plane_bytes);
    const Rsrc aux_r = field(a.aux, plane_bytes);
    const Rsrc aux2_r = field(a.aux2, plane_bytes);
    const Rsrc grad_r = field(a.grad, plane_bytes);
    const Rsrc lap_r = field(a.lap_out, plane_bytes);
    const Rsrc cur_b_r = field(a.cur_b, plane_bytes);
    const Rsrc prev_b_r = field(a.prev_b, plane_bytes);
    const Rsrc next_b_r = field(a.next_b, plane_bytes);
    const Rsrc cur_er = field(a.cur_e, eplane);
    const Rsrc prev_er = field(a.prev_e, eplane);
    const Rsrc vel_er = field(a.vel2_e, eplane);
    const Rsrc next_er = field(a.next_e, eplane);
    // This lane's quad (16 x-consecutive cells) holds output cells: it owns
    // the block exponent of kBfp16 storage.
    const int xq = x & ~15;
    const bool blk_out = (y >= kR) && (y < a.ny - kR) && (xq < a.nx - kR);

    // One plane of the star stencil for this thread's 4 points: LDS buffer b
    // holds plane zz (x and y neighbours), zq[k] is plane zz - 4 + k.
    // dt2v multiplies vv in the fused forms (dt^2, or dt^2 * 2^k when vv holds
    // the raw mantissas of a block-FP16 vel2 block).
    // lc: L p^k at the point (kAdjoint's gradient term); lap4 receives L p
    // (kWaveFwi).
    // The helpers below take the mode of the channel they serve as a
    // compile-time argument (cm: std::integral_constant): MODE itself, or for
    // the step pair kAdjoint (LDS set 0) and kWaveFwi (LDS set 1).
    // The x neighbour float4s of this lane's centre (plane in LDS buffer b).
    // Called by every lane, outside any divergent branch: the lane shifts
    // read the adjacent lanes' registers, which must be active.
    auto x_neighbours = [&](auto cm, int b, float4 ctr4, float4 edge, float4& xl, float4& xr) {
        constexpr int ch = (kPairM && decltype(cm)::value == kWaveFwi) ? 1 : 0;
        if constexpr (kShuffleX) {
            // from the adjacent lanes' centre float4 (a wave is one tile row);
            // the row's end lanes take their edge float4 instead.
            (void)b;
            xl = dev::wave_shr1(ctr4, edge);
            xr = dev::wave_shl1(ctr4, edge);
        } else {
            (void)edge;
            xl = lds[ch][b][ty + kR][tx];
            xr = lds[ch][b][ty + kR][tx + 2];
        }
    };
    auto compute = [&](auto cm, int b, const float4 (&zq)[9], float4 xl, float4 xr, float4 pv, float4 vv,
                       float dt2v, float4 lc, float4& lap4) -> float4 {
        constexpr int ch = (kPairM && decltype(cm)::value == kWaveFwi) ? 1 : 0;
        constexpr bool cFwd = decltype(cm)::value == kWaveFwi;
        float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 ctr4 = zq[kR];
        const float w[12] = {xl.x,   xl.y,   xl.z,   xl.w,   ctr4.x, ctr4.y,
                             ctr4.z, ctr4.w, xr.x,   xr.y,   xr.z,   xr.w};
        float4 ym[kR], yp[kR];
#pragma unroll
        for (int r = 1; r <= kR; ++r) {
            ym[r - 1] = lds[ch][b][ty + kR - r][tx + 1];
            yp[r - 1] = lds[ch][b][ty + kR + r][tx + 1];
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
                    ly += Fd<float>::c(r) * (comp(yp[r - 1], i) + comp(ym[r - 1], i)) * a.s.inv_dy2;
                    lz += Fd<float>::c(r) * (comp(zq[kR + r], i) + comp(zq[kR - r], i)) * a.s.inv_dz2;
                }
                if constexpr (kHeat) {
                    res = p + a.coef * (lx + ly + lz);
                } else {  // wave, FWI forward, adjoint in psi (vv = m)
                    res = 2.f * p - comp(pv, i) + a.s.dt2 * comp(vv, i) * (lx + ly + lz);
                    if constexpr (cFwd) dev::set_comp(lap4, i, lx + ly + lz);
                }
            } else {
                // Combined weights w_r = c_r / h^2, one accumulator.
                float acc = kHeat ? __builtin_fmaf(a.w0, p, p) : a.w0 * p;
#pragma unroll
                for (int r = 1; r <= kR; ++r) {
                    acc = __builtin_fmaf(a.wx[r - 1], w[4 + i + r] + w[4 + i - r], acc);
                    acc = __builtin_fmaf(a.wy[r - 1], comp(yp[r - 1], i) + comp(ym[r - 1], i), acc);
                    acc = __builtin_fmaf(a.wz[r - 1], comp(zq[kR + r], i) + comp(zq[kR - r], i), acc);
                }
                if constexpr (kHeat) {
                    res = acc;
                } else {
                    const float base = __builtin_fmaf(2.f, p, -comp(pv, i));
                    res = __builtin_fmaf(dt2v * comp(vv, i), acc, base);
                }
            }
            dev::set_comp(out, i, res);
        }
        return out;
    };
    // Steps past z_end (the tail of the last group of 9 in a chunk whose length
    // is not a multiple of 9) run on clamped planes and store nothing.
    auto finish = [&](auto cm, int zz, int b, const float4 (&zq)[9], float4 edge, float4 pv, float4 vv,
                      float dt2v, float4 lc, float4 gr, float4 l2) {
        constexpr bool cFwd = decltype(cm)::value == kWaveFwi;
        constexpr bool cAdj = decltype(cm)::value == kAdjoint;
        // Every lane computes (its inputs are loaded, from offset 0 outside
        // the grid); only output lanes store, and block-FP16 leaves the other
        // lanes out of its block maximum.
        float4 lap4 = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 xl, xr;
        x_neighbours(cm, b, zq[kR], edge, xl, xr);
        float4 out = compute(cm, b, zq, xl, xr, pv, vv, dt2v, lc, lap4);
        if (zz < z_end) {
            if constexpr (cFwd || cAdj) {
                // the source / trace cell lies in this lane's 4 cells of this plane
                const bool at_src = own_out && zz == a.src_z && y == a.src_y && a.src_x >= x && a.src_x < x + 4;
                if constexpr (cFwd) {
                    if (at_src && a.src_val) dev::set_comp(out, a.src_x - x, comp(out, a.src_x - x) + *a.src_val);
                    // receiver plane: p^{n+1} on plane rec_z into the trace (its
                    // halo is pre-zeroed); the recompute of a step pair records none
                    if constexpr (!kPairM)
                        if (a.trace && zz == a.rec_z)
                            store4<FMT, false, true>(plane_rsrc(a.trace, 0, plane_bytes), 0,
                                                     plane_rsrc(a.trace, 0, 0), 0, own_b, 0, out, 1.f, own_out,
                                                     false);
                    if (a.lap_out)  // written once, read much later: non-temporal in either mode
                        store4<FMT, true, true>(lap_r, soff(zz, plane_bytes), lap_r, 0, own_b, 0, lap4, 1.f,
                                                own_out, false);
                } else {
                    if (a.res && zz == a.rec_z && own_out) {  // + m * residual
                        const float4 r = *reinterpret_cast<const float4*>(a.res + own_off);
                        out = make_float4(out.x + vv.x * r.x, out.y + vv.y * r.y, out.z + vv.z * r.z,
                                          out.w + vv.w * r.w);
                    }
                    if (a.aux) {  // gradient: grad += dt^2 * (L p^k) * psi^{k+1}
                        const float4 c = zq[kR];
                        float4 g = gr;
                        if (a.aux2)  // first step k+1's term, dt^2 * (L p^{k+1}) * psi^{k+2}
                            g = make_float4(g.x + a.s.dt2 * l2.x * pv.x, g.y + a.s.dt2 * l2.y * pv.y,
                                            g.z + a.s.dt2 * l2.z * pv.z, g.w + a.s.dt2 * l2.w * pv.w);
                        g = make_float4(g.x + a.s.dt2 * lc.x * c.x, g.y + a.s.dt2 * lc.y * c.y,
                                        g.z + a.s.dt2 * lc.z * c.z, g.w + a.s.dt2 * lc.w * c.w);
                        store4<FMT, NT, true>(grad_r, soff(zz, plane_bytes), grad_r, 0, own_b, 0, g, 1.f,
                                              own_out, false);
                    }
                    if (at_src && a.trace) a.trace[0] = comp(out, a.src_x - x) / comp(vv, a.src_x - x);
                }
            }
            if constexpr (kPairM && cFwd)
                store4<FMT, NT, true>(next_b_r, soff(zz, plane_bytes), next_b_r, 0, own_b, 0, out, 1.f, own_out,
                                      false);
            else
                store4<FMT, NT, kExact>(next_r, soff(zz, plane_bytes), next_er, kBlk ? soff(zz, eplane) : 0, own_b,
                                        own_off >> 4, out, a.p_scale, own_out, blk_out);
        }
    };
    using ModeC = std::integral_constant<int, MODE>;

    // Register queue: plane (zz - 4 + k) lives in q[(j + k) % 9] at unroll step j.
    float4 q[9];
    if constexpr (FMT == dev::kF32) {
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] = raw_at(cur_r, z_begin - kR + k, own_ld);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            q[k] = decode<FMT, kExact>(raw_at(cur_r, z_begin - kR + k, own_ld),
                                       exp_at(cur_er, z_begin - kR + k, own_eld), p_inv);
    }

    if constexpr (kPairM) {
        // The step pair: the adjoint channel (queue q, LDS set 0) and the
        // recompute channel (queue qb, LDS set 1) march the same planes, one
        // barrier per plane for both; m is loaded once and serves both steps.
        // Every load goes out one plane ahead (queue refills: four planes
        // further), as in the single FWI steps.
        const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
        const bool lap_in = a.aux != nullptr;
        const bool lap2_in = lap_in && a.aux2 != nullptr;
        float4 qb[9];
#pragma unroll
        for (int k = 0; k < 8; ++k) qb[k] = raw_at(cur_b_r, z_begin - kR + k, own_ld);
        struct Stage {
            float4 r, rb, xe, xeb, h[kHaloPer], hb[kHaloPer], pv, pvb, vv, lc, gr, l2;
        };
        // loads of plane z (queue refills: plane z + kR)
        auto fetch = [&](int z, Stage& st) {
            st.r = raw_at(cur_r, min(z + kR, z_last), own_ld);
            st.rb = raw_at(cur_b_r, min(z + kR, z_last), own_ld);
            st.xe = st.xeb = zero4;
            if constexpr (kShuffleX) {
                st.xe = raw_at(cur_r, z, xe_ld);
                st.xeb = raw_at(cur_b_r, z, xe_ld);
            }
#pragma unroll
            for (int k = 0; k < kHaloPer; ++k) {
                st.h[k] = raw_at(cur_r, z, h_ld[k]);
                st.hb[k] = raw_at(cur_b_r, z, h_ld[k]);
            }
            st.pv = raw_stream(prev_r, z);
            st.pvb = raw_stream(prev_b_r, z);
            st.vv = raw_stream(vel_r, z);
            st.lc = st.gr = st.l2 = zero4;
            if (lap_in) {
                st.lc = raw_once(aux_r, z);
                st.gr = raw_stream(grad_r, z);
            }
            if (lap2_in) st.l2 = raw_once(aux2_r, z);
        };
        Stage cs;
        fetch(min(z_begin, z_last), cs);
        for (int zb = z_begin; zb < z_end; zb += 9) {
            dev::static_for<9>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const int zz = zb + j;
                q[(j + 8) % 9] = cs.r;
                qb[(j + 8) % 9] = cs.rb;
                Stage ns;
                fetch(min(zz + 1, z_last), ns);
                const int b = (zz - z_begin) & 1;
                lds[0][b][ty + kR][tx + 1] = q[(j + 4) % 9];
                lds[kNch - 1][b][ty + kR][tx + 1] = qb[(j + 4) % 9];
#pragma unroll
                for (int k = 0; k < kHaloPer; ++k)
                    if (tid + k * kThreads < kHalo) {
                        lds[0][b][h_row[k]][h_col[k]] = cs.h[k];
                        lds[kNch - 1][b][h_row[k]][h_col[k]] = cs.hb[k];
                    }
                __syncthreads();
                const float4 zq[9] = {q[j % 9],       q[(j + 1) % 9], q[(j + 2) % 9],
                                      q[(j + 3) % 9], q[(j + 4) % 9], q[(j + 5) % 9],
                                      q[(j + 6) % 9], q[(j + 7) % 9], q[(j + 8) % 9]};
                const float4 zqb[9] = {qb[j % 9],       qb[(j + 1) % 9], qb[(j + 2) % 9],
                                       qb[(j + 3) % 9], qb[(j + 4) % 9], qb[(j + 5) % 9],
                                       qb[(j + 6) % 9], qb[(j + 7) % 9], qb[(j + 8) % 9]};
                finish(std::integral_constant<int, kAdjoint>{}, zz, b, zq, cs.xe, cs.pv, cs.vv, a.s.dt2, cs.lc,
                       cs.gr, cs.l2);
                finish(std::integral_constant<int, kWaveFwi>{}, zz, b, zqb, cs.xeb, cs.pvb, cs.vv, a.s.dt2, zero4,
                       zero4, zero4);
                cs = ns;
            });
        }
    } else if constexpr (FMT == dev::kF32) {
        // FP32: every load is issued kPD steps ahead: the queue refill for
        // plane zz+4+kPD, halo and streams for plane zz+kPD; stage d of each
        // FIFO holds plane zz+d (queue refills: zz+4+d).  (kAdjoint streams
        // vel2 = m, and L p^k and the gradient when aux is non-null.)
        constexpr int kPD = prefetch_depth(TXV, TY, MODE);
        const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
        const bool lap_in = kAdj && a.aux != nullptr;
        const bool lap2_in = lap_in && a.aux2 != nullptr;
        float4 rq[kPD], xeq[kPD];
        float4 hq[kPD][kHaloPer];
        float4 pvq[kPD], vvq[kPD], lcq[kPD], grq[kPD], l2q[kPD];
        // Loads of plane z into FIFO stage d (queue refill: plane z + kR).
        auto fetch = [&](int z, float4& r, float4& xe, float4 (&h)[kHaloPer], float4& pv, float4& vv, float4& lc,
                         float4& gr, float4& l2) {
            r = raw_at(cur_r, min(z + kR, z_last), own_ld);
            xe = zero4;
            if constexpr (kShuffleX) xe = raw_at(cur_r, z, xe_ld);
#pragma unroll
            for (int k = 0; k < kHaloPer; ++k) h[k] = raw_at(cur_r, z, h_ld[k]);
            pv = vv = lc = gr = l2 = zero4;
            if constexpr (!kHeat) {
                pv = raw_stream(prev_r, z);
                vv = raw_stream(vel_r, z);
            }
            if constexpr (kAdj) {
                if (lap_in) {
                    lc = raw_once(aux_r, z);
                    gr = raw_stream(grad_r, z);
                }
                if (lap2_in) l2 = raw_once(aux2_r, z);
            }
        };
#pragma unroll
        for (int d = 0; d < kPD; ++d)
            fetch(min(z_begin + d, z_last), rq[d], xeq[d], hq[d], pvq[d], vvq[d], lcq[d], grq[d], l2q[d]);
        for (int zb = z_begin; zb < z_end; zb += 9) {
            // The 9 steps are expanded at compile time (j is a constant in
            // every queue index), so the queue arrays stay in registers.
            dev::static_for<9>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const int zz = zb + j;
                q[(j + 8) % 9] = rq[0];
                float4 r_n, xe_n, h_n[kHaloPer], pv_n, vv_n, lc_n, gr_n, l2_n;
                fetch(min(zz + kPD, z_last), r_n, xe_n, h_n, pv_n, vv_n, lc_n, gr_n, l2_n);
                const int b = (zz - z_begin) & 1;
                lds[0][b][ty + kR][tx + 1] = q[(j + 4) % 9];
#pragma unroll
                for (int k = 0; k < kHaloPer; ++k)
                    if (tid + k * kThreads < kHalo) lds[0][b][h_row[k]][h_col[k]] = hq[0][k];
                __syncthreads();
                const float4 zq[9] = {q[j % 9],       q[(j + 1) % 9], q[(j + 2) % 9],
                                      q[(j + 3) % 9], q[(j + 4) % 9], q[(j + 5) % 9],
                                      q[(j + 6) % 9], q[(j + 7) % 9], q[(j + 8) % 9]};
                finish(ModeC{}, zz, b, zq, xeq[0], pvq[0], vvq[0], a.s.dt2, lcq[0], grq[0], l2q[0]);
                // FIFO shift (register renaming under the unroll)
#pragma unroll
                for (int d = 0; d + 1 < kPD; ++d) {
                    rq[d] = rq[d + 1];
                    xeq[d] = xeq[d + 1];
#pragma unroll
                    for (int k = 0; k < kHaloPer; ++k) hq[d][k] = hq[d + 1][k];
                    pvq[d] = pvq[d + 1];
                    vvq[d] = vvq[d + 1];
                    lcq[d] = lcq[d + 1];
                    grq[d] = grq[d + 1];
                    l2q[d] = l2q[d + 1];
                }
                rq[kPD - 1] = r_n;
                xeq[kPD - 1] = xe_n;
#pragma unroll
                for (int k = 0; k < kHaloPer; ++k) hq[kPD - 1][k] = h_n[k];
                pvq[kPD - 1] = pv_n;
                vvq[kPD - 1] = vv_n;
                lcq[kPD - 1] = lc_n;
                grq[kPD - 1] = gr_n;
                l2q[kPD - 1] = l2_n;
            });
        }
    } else {
        // 16-bit storage: every load is issued one step ahead as raw 16-bit
        // words (the queue refill for plane zz+5, halo and streams for plane
        // zz+1) and decoded in the step that uses it, just before it is
        // reloaded.  Before the plane's barrier a step does only what LDS
        // needs (decode and store the halo, reload it); the queue's newest
        // plane and the p_prev / vel2 streams are decoded and reloaded after
        // the barrier, so the barrier phase is short and the waves holding
        // halo items do not hold the others up (+6 % block-FP16, +7 % bf16 at
        // 512^3 against decoding everything before the barrier).
        // Halo items exist for the first kHalo threads (whole waves but one),
        // so the others skip the decodes.  The loads stay unconditional: under
        // a divergent branch the compiler stops treating the plane descriptors
        // as uniform and wraps every buffer access in a readfirstlane loop.
        bool h_use[kHaloPer];
#pragma unroll
        for (int k = 0; k < kHaloPer; ++k) h_use[k] = tid + k * kThreads < kHalo;
        Raw<FMT> rq = raw_at(cur_r, z_begin + kR, own_ld);
        int rq_e = exp_at(cur_er, z_begin + kR, own_eld);
        Raw<FMT> rh[kHaloPer];
        int rh_e[kHaloPer];
#pragma unroll
        for (int k = 0; k < kHaloPer; ++k) {
            rh[k] = raw_at(cur_r, z_begin, h_ld[k]);
            rh_e[k] = exp_at(cur_er, z_begin, h_eld[k]);
        }
        Raw<FMT> rp = raw_stream(prev_r, z_begin), rv = raw_stream(vel_r, z_begin);
        int rp_e = exp_at(prev_er, z_begin, own_eld);
        int rv_e = exp_at(vel_er, z_begin, own_eld);
        for (int zb = z_begin; zb < z_end; zb += 9) {
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int zz = zb + j;
                const int z1 = min(zz + 1, z_last);
                const int z5 = min(zz + kR + 1, z_last);
                // before the barrier only what LDS needs: the halo of plane zz
                float4 h[kHaloPer];
#pragma unroll
                for (int k = 0; k < kHaloPer; ++k)
                    if (h_use[k]) h[k] = decode<FMT, kExact>(rh[k], rh_e[k], p_inv);
#pragma unroll
                for (int k = 0; k < kHaloPer; ++k) {
                    rh[k] = raw_at(cur_r, z1, h_ld[k]);
                    rh_e[k] = exp_at(cur_er, z1, h_eld[k]);
                }
                const int b = (zz - z_begin) & 1;
                lds[0][b][ty + kR][tx + 1] = q[(j + 4) % 9];
#pragma unroll
                for (int k = 0; k < kHaloPer; ++k)
                    if (h_use[k]) lds[0][b][h_row[k]][h_col[k]] = h[k];
                __syncthreads();
                // after it: the queue's newest plane and the streams of plane zz
                q[(j + 8) % 9] = decode<FMT, kExact>(rq, rq_e, p_inv);
                const float4 pv = decode<FMT, kExact>(rp, rp_e, p_inv);
                float4 vv;
                float dt2v = a.s.dt2;
                if constexpr (kBlk && !kExact) {
                    // fused form: fold vel2's block scale into dt^2
                    vv = dev::bfp16x4_mantissas(rv);
                    dt2v = __builtin_ldexpf(a.s.dt2, rv_e);
                } else {
                    vv = decode<FMT, kExact>(rv, rv_e, v_inv);
                }
                rq = raw_at(cur_r, z5, own_ld);
                rq_e = exp_at(cur_er, z5, own_eld);
                rp = raw_stream(prev_r, z1);
                rv = raw_stream(vel_r, z1);
                rp_e = exp_at(prev_er, z1, own_eld);
                rv_e = exp_at(vel_er, z1, own_eld);
                const float4 zq[9] = {q[j % 9],       q[(j + 1) % 9], q[(j + 2) % 9],
                                      q[(j + 3) % 9], q[(j + 4) % 9], q[(j + 5) % 9],
                                      q[(j + 6) % 9], q[(j + 7) % 9], q[(j + 8) % 9]};
                finish(ModeC{}, zz, b, zq, pv, pv, vv, dt2v, pv, pv, pv);
            }
        }
    }
}

template <int TXV, int TY, int MODE, bool NT, int FMT = dev::kF32>
__global__ __launch_bounds__(TXV* TY) __attribute__((amdgpu_waves_per_eu(min_waves(TXV, TY, FMT, MODE))))
void zq_kernel(ZqArgs a) {
    zq_tile<TXV, TY, MODE, NT, FMT>(a, blockIdx.x);
}

// The whole ragged batch run in one launch.  Work item j of XCD run r
// (r = 0..7) is item 8 j + r of the interleaved item table, run_len =
// ceil(n_items / 8) items per run.  In each step a workgroup of run r
// (blockIdx.x % 8 == r, so it sits on that run's XCD) takes the next item of
// its run from a queue, as the dispatcher of a plain launch would hand it
// the next workgroup, until the run is empty; then the grid meets at a
// barrier before the next step reads what this one wrote.
//
// The counters only grow: per step, run r's queue hands out run_len items
// and one empty slot to each of its G / 8 workgroups, so step n's slots start
// at n (run_len + G / 8); the barrier counter reaches (n + 1) G when every
// workgroup has finished step n.  One lane per workgroup spins, sleeping
// between polls, for at most max_polls polls (by default kBarrierPolls,
// seconds; a step takes milliseconds): the host launches no more workgroups
// than stay resident
// together, and a barrier that still times out sets the abort flag, which
// makes every workgroup leave at its next barrier, so the kernel always
// ends.  The fences order this workgroup's stores of the step before its
// arrival and the next step's loads after the barrier, at device scope
// (across the XCDs' L2s).
//
// sync: [0] abort flag, [1] barrier counter, [2 + r] queue of run r; zero
// first.
__device__ bool grid_barrier(unsigned* sync, unsigned target, unsigned max_polls) {
    __shared__ int s_ok;
    __syncthreads();  // every lane's stores of the step issued
    if (threadIdx.x == 0) {
        __threadfence();
        __hip_atomic_fetch_add(sync + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        for (unsigned polls = 0; __hip_atomic_load(sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target;
             ++polls) {
            if (polls >= max_polls ||
                ((polls & 63u) == 0 && __hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(sync, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

template <int TXV, int TY, bool NT>
__global__ __launch_bounds__(TXV* TY) __attribute__((amdgpu_waves_per_eu(min_waves(TXV, TY, dev::kF32, kWaveFwi))))
void zq_persist_kernel(const RaggedShot* shots, const RaggedItem* items, int n_steps, int n_items, int alternate,
                       unsigned* sync, unsigned max_polls) {
    __shared__ int s_item;
    const unsigned run = blockIdx.x % 8u;
    const unsigned run_len = (static_cast<unsigned>(n_items) + 7u) / 8u;
    const unsigned per_step = run_len + gridDim.x / 8u;  // queue slots of one step
    unsigned* const queue = sync + 2 + run;
    for (int n = 0; n < n_steps; ++n) {
        ZqArgs s{};  // fields the item does not set stay known zeros (no trace, no L p output)
        s.items = items;
        s.rshots = shots;
        s.rot = n % 3;
        s.step = n;
        s.mirror = alternate ? (n & 1) : 0;
        const unsigned first = static_cast<unsigned>(n) * per_step;
        for (;;) {
            if (threadIdx.x == 0) {
                const unsigned k =
                    __hip_atomic_fetch_add(queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - first;
                s_item = k < run_len ? static_cast<int>(k * 8u + run) : -1;
            }
            __syncthreads();
            // one value for the workgroup: readfirstlane tells the compiler so
            const int item = __builtin_amdgcn_readfirstlane(s_item);
            __syncthreads();  // s_item read before lane 0 writes the next
            if (item < 0) break;
            if (item < n_items) zq_tile<TXV, TY, kWaveFwi, NT>(s, static_cast<unsigned>(item));
        }
        if (n + 1 < n_steps && !grid_barrier(sync, static_cast<unsigned>(n + 1) * gridDim.x, max_polls)) return;
    }
}

using KernelFn = void (*)(ZqArgs);
using PersistFn = void (*)(const RaggedShot*, const RaggedItem*, int, int, int, unsigned*, unsigned);

template <bool NT>
PersistFn pick_persist(int txv, int ty) {
    if (txv == 16 && ty == 16) return zq_persist_kernel<16, 16, NT>;
    if (txv == 16 && ty == 32) return zq_persist_kernel<16, 32, NT>;
    if (txv == 32 && ty == 8) return zq_persist_kernel<32, 8, NT>;
    if (txv == 32 && ty == 16) return zq_persist_kernel<32, 16, NT>;
    if (txv == 64 && ty == 4) return zq_persist_kernel<64, 4, NT>;
    if (txv == 64 && ty == 8) return zq_persist_kernel<64, 8, NT>;
    if (txv == 64 && ty == 12) return zq_persist_kernel<64, 12, NT>;
    if (txv == 64 && ty == 16) return zq_persist_kernel<64, 16, NT>;
    return nullptr;
}

// Tile shapes (float4 columns x rows) the library is built with.
struct TileShape {
    int txv, ty;
};
constexpr TileShape kShapes[] = {{16, 16}, {16, 32}, {32, 8}, {32, 16}, {64, 4}, {64, 8}, {64, 12}, {64, 16}};

template <int MODE, bool NT>
KernelFn pick_shape(int txv, int ty) {
    if (txv == 16 && ty == 16) return zq_kernel<16, 16, MODE, NT>;
    if (txv == 16 && ty == 32) return zq_kernel<16, 32, MODE, NT>;
    if (txv == 32 && ty == 8) return zq_kernel<32, 8, MODE, NT>;
    if (txv == 32 && ty == 16) return zq_kernel<32, 16, MODE, NT>;
    if (txv == 64 && ty == 8) return zq_kernel<64, 8, MODE, NT>;
    if (txv == 64 && ty == 12) return zq_kernel<64, 12, MODE, NT>;
    if constexpr (MODE != kAdjoint) {  // the adjoint launcher maps these two to 64 x 8
        if (txv == 64 && ty == 4) return zq_kernel<64, 4, MODE, NT>;
        if (txv == 64 && ty == 16) return zq_kernel<64, 16, MODE, NT>;
    }
    return nullptr;
}

// The step pair's tiles: two queues and two LDS planes per workgroup fit 256
// VGPRs and 68 KB of LDS at these shapes.
constexpr TileShape kPairShapes[] = {{32, 8}, {32, 16}, {64, 4}, {64, 8}};

template <bool NT>
KernelFn pick_pair(int txv, int ty) {
    if (txv == 32 && ty == 8) return zq_kernel<32, 8, kPair, NT>;
    if (txv == 32 && ty == 16) return zq_kernel<32, 16, kPair, NT>;
    if (txv == 64 && ty == 4) return zq_kernel<64, 4, kPair, NT>;
    if (txv == 64 && ty == 8) return zq_kernel<64, 8, kPair, NT>;
    return nullptr;
}

KernelFn pick(int txv, int ty, int mode, bool nt) {
    switch (mode) {
        case kPair: return nt ? pick_pair<true>(txv, ty) : pick_pair<false>(txv, ty);
        case kWaveFwi: return nt ? pick_shape<kWaveFwi, true>(txv, ty) : pick_shape<kWaveFwi, false>(txv, ty);
        case kAdjoint: return nt ? pick_shape<kAdjoint, true>(txv, ty) : pick_shape<kAdjoint, false>(txv, ty);
        case kWaveFused: return nt ? pick_shape<kWaveFused, true>(txv, ty) : pick_shape<kWaveFused, false>(txv, ty);
        case kWaveExact: return nt ? pick_shape<kWaveExact, true>(txv, ty) : pick_shape<kWaveExact, false>(txv, ty);
        case kHeatFused: return nt ? pick_shape<kHeatFused, true>(txv, ty) : pick_shape<kHeatFused, false>(txv, ty);
        default: return nt ? pick_shape<kHeatExact, true>(txv, ty) : pick_shape<kHeatExact, false>(txv, ty);
    }
}

template <int MODE, int FMT, bool NT>
KernelFn pick_shape16(int txv, int ty) {
    if (txv == 16 && ty == 16) return zq_kernel<16, 16, MODE, NT, FMT>;
    if (txv == 16 && ty == 32) return zq_kernel<16, 32, MODE, NT, FMT>;
    if (txv == 32 && ty == 8) return zq_kernel<32, 8, MODE, NT, FMT>;
    if (txv == 32 && ty == 16) return zq_kernel<32, 16, MODE, NT, FMT>;
    if (txv == 64 && ty == 4) return zq_kernel<64, 4, MODE, NT, FMT>;
    if (txv == 64 && ty == 8) return zq_kernel<64, 8, MODE, NT, FMT>;
    if (txv == 64 && ty == 12) return zq_kernel<64, 12, MODE, NT, FMT>;
    if (txv == 64 && ty == 16) return zq_kernel<64, 16, MODE, NT, FMT>;
    return nullptr;
}

template <int MODE>
KernelFn pick16_mode(int txv, int ty, int fmt, bool nt) {
    switch (fmt) {
        case dev::kBf16:
            return nt ? pick_shape16<MODE, dev::kBf16, true>(txv, ty) : pick_shape16<MODE, dev::kBf16, false>(txv, ty);
        case dev::kFp16s:
            return nt ? pick_shape16<MODE, dev::kFp16s, true>(txv, ty) : pick_shape16<MODE, dev::kFp16s, false>(txv, ty);
        case dev::kBfp16:
            return nt ? pick_shape16<MODE, dev::kBfp16, true>(txv, ty) : pick_shape16<MODE, dev::kBfp16, false>(txv, ty);
        default: return nullptr;
    }
}

KernelFn pick16(int txv, int ty, int fmt, int mode, bool nt) {
    return mode == kWaveExact ? pick16_mode<kWaveExact>(txv, ty, fmt, nt)
                              : pick16_mode<kWaveFused>(txv, ty, fmt, nt);
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

// Sweep direction of a launch: the opposite of the launch that wrote this
// one's input field (a chained time step), else upwards.  Remembers the last 8
// launches of the calling thread; STENCIL_ALTERNATE_SWEEP=0 keeps every launch
// upwards.  Direction never changes results, only which planes come first.
bool sweep_alternation() {
    static const bool enabled = [] {
        const char* e = std::getenv("STENCIL_ALTERNATE_SWEEP");
        return !(e && e[0] == '0');
    }();
    return enabled;
}

int sweep_direction(const void* cur, const void* next) {
    struct Recent {
        const void* next;
        int mirror;
    };
    static thread_local Recent recent[8] = {};
    static thread_local int head = 0;
    if (!sweep_alternation() || !cur) return 0;
    int m = 0;
    for (const Recent& r : recent)
        if (r.next == cur) {
            m = 1 - r.mirror;
            break;
        }
    recent[head] = Recent{next, m};
    head = (head + 1) & 7;
    return m;
}

hipError_t launch_zq(ZqArgs a, const Dims& d, ZqShape shape, int mode, hipStream_t st,
                     int fmt = dev::kF32, int n_shots = 1) {
    if (d.nx <= 2 * kR || d.ny <= 2 * kR || d.nz <= 2 * kR) return hipSuccess;
    shape = zq_resolve(d, shape, fmt != dev::kF32 ? 2 + fmt : (mode == kHeatFused || mode == kHeatExact ? 1 : 0));
    KernelFn fn = fmt == dev::kF32 ? pick(shape.txv, shape.ty, mode, shape.nt != 0)
                                   : pick16(shape.txv, shape.ty, fmt, mode, shape.nt != 0);
    if (!fn) return hipErrorInvalidValue;
    const int64_t inz = d.nz - 2 * kR;
    a.nx = static_cast<int>(d.nx);
    a.ny = static_cast<int>(d.ny);
    a.nz = static_cast<int>(d.nz);
    a.tiles_x = static_cast<int>((d.nx + 4 * shape.txv - 1) / (4 * shape.txv));
    a.tiles_y = static_cast<int>((d.ny - 2 * kR + shape.ty - 1) / shape.ty);
    a.chunk = zq_chunk_planes(inz, shape.n_chunks);
    // A workgroup's buffer resources span its chunk plus 21 planes (queue fill
    // and prefetch); keep that below 2^32 bytes with shorter chunks if needed.
    const int64_t max_chunk = (int64_t(0xFFFFFFFF) / (d.nx * d.ny * 4) - 21) / 9 * 9;
    if (max_chunk < 9) return hipErrorInvalidValue;  // zq_supported() excludes such planes
    if (a.chunk > max_chunk) a.chunk = static_cast<int>(max_chunk);
    const int64_t n_chunks = (inz + a.chunk - 1) / a.chunk;
    const int64_t nwg = int64_t(a.tiles_x) * a.tiles_y * n_chunks;
    a.xcd_remap = shape.xcd_remap != 0 ? 1 : 0;
    if (a.batch) {
        a.mirror = sweep_alternation() ? (a.step & 1) : 0;  // one shot's nodes chain step after step
    } else {
        a.mirror = sweep_direction(a.cur, a.next);
        if (a.mirror) {
            a.src_z = a.nz - 1 - a.src_z;
            a.rec_z = a.nz - 1 - a.rec_z;
        }
    }
    hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(nwg), static_cast<unsigned>(n_shots)), dim3(shape.txv * shape.ty),
                       0, st, a);
    return hipGetLastError();
}

}  // namespace

// Chunks are whole groups of 9 planes (the kernel's step group), so only the
// last chunk of a column runs idle steps.
int zq_chunk_planes(int64_t interior_planes, int n_chunks) {
    if (n_chunks < 1) n_chunks = 1;
    return static_cast<int>(((interior_planes + n_chunks - 1) / n_chunks + 8) / 9 * 9);
}

bool zq_supported(const Dims& d, const void* const* ptrs, int n_ptrs) {
    if (d.nx % 4 != 0) return false;
    // 32-bit byte offsets within a workgroup's planes: at least 30 planes
    // (one 9-plane chunk + 21) must fit in 2^32 bytes.
    if (d.nx * d.ny * 4 * 30 >= (int64_t(1) << 32)) return false;
    for (int i = 0; i < n_ptrs; ++i)
        if (reinterpret_cast<uintptr_t>(ptrs[i]) % 16 != 0) return false;
    return true;
}

bool zq_shape_valid(int txv, int ty) {
    for (const auto& s : kShapes)
        if (s.txv == txv && s.ty == ty) return true;
    return false;
}

bool zq_pair_shape_valid(int txv, int ty) {
    for (const auto& s : kPairShapes)
        if (s.txv == txv && s.ty == ty) return true;
    return false;
}

ZqShape zq_resolve(const Dims& d, ZqShape s, int kind) {
    if (!zq_shape_valid(s.txv, s.ty)) {
        // Widest tile whose row padding past nx is smallest (ties go wide):
        // 256-cell rows at nx = 512, 128 at nx = 128, ...  Rows: 16-bit storage
        // and wide grids take 64 x 16 (one 1024-thread workgroup per CU),
        // otherwise 512 threads per workgroup.
        int best = 16;
        int64_t best_waste = -1;
        for (int t : {64, 32, 16}) {
            const int64_t w = ((d.nx + 4 * t - 1) / (4 * t)) * (4 * t) - d.nx;
            if (best_waste < 0 || w < best_waste) {
                best = t;
                best_waste = w;
            }
        }
        if (kind == 1 && best == 64) best = 32;  // heat: 128 x 16 cells (same padding)
        s.txv = best;
        s.ty = best == 64 ? ((kind >= 2 && kind <= 5) || d.nx >= 1024 ? 16 : 8) : best == 32 ? 16 : 32;
    }
    // 64 x 16 (1024 threads, 128 VGPRs) and 64 x 4 (three halo items per
    // thread) spill with the adjoint's extra streams: it runs 64 x 8 instead.
    if (kind == kKindFwiAdjoint && s.txv == 64 && (s.ty == 16 || s.ty == 4)) {
        s.ty = 8;
        s.n_chunks = 0;
    }
    // The step pair is built for four tiles: 64-wide tiles keep 64 x 4 or
    // become 64 x 8, narrower ones 32 x 8 or 32 x 16.
    if (kind == kKindFwiPair && !zq_pair_shape_valid(s.txv, s.ty)) {
        s.ty = s.txv == 64 ? 8 : (s.txv == 32 && s.ty == 8 ? 8 : 16);
        s.txv = s.txv == 64 ? 64 : 32;
        s.n_chunks = 0;
    }
    if (s.nt < 0) s.nt = 1;
    const int64_t inz = d.nz - 2 * kR;
    if (s.n_chunks <= 0) {
        // z-chunks for about one workgroup per CU (256 CUs), chunks of at
        // least 32 planes: long z-marches stream best.  These are the
        // autotuner's picks at 512^3 (wave 256 x 8 cells, heat 128 x 16,
        // both 2 chunks).
        const int64_t tiles =
            ((d.nx + 4 * s.txv - 1) / (4 * s.txv)) * ((d.ny - 2 * kR + s.ty - 1) / s.ty);
        const int64_t target = 256;
        int64_t want = (target + tiles / 2) / tiles;
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
                      float v_scale, bool exact, ZqShape shape, hipStream_t st) {
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
    if (!exact) {
        // Fused forms run on the stored values: p in its scaled domain (the
        // update is linear in p), the vel2 scale folded into dt^2.
        if (fmt == dev::kFp16s) a.s.dt2 = s.dt2 * a.v_inv;
        fill_weights(a, 1.0);
    }
    return launch_zq(a, d, shape, exact ? kWaveExact : kWaveFused, st, fmt);
}

hipError_t wave_zq_fwi_f32(const float* p, const float* pp, const float* m, float* pn, float* lap_out,
                           int src_z, int src_y, int src_x, const float* src_val, float* rec_plane, int rec_z,
                           const Dims& d, const Scalars<float>& s, ZqShape shape, hipStream_t st) {
    ZqArgs a{};
    a.trace = rec_plane;
    a.rec_z = rec_z;
    a.cur = p;
    a.prev = pp;
    a.vel2 = m;
    a.next = pn;
    a.lap_out = lap_out;
    a.src_val = src_val;
    a.src_z = src_z;
    a.src_y = src_y;
    a.src_x = src_x;
    a.s = s;
    if (shape.nt < 0) shape.nt = 1;
    return launch_zq(a, d, shape, kWaveFwi, st);
}

hipError_t batch_zq_fwi_f32(const BatchShot* table, int n_shots, int rot, int step, const Dims& d,
                            const Scalars<float>& s, ZqShape shape, hipStream_t st) {
    ZqArgs a{};
    a.batch = table;
    a.rot = rot;
    a.step = step;
    a.s = s;
    shape.nt = 1;
    return launch_zq(a, d, shape, kWaveFwi, st, dev::kF32, n_shots);
}

hipError_t ragged_zq_fwi_f32(const RaggedShot* shots, const RaggedItem* items, int n_items, int rot, int step,
                             int txv, int ty, bool nt, hipStream_t st) {
    if (n_items <= 0) return hipSuccess;
    KernelFn fn = pick(txv, ty, kWaveFwi, nt);
    if (!fn || !shots || !items) return hipErrorInvalidValue;
    ZqArgs a{};
    a.items = items;
    a.rshots = shots;
    a.rot = rot;
    a.step = step;
    a.mirror = sweep_alternation() ? (step & 1) : 0;  // every shot chains step after step
    hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(n_items)), dim3(txv * ty), 0, st, a);
    return hipGetLastError();
}

int ragged_persist_grid(int txv, int ty) {
    PersistFn fn = pick_persist<true>(txv, ty);
    if (!fn) return 0;
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), txv * ty, 0) !=
            hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return (per_cu * cus) / 8 * 8;
}

hipError_t ragged_persist_fwi_f32(const RaggedShot* shots, const RaggedItem* items, int n_items, int n_steps, int txv,
                                  int ty, int grid, unsigned* sync, unsigned max_polls, hipStream_t st) {
    if (n_items <= 0 || n_steps <= 0) return hipSuccess;
    PersistFn fn = pick_persist<true>(txv, ty);
    // the queues ((run_len + grid / 8) slots per step) and the barrier
    // counter (grid per step) stay below 2^32
    const uint64_t bound = (static_cast<uint64_t>(n_items) + static_cast<uint64_t>(grid)) * static_cast<uint64_t>(n_steps);
    if (!fn || !shots || !items || !sync || grid < 8 || grid % 8 != 0 || bound >= (uint64_t(1) << 32))
        return hipErrorInvalidValue;
    const int alternate = sweep_alternation() ? 1 : 0;
    hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(grid)), dim3(static_cast<unsigned>(txv * ty)), 0, st, shots,
                       items, n_steps, n_items, alternate, sync, max_polls);
    return hipGetLastError();
}

hipError_t adjoint_zq_f32(const float* psi1, const float* psi2, const float* m, const float* lap, const float* lap2,
                          float* grad, float* psi_out, const float* res, int rec_z, int src_z, int src_y, int src_x,
                          float* trace, const Dims& d, const Scalars<float>& s, ZqShape shape, hipStream_t st) {
    if (lap2 && !lap) return hipErrorInvalidValue;
    ZqArgs a{};
    a.cur = psi1;
    a.prev = psi2;
    a.vel2 = m;
    a.aux = lap;
    a.aux2 = lap2;
    a.grad = grad;
    a.next = psi_out;
    a.res = res;
    a.rec_z = rec_z;
    a.trace = trace;
    a.src_z = src_z;
    a.src_y = src_y;
    a.src_x = src_x;
    a.s = s;
    if (shape.nt < 0) shape.nt = 1;
    return launch_zq(a, d, zq_resolve(d, shape, kKindFwiAdjoint), kAdjoint, st);
}

hipError_t pair_zq_f32(const float* psi1, const float* psi2, const float* m, const float* lap, const float* lap2,
                       float* grad, float* psi_out, const float* res, int rec_z, int src_z, int src_y, int src_x,
                       float* trace, const float* p, const float* pp, float* pn, float* lap_out,
                       const float* src_val, const Dims& d, const Scalars<float>& s, ZqShape shape, hipStream_t st) {
    if ((lap2 && !lap) || !p || !pp || !pn) return hipErrorInvalidValue;
    ZqArgs a{};
    a.cur = psi1;
    a.prev = psi2;
    a.vel2 = m;
    a.aux = lap;
    a.aux2 = lap2;
    a.grad = grad;
    a.next = psi_out;
    a.res = res;
    a.rec_z = rec_z;
    a.trace = trace;
    a.cur_b = p;
    a.prev_b = pp;
    a.next_b = pn;
    a.lap_out = lap_out;
    a.src_val = src_val;
    a.src_z = src_z;
    a.src_y = src_y;
    a.src_x = src_x;
    a.s = s;
    if (shape.nt < 0) shape.nt = 1;
    return launch_zq(a, d, zq_resolve(d, shape, kKindFwiPair), kPair, st);
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
