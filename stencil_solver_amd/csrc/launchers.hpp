// Host-side launchers of the gfx950 stencil kernels (internal to the library;
// the public surface is include/stencil_capi.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <vector>

namespace stencil::hipk {

// Total cells per axis, halo included.
struct Dims {
    int64_t nx = 0, ny = 0, nz = 0;
};

// The per-launch scalars the reference launchers form on the host
// (src/kernels/stencil_naive.cu:96-99): inv_h2 = 1/(h*h), dt2 = dt*dt in T.
template <class T>
struct Scalars {
    T inv_dx2, inv_dy2, inv_dz2, dt2;
};

template <class T>
Scalars<T> make_scalars(T dx, T dy, T dz, T dt) {
    Scalars<T> s;
    s.inv_dx2 = T(1) / (dx * dx);
    s.inv_dy2 = T(1) / (dy * dy);
    s.inv_dz2 = T(1) / (dz * dz);
    s.dt2 = dt * dt;
    return s;
}

// Launch shape of the LDS + z-register-queue kernel.  Zero fields = choose
// automatically (see zq_resolve).
struct ZqShape {
    int txv = 0;       // tile width in float4 (tile is 4*txv x ty cells, txv*ty threads)
    int ty = 0;        // tile rows; (txv, ty) must be one of the built shapes
    int n_chunks = 0;  // z-chunks per column (each re-reads 8 halo planes)
    int xcd_remap = -1;  // 1 = XCD-contiguous tile order, 0 = off, -1 = auto
    int nt = -1;       // 1 = non-temporal p_prev / vel2 / p_next, 0 = cached, -1 = auto
};

// Whether the kernel can take these dims and pointers (nx % 4 == 0, 16-byte
// aligned bases, 30 planes below 2^32 bytes: nx * ny < 35.7 M cells).
bool zq_supported(const Dims& d, const void* const* ptrs, int n_ptrs);
// Whether (txv, ty) is a tile shape the library is built with.
bool zq_shape_valid(int txv, int ty);
// Shape-cache kinds: 0 wave step, 1 heat step, 2 + storage the 16-bit wave
// step, then the FWI kernels (forward, adjoint, recompute, step pair).
constexpr int kKindFwiForward = 6;
constexpr int kKindFwiAdjoint = 7;
constexpr int kKindFwiRecompute = 8;  // forward steps beside the adjoint steps
constexpr int kKindFwiPair = 9;       // the backward pass's step pair (recompute + adjoint, one launch)
// Whether (txv, ty) is a tile shape the step pair is built with.
bool zq_pair_shape_valid(int txv, int ty);
// Fills unset fields of a shape for a kernel of this kind; the adjoint kind
// also maps the tiles it is not built for (64 x 4, 64 x 16) to 64 x 8.
ZqShape zq_resolve(const Dims& d, ZqShape s, int kind);
// z-planes per chunk of the tiled kernel for n_chunks chunks of a column.
int zq_chunk_planes(int64_t interior_planes, int n_chunks);

// Autotuner cache (tune.cpp): the fastest shape measured for a grid size
// and kernel kind.
bool zq_cached_shape(const Dims& d, int kind, ZqShape* out);
void zq_cache_put(const Dims& d, int kind, const ZqShape& s);
void zq_cache_clear();
// Tuner helpers: the launch shapes swept for these dims, and the time of one
// call of `fn` (ms per call over reps calls after 2 warmups; < 0 if a call
// failed).
std::vector<ZqShape> zq_tune_candidates(const Dims& d);
float zq_time_launches(const std::function<hipError_t()>& fn, int reps, hipStream_t st);

// The acoustic step.  `exact` = the reference association (bit-identical to
// its CPU path); otherwise FMA-fused weights.
hipError_t wave_naive_f32(const float* cur, const float* prev, const float* vel2, float* next,
                          const Dims& d, const Scalars<float>& s, hipStream_t st);
hipError_t wave_naive_f64(const double* cur, const double* prev, const double* vel2,
                          double* next, const Dims& d, const Scalars<double>& s, hipStream_t st);
hipError_t wave_zq_f32(const float* cur, const float* prev, const float* vel2, float* next,
                       const Dims& d, const Scalars<float>& s, bool exact, ZqShape shape,
                       hipStream_t st);

// A field in 16-bit storage: data = nx*ny*nz uint16; exp = nx*ny*nz/16 int8
// block exponents (block floating point only).
struct Field16 {
    void* data;
    int8_t* exp;
};
// The acoustic step on 16-bit fields (fmt: 1 bf16, 2 scaled fp16, 3 block
// fp16), FP32 arithmetic in the reference association (exact) or the
// fused-weight form.  p_scale applies to p_cur / p_prev / p_next, v_scale to
// vel2 (powers of two; fp16 only).
hipError_t wave_zq_16(int fmt, const Field16& cur, const Field16& prev, const Field16& vel2,
                      const Field16& next, const Dims& d, const Scalars<float>& s, float p_scale,
                      float v_scale, bool exact, ZqShape shape, hipStream_t st);
// FWI steps on the tiled kernel (FP32, reference association; bit-identical
// to the global-memory FWI kernels).  Forward: p_next = 2p - pp + dt^2 m L p
// (+ *src_val at cell (src_z, src_y, src_x) when src_val is non-null), L p
// into lap_out when non-null, the interior of plane rec_z of p_next into
// rec_plane (nx*ny, halo left as is) when non-null.  Adjoint, in psi = m lambda:
// psi_out = 2 psi1 - psi2 + dt^2 m L psi1 (+ m res on plane rec_z when res is
// non-null); grad += dt^2 lap psi1 when lap is non-null, preceded by
// grad += dt^2 lap2 psi2 when lap2 (L p^{k+1}, needs lap) is non-null: the
// gradient terms of two steps in one pass, added as the two steps would;
// *trace = psi_out / m at the source cell when trace is non-null.
hipError_t wave_zq_fwi_f32(const float* p, const float* pp, const float* m, float* pn, float* lap_out,
                           int src_z, int src_y, int src_x, const float* src_val, float* rec_plane, int rec_z,
                           const Dims& d, const Scalars<float>& s, ZqShape shape, hipStream_t st);
hipError_t adjoint_zq_f32(const float* psi1, const float* psi2, const float* m, const float* lap, const float* lap2,
                          float* grad, float* psi_out, const float* res, int rec_z, int src_z, int src_y, int src_x,
                          float* trace, const Dims& d, const Scalars<float>& s, ZqShape shape, hipStream_t st);
// The backward pass's step pair: one launch runs the adjoint step above
// (psi1 .. trace) and the forward step p, pp -> pn with L p into lap_out and
// *src_val at the source cell (no receiver plane) on the same tiles, reading
// m once; each field gets the bits the two single steps would give it.  The
// two steps' fields must be distinct buffers.
hipError_t pair_zq_f32(const float* psi1, const float* psi2, const float* m, const float* lap, const float* lap2,
                       float* grad, float* psi_out, const float* res, int rec_z, int src_z, int src_y, int src_x,
                       float* trace, const float* p, const float* pp, float* pn, float* lap_out,
                       const float* src_val, const Dims& d, const Scalars<float>& s, ZqShape shape, hipStream_t st);
// One shot of a batched forward launch (device table entry): three time-level
// buffers, vel2, the source cell and its wavelet (nullable: no source).
struct BatchShot {
    float* ring[3];
    const float* vel2;
    const float* wavelet;
    int src_z, src_y, src_x, pad;
};
// One shot of a ragged batched launch (shots of any dims in one launch): its
// buffers and source as in BatchShot, plus its own dims, tiling (tiles per
// plane row / column, z-chunk length) and scalars.
struct RaggedShot {
    float* ring[3];
    const float* vel2;
    const float* wavelet;
    int src_z, src_y, src_x;
    int nx, ny, nz;
    int tiles_x, tiles_y, chunk;
    Scalars<float> s;
};
// One workgroup of a ragged launch: tile (tile_x, tile_y) of z-chunk chunk_id
// of shot `shot`; shot < 0 = no work (padding of a shorter per-XCD list).
struct RaggedItem {
    int shot, tile_x, tile_y, chunk_id;
};
// The forward FWI step of every shot of a ragged batch in one launch of
// n_items workgroups of txv * ty threads: workgroup b steps items[b], shot i
// stepping ring[rot] / ring[rot+1] -> ring[rot+2] (mod 3) and adding
// wavelet[step].  Every shot's tiling must be the one its items were built
// with for this (txv, ty).
hipError_t ragged_zq_fwi_f32(const RaggedShot* shots, const RaggedItem* items, int n_items, int rot, int step,
                             int txv, int ty, bool nt, hipStream_t st);
// All n_steps steps of the same ragged batch in ONE launch of `grid`
// workgroups (a multiple of 8, at most ragged_persist_grid(): all resident
// together) that take each step's items from per-XCD queues and meet at a
// grid barrier between steps; step n as ragged_zq_fwi_f32(..., n % 3, n, ...)
// with non-temporal streams.  sync: 10 zeroed words (abort flag, barrier
// counter, 8 queues); the flag is non-zero after the run when a barrier
// timed out (the fields are then not valid).  A workgroup gives up waiting at
// a barrier after max_polls polls (kBarrierPolls: seconds).
constexpr unsigned kBarrierPolls = 1u << 22;
int ragged_persist_grid(int txv, int ty);
hipError_t ragged_persist_fwi_f32(const RaggedShot* shots, const RaggedItem* items, int n_items, int n_steps, int txv,
                                  int ty, int grid, unsigned* sync, unsigned max_polls, hipStream_t st);
// The forward FWI step of n_shots shots of equal dims in one launch: shot i
// steps ring[rot] / ring[rot+1] -> ring[rot+2] (mod 3), adding wavelet[step].
hipError_t batch_zq_fwi_f32(const BatchShot* table, int n_shots, int rot, int step, const Dims& d,
                            const Scalars<float>& s, ZqShape shape, hipStream_t st);
// Elementwise conversion between FP32 and a 16-bit format (n % 16 == 0).
hipError_t encode16(int fmt, const float* src, Field16 dst, size_t n, float scale, hipStream_t st);
hipError_t decode16(int fmt, Field16 src, float* dst, size_t n, float scale, hipStream_t st);

// The heat step t_next = t + coef * Lap(t), coef = dt2 * alpha formed in float
// (the reference's dt2 * vel2[c] with vel2 == alpha).
hipError_t heat_naive_f32(const float* t, float* tn, const Dims& d, const Scalars<float>& s,
                          float coef, hipStream_t st);
hipError_t heat_naive_f64(const double* t, double* tn, const Dims& d, const Scalars<double>& s,
                          double coef, hipStream_t st);
hipError_t heat_zq_f32(const float* t, float* tn, const Dims& d, const Scalars<float>& s,
                       float coef, bool exact, ZqShape shape, hipStream_t st);

// STENCIL_OK, or STENCIL_ERR_HIP with e kept for stencil_last_hip_error().
int hip_status(hipError_t e);

}  // namespace stencil::hipk
