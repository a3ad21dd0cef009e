/*
 * include/stencil_capi.h -- the C-ABI of the MI355X stencil library
 * (libstencil_hip.so, built from stencil_solver_amd/csrc/).
 *
 * Plain pointers and sizes only.  Every launcher is asynchronous on the HIP
 * stream it is given (NULL = the default stream), never synchronises, never
 * allocates, never exits, and returns STENCIL_OK or an error code
 * (stencil_strerror() names it).  Buffers are caller-owned device pointers in
 * the reference layout: x fastest, idx = z*nx*ny + y*nx + x, dims include a
 * radius-4 halo on every face that the launchers never write
 * (include/stencil/types.hpp:110-136 of the reference).  p_next must not
 * alias an input.  Beyond the launchers: the tuners synchronise, the batch
 * object allocates its small pointer tables at create, and the FWI backward
 * pass runs on a second stream of its own (created once per device) joined
 * to the caller's by events it creates per call; device memory beyond the
 * inputs and outputs is the caller's (workspaces), except for the static
 * per-device partial-sum array of stencil_energy_f32_noalloc.
 */
#ifndef STENCIL_CAPI_H
#define STENCIL_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STENCIL_CAPI_RADIUS 4

/* Grid descriptor, the C form of stencil::GridDims (reference
 * include/stencil/types.hpp:115-140).  Spacings and dt are carried in double;
 * the f32 launchers narrow them to float first, so a float GridDims passes
 * through unchanged. */
typedef struct stencil_grid {
    uint64_t nx, ny, nz; /* total cells, halos included */
    double dx, dy, dz;   /* spacing [m] */
    double dt;           /* time step [s] */
} stencil_grid;

/* KernelVariant (reference include/stencil/config.hpp:25-29). */
enum stencil_variant {
    STENCIL_VARIANT_NAIVE = 0,    /* coalesced global-memory kernel */
    STENCIL_VARIANT_TILED = 1,    /* LDS plane + z register queue */
    STENCIL_VARIANT_MULTIGPU = 2  /* one GPU here: same as TILED (kernels.cuh:78-81) */
};

enum stencil_status {
    STENCIL_OK = 0,
    STENCIL_ERR_INVALID = 1,     /* bad dims / null pointer / bad enum */
    STENCIL_ERR_HIP = 2,         /* a HIP runtime call failed */
    STENCIL_ERR_UNSUPPORTED = 3, /* shape the kernel cannot take */
    STENCIL_ERR_NOMEM = 4,
    STENCIL_ERR_TIMEOUT = 5      /* a device-side wait gave up (results invalid) */
};

const char* stencil_strerror(int status);
/* The HIP error code behind the last STENCIL_ERR_HIP on this thread. */
int stencil_last_hip_error(void);

/* Replaces upload_fd_coefficients() (reference include/stencil/kernels.cuh:33,
 * src/kernels/stencil_naive.cu:38-46).  The coefficients are compile-time
 * constants of the gfx950 kernels, so this only checks that a device is
 * usable; kept so callers written against the reference still link. */
int stencil_upload_fd_coefficients(void);

/* Replaces launch_naive (reference include/stencil/kernels.cuh:42-43,
 * src/kernels/stencil_naive.cu:92-114): one acoustic leapfrog step
 *   p_next = 2 p_cur - p_prev + dt^2 vel2 Lap(p_cur)
 * over the interior, in the reference's exact association (bit-identical to
 * its CPU path built without FMA contraction). */
int stencil_launch_naive_f32(const float* p_cur, const float* p_prev, const float* vel2,
                             float* p_next, const stencil_grid* grid, void* hip_stream);
/* Replaces launch_tiled (reference include/stencil/kernels.cuh:54-55,
 * src/kernels/stencil_tiled.cu:147-168): same step, LDS-staged plane + z
 * register queue, same association (bit-identical to launch_naive).  Shapes
 * the vector kernel cannot take (nx % 4 != 0, unaligned bases) run the
 * coalesced kernel. */
int stencil_launch_tiled_f32(const float* p_cur, const float* p_prev, const float* vel2,
                             float* p_next, const stencil_grid* grid, void* hip_stream);
/* Replaces launch(KernelVariant, ...) (reference include/stencil/kernels.cuh:68-83). */
int stencil_launch_f32(int variant, const float* p_cur, const float* p_prev, const float* vel2,
                       float* p_next, const stencil_grid* grid, void* hip_stream);

/* The STENCIL_FP64 build of the same three (reference types.hpp:33-37);
 * the FP64 oracle run. */
int stencil_launch_naive_f64(const double* p_cur, const double* p_prev, const double* vel2,
                             double* p_next, const stencil_grid* grid, void* hip_stream);
int stencil_launch_tiled_f64(const double* p_cur, const double* p_prev, const double* vel2,
                             double* p_next, const stencil_grid* grid, void* hip_stream);
int stencil_launch_f64(int variant, const double* p_cur, const double* p_prev,
                       const double* vel2, double* p_next, const stencil_grid* grid,
                       void* hip_stream);

/* Replaces the heat step of HeatDiffusionSolver::step_impl (reference
 * src/solver/heat_diffusion.cpp:103-122: wave kernel with vel2 = alpha and
 * p_prev = p_cur, then a full-field D2D copy) with one single-level kernel:
 *   t_next = t + (dt^2 alpha) Lap(t)
 * where grid->dt is the heat solver's dt_eff = (float)sqrt((double)dt)
 * (heat_diffusion.cpp:86-87).  Bit-identical to the reference for every
 * variant. */
int stencil_heat_step_f32(const float* t, float* t_next, float alpha, const stencil_grid* grid,
                          int variant, void* hip_stream);

/* The STENCIL_FP64 heat step (naive kernel for every variant). */
int stencil_heat_step_f64(const double* t, double* t_next, double alpha, const stencil_grid* grid,
                          int variant, void* hip_stream);

/* Algorithmic HBM bytes one launch moves (reference types.hpp:209-212:
 * 4 fields x interior x sizeof(real_t) for the wave step; 2 fields for heat). */
double stencil_wave_bytes_per_step(const stencil_grid* grid, int bytes_per_value);
double stencil_heat_bytes_per_step(const stencil_grid* grid, int bytes_per_value);

/* ---- Extensions (no counterpart in the reference) ---- */

/* Launch shape of the tiled kernel (autotuner / measurement).  Zero or -1
 * fields are chosen by the library.  (tile_x4, tile_y) must be one of the
 * built shapes: (16,16) (16,32) (32,8) (32,16) (64,4) (64,8) (64,12)
 * (64,16) -- a tile is 4*tile_x4 cells wide and tile_y rows tall, one
 * thread per 4 cells. */
typedef struct stencil_tile_shape {
    int tile_x4;     /* float4 columns per tile row (0 = default) */
    int tile_y;      /* rows per tile (0 = default) */
    int n_chunks;    /* z-chunks per column (0 = automatic) */
    int xcd_remap;   /* 1 on, 0 off, -1 automatic */
    int nontemporal; /* 1 non-temporal p_prev/vel2/p_next, 0 cached, -1 automatic */
} stencil_tile_shape;

/* The tiled step with an explicit launch shape (NULL = defaults).  exact != 0
 * selects the reference association (the default of stencil_launch_tiled_f32);
 * exact == 0 an FMA form with combined weights c_r/h^2 (fewer VALU ops; its
 * rounding drifts from the reference by ~1e-5 relative L2 over 50-200 steps,
 * so it is opt-in).  Returns STENCIL_ERR_UNSUPPORTED when nx % 4 != 0 or a
 * base is not 16-byte aligned (stencil_launch_tiled_f32 then runs the naive
 * kernel instead). */
int stencil_launch_tiled_ex_f32(const float* p_cur, const float* p_prev, const float* vel2,
                                float* p_next, const stencil_grid* grid, int exact,
                                const stencil_tile_shape* tile, void* hip_stream);
int stencil_heat_step_ex_f32(const float* t, float* t_next, float alpha, const stencil_grid* grid,
                             int exact, const stencil_tile_shape* tile, void* hip_stream);

/* Autotuner (replaces the reference's hand-run tuning_results/ sweeps): times
 * every built tile shape x z-chunk count x streaming mode of the tiled kernel
 * (reps launches each, hipEvents on hip_stream) on the caller's buffers --
 * p_next / t_next are overwritten -- and caches the fastest for this grid
 * size; later tiled launches with a NULL shape on the same dims use it.
 * Writes the sweep as CSV to csv_path when it is not NULL.  Synchronises. */
int stencil_tune_tiled_f32(const float* p_cur, const float* p_prev, const float* vel2,
                           float* p_next, const stencil_grid* grid, int reps,
                           stencil_tile_shape* best, const char* csv_path, void* hip_stream);
int stencil_tune_heat_f32(const float* t, float* t_next, const stencil_grid* grid, int reps,
                          stencil_tile_shape* best, const char* csv_path, void* hip_stream);
/* The shape a NULL-shape launch would use for these dims (kind 0 wave, 1
 * heat, 2 + storage for the 16-bit wave step, 6 FWI forward step, 7 adjoint
 * step, 8 forward step recomputed beside the adjoint, 9 the backward pass's
 * step pair: a recompute and an adjoint step in one launch, built for tiles
 * of 128 x 8, 128 x 16, 256 x 4 and 256 x 8 cells): returns 1 if it came
 * from the tuner cache, 0 for the default. */
int stencil_tuned_shape(const stencil_grid* grid, int kind, stencil_tile_shape* out);
/* Puts a shape into the tuner cache for these dims and kind (as above; kind 2
 * is unused), e.g. one saved from an earlier run's tuning; zero n_chunks and
 * -1 flags stay automatic.  STENCIL_ERR_INVALID for a shape the library is
 * not built with. */
int stencil_set_tuned_shape(const stencil_grid* grid, int kind, const stencil_tile_shape* shape);
void stencil_clear_tuning(void);

/* HBM calibration (the reference's tests/perf/bench_stream.cu:53-101):
 * op 0 = copy src -> dst, 1 = read src (dst receives at most 2048 partial
 * sums), 2 = write dst := 1, 3 = dst = src[0:n] + src[n:2n] + src[2n:3n]
 * (the stencil's 3-read / 1-write mix; src holds 3n floats).  n floats per
 * array, n % 4 == 0, 16-byte aligned. */
int stencil_stream_f32(int op, float* dst, const float* src, uint64_t n, void* hip_stream);

/* ---- 16-bit field storage: p, p_prev and vel2 stored in 16 bits, arithmetic
 * in FP32.  Decoding is exact; with exact = 1 the step uses the reference
 * association and is bit-reproducible against a CPU restatement, with exact = 0
 * it uses the fused-weight form (fewer operations per point; the FP32
 * arithmetic differs by rounding only, far below the 16-bit storage error).
 *   STENCIL_STORE_BF16  bfloat16, round to nearest even
 *   STENCIL_STORE_FP16  IEEE half of value*scale, scale a power of two per field
 *   STENCIL_STORE_BFP16 block floating point: int16 mantissas + one int8
 *                       exponent per 16 x-consecutive values (v = m*2^(e-15),
 *                       e = frexp exponent of the block max, clamped to
 *                       [-112, 127]); needs the
 *                       zero halo of Dirichlet walls, and nx % 16 == 0. ---- */
enum stencil_storage {
    STENCIL_STORE_BF16 = 1,
    STENCIL_STORE_FP16 = 2,
    STENCIL_STORE_BFP16 = 3
};
typedef struct stencil_field16 {
    void* data;  /* nx*ny*nz uint16 */
    int8_t* exp; /* nx*ny*nz/16 block exponents (BFP16 only, else NULL) */
} stencil_field16;

/* Bytes of the data and exponent arrays of one field of this grid. */
int stencil_field16_bytes(int storage, const stencil_grid* grid, uint64_t* data_bytes,
                          uint64_t* exp_bytes);
/* n values (n % 16 == 0) between FP32 and 16-bit storage; scale is the
 * power-of-two field scale (FP16 only, 1 otherwise). */
int stencil_encode16(int storage, const float* src, const stencil_field16* dst, uint64_t n,
                     float scale, void* hip_stream);
int stencil_decode16(int storage, const stencil_field16* src, float* dst, uint64_t n, float scale,
                     void* hip_stream);
/* One acoustic step on 16-bit fields with the tiled kernel.  p_scale is the
 * scale of p_cur / p_prev / p_next, v_scale of vel2 (both 1 unless FP16). */
int stencil_launch_tiled_16(int storage, const stencil_field16* p_cur, const stencil_field16* p_prev,
                            const stencil_field16* vel2, const stencil_field16* p_next,
                            const stencil_grid* grid, float p_scale, float v_scale, int exact,
                            const stencil_tile_shape* tile, void* hip_stream);
/* Autotuner of the 16-bit step (fused form), as stencil_tune_tiled_f32: the
 * best shape is cached for NULL-shape stencil_launch_tiled_16 calls on this
 * grid and storage (the sweep CSV has naive_ms = 0). */
int stencil_tune_tiled_16(int storage, const stencil_field16* p_cur, const stencil_field16* p_prev,
                          const stencil_field16* vel2, const stencil_field16* p_next,
                          const stencil_grid* grid, float p_scale, float v_scale, int reps,
                          stencil_tile_shape* best, const char* csv_path, void* hip_stream);

/* ---- FP64 reductions over the interior (deterministic: fixed-order two-pass
 * reduction through `work`, stencil_reduce_workspace_bytes() of device
 * memory; results land in out[0..2] on the device). ---- */
uint64_t stencil_reduce_workspace_bytes(void);
/* Leapfrog energy of the pair (p^n, p^{n+1}): out[0] = sum (p^{n+1}-p^n)^2 /
 * (vel2 dt^2), out[1] = -sum p^{n+1} * Lap(p^n); out[0] + out[1] is conserved
 * by the exact scheme. */
int stencil_energy_f32(const float* p_n, const float* p_np1, const float* vel2,
                       const stencil_grid* grid, double* out, double* work, void* hip_stream);
int stencil_energy_f64(const double* p_n, const double* p_np1, const double* vel2,
                       const stencil_grid* grid, double* out, double* work, void* hip_stream);
/* No counterpart in the reference.  stencil_energy_f32 for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * or time-step loops, harnesses that forbid hipMalloc): the same kernels and
 * results, with the partial sums in a static per-device array of the library
 * (24 KiB: 3 x 1024 doubles, the same for every grid) instead of `work`.  The
 * array's address is resolved on the first call on a device, with that device
 * current (call with the stream's device current).  Allocates nothing;
 * asynchronous on hip_stream like the other reductions.  Calls that overlap on
 * one device (other streams, other host threads) share the array and race:
 * order them.  Grids the reductions cannot take (an axis of <= 8 or >= 2^20
 * cells) return STENCIL_ERR_INVALID. */
int stencil_energy_f32_noalloc(const float* p_n, const float* p_np1, const float* vel2,
                               const stencil_grid* grid, double* out, void* hip_stream);
/* out[0] = sum (a-b)^2, out[1] = sum b^2 (relative L2 = sqrt(out0/out1)). */
int stencil_diff_norms_f32_f64(const float* a, const double* b, const stencil_grid* grid,
                               double* out, double* work, void* hip_stream);
/* out[0] = sum a*b. */
int stencil_dot_f64(const double* a, const double* b, const stencil_grid* grid, double* out,
                    double* work, void* hip_stream);
int stencil_dot_f32(const float* a, const float* b, const stencil_grid* grid, double* out,
                    double* work, void* hip_stream);

/* Seeded host-side input generators (OpenMP), the same values on every
 * platform: see stencil::init_gaussian / init_layered_noise_vel2 /
 * init_noise_cube in include/stencil/grid.hpp. */
int stencil_host_init_gaussian_f32(float* host, uint64_t nx, uint64_t ny, uint64_t nz, float cx,
                                   float cy, float cz, float sigma);
int stencil_host_layered_noise_vel2_f32(float* host, uint64_t nx, uint64_t ny, uint64_t nz,
                                        int n_layers, float vmin, float vmax, float noise,
                                        uint64_t seed);
int stencil_host_noise_cube_f32(float* host, uint64_t nx, uint64_t ny, uint64_t nz,
                                uint64_t side, float hot, uint64_t seed);

/* ---- Adjoint-state FWI gradient (no reference counterpart: the reference
 * models forward only).  Scheme, with m = vel2, L the stencil Laplacian and a
 * point source of amplitude wavelet[n] added to p^{n+1}:
 *   p^{n+1} = 2p^n - p^{n-1} + dt^2 m L p^n + wavelet[n] e_src,  p^0 = p^-1 = 0
 *   J = 1/2 sum_{k=1..N} |p^k(plane rec_z) - d_obs^k|^2
 * The gradient is the exact discrete adjoint: dJ/dm = sum_k dt^2 (L p^k)
 * lambda^{k+1}, lambda^k = 2lambda^{k+1} - lambda^{k+2} + dt^2 L(m lambda^{k+1})
 * + residual^k on plane rec_z, computed in psi = m lambda (psi^k = 2psi^{k+1} -
 * psi^{k+2} + dt^2 m L psi^{k+1} + m residual^k; dJ/dm = sum_k dt^2 (L p^k)
 * psi^{k+1} / m).  The forward pass keeps the level pair at
 * every ckpt_every-th step; the backward pass recomputes each segment from
 * it.  With ckpt_every = n_steps (one segment: n_steps fields of workspace,
 * e.g. 134 GB for 256^3 x 2000 FP32) the misfit forward stores every L p^k
 * and the backward runs the adjoint steps only, with the same results.
 * With more than one segment the backward pass recomputes segment s-1 on a
 * second stream of its own (created once per device, default priority),
 * overlapped with the adjoint steps of segment s on hip_stream and joined to
 * it by events: the call stays asynchronous and its work ends on hip_stream.
 * With STENCIL_FWI_OVERLAP=0 in the environment everything runs on
 * hip_stream, recompute then adjoint steps per segment (same results).
 * All device memory is the caller's workspace.  Arrays: vel2 and grad
 * nx*ny*nz, wavelet n_steps (device), traces / d_obs n_steps*ny*nx with trace
 * n holding p^{n+1} on plane rec_z. ---- */
typedef struct stencil_fwi_config {
    uint64_t n_steps;
    uint64_t src_x, src_y, src_z; /* source cell (interior) */
    uint64_t rec_z;               /* receiver plane (interior z), all nx*ny cells */
    uint64_t ckpt_every;          /* segment length; 0 = ceil(sqrt(n_steps)); n_steps = no recompute */
    uint64_t variant;             /* stencil_variant of the FP32 steps (FP64: naive kernels) */
} stencil_fwi_config;
/* Workspace bytes (256-byte aligned base) for this grid, config and value
 * size (4 or 8); 0 if the arguments are invalid. */
uint64_t stencil_fwi_workspace_bytes(const stencil_grid* grid, const stencil_fwi_config* cfg,
                                     int bytes_per_value);
/* Forward modelling with checkpoints.  traces_out (nullable) receives the
 * synthetic traces; with d_obs (nullable) the residual is kept in the
 * workspace and J is written to misfit_out (device double, nullable). */
int stencil_fwi_forward_f32(const float* vel2, const stencil_grid* grid, const stencil_fwi_config* cfg,
                            const float* wavelet, const float* d_obs, float* traces_out, double* misfit_out,
                            void* work, void* hip_stream);
int stencil_fwi_forward_f64(const double* vel2, const stencil_grid* grid, const stencil_fwi_config* cfg,
                            const double* wavelet, const double* d_obs, double* traces_out, double* misfit_out,
                            void* work, void* hip_stream);
/* Adjoint sweep after a forward call with d_obs on the same workspace: grad
 * (nx*ny*nz) = dJ/dvel2; src_trace (nullable, n_steps) = lambda^{n+1} at the
 * source cell, the adjoint of the source injection (dot-product test). */
int stencil_fwi_backward_f32(const float* vel2, const stencil_grid* grid, const stencil_fwi_config* cfg,
                             const float* wavelet, float* grad, float* src_trace, void* work,
                             void* hip_stream);
int stencil_fwi_backward_f64(const double* vel2, const stencil_grid* grid, const stencil_fwi_config* cfg,
                             const double* wavelet, double* grad, double* src_trace, void* work,
                             void* hip_stream);
/* Autotuner of the FP32 FWI kernels (no reference counterpart): the sweep of
 * stencil_tune_tiled_f32 (both streaming modes) applied to the forward
 * step as the forward pass runs it (recording the receiver plane), to the
 * adjoint step, and to the forward step writing L p while the chosen adjoint
 * step runs beside it on the second stream, as the backward pass recomputes
 * (with one segment, or ckpt_every = 1, there is no such overlap and
 * best_recompute repeats best_forward).  Each timed chain rotates three
 * buffers like the time loop.  Runs on the workspace's wavefield, adjoint
 * and Laplacian buffers; the checkpoints, traces and stored L p of an
 * earlier forward pass stay intact.  The shapes are cached for this grid
 * (stencil_tuned_shape kinds 6, 7, 8) and used by the tiled FWI calls on
 * it.  CSV rows: kernel, grid, shape, ms per step.  Synchronises. */
int stencil_tune_fwi_f32(const stencil_grid* grid, const stencil_fwi_config* cfg, void* work, int reps,
                         stencil_tile_shape* best_forward, stencil_tile_shape* best_adjoint,
                         stencil_tile_shape* best_recompute, const char* csv_path, void* hip_stream);
/* The single steps: the acoustic step (reference association) plus
 * wavelet[n] at linear cell src_index (< 0: none), writing L p_cur into
 * lap_out (nullable); and the adjoint step in the variable psi = vel2 *
 * lambda (lambda the adjoint field): psi_out = 2 psi1 - psi2 +
 * dt^2 vel2 L(psi1) (+ vel2 * res_plane on plane rec_z when non-null), grad +=
 * dt^2 lap psi1 when lap is non-null (dJ/dvel2 is the sum divided by vel2),
 * *src_value = psi_out[src_index] / vel2[src_index] (= lambda there) when
 * src_value is non-null. */
int stencil_forward_step_f32(const float* p_cur, const float* p_prev, const float* vel2, float* p_next,
                             float* lap_out, const stencil_grid* grid, int64_t src_index,
                             const float* wavelet, int64_t n, int variant, void* hip_stream);
int stencil_forward_step_f64(const double* p_cur, const double* p_prev, const double* vel2,
                             double* p_next, double* lap_out, const stencil_grid* grid, int64_t src_index,
                             const double* wavelet, int64_t n, int variant, void* hip_stream);
int stencil_adjoint_step_f32(const float* psi1, const float* psi2, const float* vel2, const float* lap,
                             float* grad, float* psi_out, const float* res_plane, int64_t rec_z,
                             const stencil_grid* grid, int64_t src_index, float* src_value, int variant,
                             void* hip_stream);
int stencil_adjoint_step_f64(const double* psi1, const double* psi2, const double* vel2, const double* lap,
                             double* grad, double* psi_out, const double* res_plane, int64_t rec_z,
                             const stencil_grid* grid, int64_t src_index, double* src_value, int variant,
                             void* hip_stream);

/* ---- Batched multi-shot forward modelling (no reference counterpart).
 * Each shot is the FWI forward scheme on its own grid: three caller-owned
 * time-level buffers, vel2, and an optional point source (wavelet of n_steps
 * device samples; NULL = none).  Every time step of every shot, whatever its
 * grid, is ONE launch of the tiled kernel over a table of work items (shot,
 * tile, z-chunk) built at create time: one tile shape for the batch, z-chunks
 * of about a target length, the items cut into eight equal-cost runs, one
 * per XCD.  With tune non-zero the batch autotuner times that launch on the
 * batch's buffers for every built tile shape, a few chunk lengths and with /
 * without the XCD runs, and keeps the fastest (else a default plan).  The
 * whole run -- zeroing all levels, then n_steps launches -- is captured into
 * one hipGraph at create time.  run() is one graph launch; afterwards
 * p^{n_steps} of every shot is in ring[stencil_batch_newest()].  Every run
 * starts again from zero fields.  Needs nx % 4 == 0 and 16-byte aligned
 * buffers.  With STENCIL_BATCH_MODE=grouped in the environment at create
 * time, shots of equal grids instead form groups with one launch per group
 * per step, each group's shape from the wave autotuner (same results). ---- */
typedef struct stencil_shot {
    stencil_grid grid;
    float* ring[3];
    const float* vel2;
    const float* wavelet;         /* n_steps samples or NULL */
    uint64_t src_x, src_y, src_z; /* interior cell (when wavelet is set) */
} stencil_shot;
typedef struct stencil_batch stencil_batch;
int stencil_batch_create(const stencil_shot* shots, int n_shots, int n_steps, int tune,
                         stencil_batch** out, void* hip_stream);
int stencil_batch_run(stencil_batch* batch, void* hip_stream);
/* Ring index holding the newest level (same for every shot). */
int stencil_batch_newest(const stencil_batch* batch, int shot);
/* Groups of equal grids, and per group its grid, launch shape and shot count
 * (ragged: the batch's tile shape, this grid's z-chunk count, xcd_remap = the
 * per-XCD runs). */
int stencil_batch_groups(const stencil_batch* batch);
int stencil_batch_group(const stencil_batch* batch, int group, stencil_grid* grid,
                        stencil_tile_shape* shape, int* n_shots);
/* Re-plans a (ragged) batch with an explicit plan and recaptures its graph:
 * the tile shape (tile_x4, tile_y), z-chunks of about chunk_planes planes,
 * xcd_remap 1 for the per-XCD runs (0: one list, longest items first);
 * n_chunks and nontemporal are ignored.  STENCIL_ERR_UNSUPPORTED for a
 * grouped batch.  Synchronises. */
int stencil_batch_set_plan(stencil_batch* batch, const stencil_tile_shape* shape, int chunk_planes,
                           void* hip_stream);
/* 1 for a ragged batch (one launch per step, or one persistent launch for
 * the run), 0 for a grouped one. */
int stencil_batch_is_ragged(const stencil_batch* batch);
/* Workgroups of the persistent launch (STENCIL_BATCH_MODE=persistent: one
 * launch of workgroups that are all resident together steps every shot
 * n_steps times, taking each step's work items from per-XCD queues and
 * meeting at a grid barrier between steps); 0 when the run is one launch per
 * step. */
int stencil_batch_persistent_grid(const stencil_batch* batch);
/* Waits for the stream, then STENCIL_OK, or STENCIL_ERR_TIMEOUT when a grid
 * barrier of the last persistent run gave up (its fields are then not
 * valid). */
int stencil_batch_check(const stencil_batch* batch, void* hip_stream);
void stencil_batch_destroy(stencil_batch* batch);

/* ---- Solver handles: ISolver (reference include/stencil/solver.hpp:13-24)
 * for callers that cannot use the C++ API. ---- */
typedef struct stencil_config {
    uint64_t nx, ny, nz;
    double dx, dy, dz;
    double dt; /* 0 = CFL-derived (Config::make_grid) */
    int steps, warmup_steps;
    double velocity, diffusivity;
    double pulse_cx, pulse_cy, pulse_cz, pulse_sigma;
    int physics; /* 0 wave, 1 heat */
    int variant; /* enum stencil_variant */
    int device;
    int vmodel; /* 0 constant, 1 layered-noise */
    int n_layers;
    double vmin, vmax, vnoise;
    int ic; /* 0 gaussian, 1 noise-cube */
    uint64_t cube_side;
    double cube_value;
    uint64_t seed;
    int fp64; /* 0: float fields (real_t = float), 1: the STENCIL_FP64 solver */
} stencil_config;

typedef struct stencil_perf {
    double time_ms, bandwidth_gbs, gflops, model_bandwidth_gbs;
    uint64_t grid_size;
    char label[32];
} stencil_perf;

typedef struct stencil_solver stencil_solver;

/* Fills the reference Config defaults (config.hpp:34-70). */
void stencil_config_default(stencil_config* cfg);
int stencil_solver_create(const stencil_config* cfg, stencil_solver** out);
int stencil_solver_step(stencil_solver* s);
int stencil_solver_run(stencil_solver* s, int n_steps, int warmup, stencil_perf* out);
/* Copies the newest time level into host memory of nx*ny*nz values of the
 * solver's type (bytes must match) and synchronises. */
int stencil_solver_download(stencil_solver* s, void* host, uint64_t bytes);
int stencil_solver_dims(const stencil_solver* s, stencil_grid* out);
int stencil_solver_name(const stencil_solver* s, char* buf, uint64_t len);
void stencil_solver_destroy(stencil_solver* s);

#ifdef __cplusplus
}
#endif
#endif /* STENCIL_CAPI_H */
