/*
 * oracle/stencil_oracle.h -- CPU restatement of the reference stencil path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library checks results: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (stencil_solver_amd/, include/) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the somya2703/stencil_solver tree).  Layout everywhere is the
 * reference's: x fastest, idx = z*nx*ny + y*nx + x, dims include a radius-4
 * zero halo on every face (include/stencil/types.hpp:110-136).
 *
 * How each function is pinned (checked against something other than itself:
 * the reference's own outputs, closed forms, independent implementations) is
 * enforced by tests/test_oracle.py and tests/test_oracle_prec16.py.
 */
#ifndef STENCIL_ORACLE_H
#define STENCIL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_R 4

/* FDCoeffs<4> (include/stencil/types.hpp:175-181): c[0]=c0, c[r]=c_r. */
void oracle_fd_coeffs_f32(float c[5]);
void oracle_fd_coeffs_f64(double c[5]);

/* CPU launch_naive (src/kernels/stencil_naive.cu:136-170): one leapfrog
 * step over the interior, halo of p_next untouched.  Same loop, same
 * association; compiled with -ffp-contract=off. */
void oracle_wave_step_f32(const float* p_cur, const float* p_prev, const float* vel2,
                          float* p_next, size_t nx, size_t ny, size_t nz, float dx, float dy,
                          float dz, float dt);
/* The same under STENCIL_FP64 (include/stencil/types.hpp:33-37). */
void oracle_wave_step_f64(const double* p_cur, const double* p_prev, const double* vel2,
                          double* p_next, size_t nx, size_t ny, size_t nz, double dx, double dy,
                          double dz, double dt);

/* Config::make_grid dt (include/stencil/config.hpp:75-87). */
float oracle_make_grid_dt_f32(float dx, float dy, float dz, float velocity);

/* HeatDiffusionSolver (src/solver/heat_diffusion.cpp:73-122):
 *   dt_eff = (real_t)sqrt((double)dt)  (:86-87), the wave kernel run with
 *   vel2 := alpha and p_prev == p_cur (:103-121). */
float oracle_heat_dt_eff_f32(float dt);
void oracle_heat_step_f32(const float* t, float* t_next, size_t nx, size_t ny, size_t nz,
                          float dx, float dy, float dz, float dt_eff, float alpha);

/* init_gaussian (include/stencil/grid.hpp:57-78): interior only, expf in
 * real_t, centre at c*(n-1). */
void oracle_init_gaussian_f32(float* g, size_t nx, size_t ny, size_t nz, float cx, float cy,
                              float cz, float sigma);
void oracle_init_gaussian_f64(double* g, size_t nx, size_t ny, size_t nz, double cx, double cy,
                              double cz, double sigma);
/* init_constant_velocity / init_layered_velocity (grid.hpp:83-99). */
void oracle_init_constant_velocity_f32(float* g, size_t n, float v);
void oracle_init_layered_velocity_f32(float* g, size_t nx, size_t ny, size_t nz, float v_top,
                                      float v_bot);

/* l2_error (absolute RMS over ALL cells) and max_abs_error
 * (src/utils/grid.cpp:86-115). */
double oracle_l2_error_f32(const float* a, const float* b, size_t n);
double oracle_max_abs_error_f32(const float* a, const float* b, size_t n);

/* SolverBase::init + run/step + download for the wave physics
 * (src/solver/solver_base.cpp:28-116, src/solver/wave_prop.cpp:67-75):
 * p_prev = p_cur = p0, then n_steps of step+rotate; out = newest level.
 * vel2 is a full field.  Scratch: caller passes three field buffers. */
void oracle_wave_run_f32(const float* p0, const float* vel2, float* buf0, float* buf1,
                         float* buf2, float* out, size_t nx, size_t ny, size_t nz, float dx,
                         float dy, float dz, float dt, int n_steps);
void oracle_wave_run_f64(const double* p0, const double* vel2, double* buf0, double* buf1,
                         double* buf2, double* out, size_t nx, size_t ny, size_t nz, double dx,
                         double dy, double dz, double dt, int n_steps);
/* Heat physics (heat_diffusion.cpp:103-122): out = T after n_steps. */
void oracle_heat_run_f32(const float* t0, float* buf0, float* buf1, float* out, size_t nx,
                         size_t ny, size_t nz, float dx, float dy, float dz, float dt_eff,
                         float alpha, int n_steps);

/* ---- Seeded synthetic inputs (no reference counterpart: the reference has
 * only constant / two-layer models and a Gaussian pulse).  Counter-based so
 * every cell's value depends only on (seed, stream, index). ---- */
uint64_t oracle_splitmix64(uint64_t x);
float oracle_u01(uint64_t seed, uint64_t stream, uint64_t idx);
/* vel2 = v^2 with v = v_layer[z-layer]*(1 + noise*U[-1,1]) over ALL cells;
 * layers are equal z-bands of nz/n_layers planes, v_layer ~ U[vmin,vmax]. */
void oracle_layered_noise_vel2_f32(float* vel2, size_t nx, size_t ny, size_t nz, int n_layers,
                                   float vmin, float vmax, float noise, uint64_t seed);
/* Heat IC: interior U[0,1] plus a centred cube of side `side` set to `hot`. */
void oracle_heat_noise_cube_f32(float* t, size_t nx, size_t ny, size_t nz, size_t side,
                                float hot, uint64_t seed);

/* Relative L2 over interior cells, FP64 accumulation: ||a-b|| / ||b||. */
double oracle_rel_l2_interior_f32(const float* a, const float* b, size_t nx, size_t ny,
                                  size_t nz);
double oracle_rel_l2_interior_f64f32(const float* a, const double* b, size_t nx, size_t ny,
                                     size_t nz);

/* ---- 16-bit field storage (no reference counterpart: the reference only
 * has the compile-time FP64 switch).  Formats: 1 bfloat16 (round to nearest
 * even), 2 IEEE half of value*scale (scale a power of two), 3 block floating
 * point (int16 mantissa m per value, one int8 block scale k per 16
 * x-consecutive values, v = m * 2^k; k = e - 15 with e the frexp exponent of
 * fl(max|v| * (1 + 2^-15)) clamped to [-112, 127] (e = 0 for an all-zero
 * block), so 2^-k is a normal float and |m| <= 32767 without clamping). ---- */
uint16_t oracle_f32_to_bf16(float f);
float oracle_bf16_to_f32(uint16_t h);
uint16_t oracle_f32_to_fp16(float f);
float oracle_fp16_to_f32(uint16_t h);
void oracle_encode16(int fmt, const float* src, uint16_t* dst, int8_t* exp, size_t n, float scale);
void oracle_decode16(int fmt, const uint16_t* src, const int8_t* exp, float* dst, size_t n,
                     float scale);
/* One acoustic step on 16-bit fields: decode p_cur, p_prev, vel2 (exact), the
 * FP32 step in the reference association (oracle_wave_step_f32), then encode
 * the interior of p_next; halo cells and the exponents of blocks with no
 * interior cell are left as they are.  Needs nx % 16 == 0. */
void oracle_wave_step16(int fmt, const uint16_t* cur, const int8_t* cur_e, const uint16_t* prev,
                        const int8_t* prev_e, const uint16_t* vel2, const int8_t* vel2_e,
                        uint16_t* next, int8_t* next_e, size_t nx, size_t ny, size_t nz, float dx,
                        float dy, float dz, float dt, float p_scale, float v_scale);

/* ---- Adjoint-state FWI (no reference counterpart): forward modelling with a
 * point source (wavelet[n] added to p^{n+1} at linear cell src) and the
 * receiver plane rec_z recorded as traces[n] = p^{n+1}(rec_z, :, :); the
 * misfit J = 1/2 sum |traces - d_obs|^2 (returned, FP64 accumulation) and its
 * exact discrete gradient dJ/dm (m = vel2) by the adjoint recursion with
 * every level kept.  src_trace (nullable) = lambda^{n+1} at src.  Tiny
 * grids only: (n_steps+1) fields of levels and Laplacians are held. ---- */
void oracle_fwi_forward_f32(const float* m, size_t nx, size_t ny, size_t nz, float dx, float dy,
                            float dz, float dt, int n_steps, size_t src, size_t rec_z,
                            const float* wavelet, float* traces);
void oracle_fwi_forward_f64(const double* m, size_t nx, size_t ny, size_t nz, double dx, double dy,
                            double dz, double dt, int n_steps, size_t src, size_t rec_z,
                            const double* wavelet, double* traces);
double oracle_fwi_gradient_f32(const float* m, size_t nx, size_t ny, size_t nz, float dx, float dy,
                               float dz, float dt, int n_steps, size_t src, size_t rec_z,
                               const float* wavelet, const float* d_obs, float* grad,
                               float* src_trace);
double oracle_fwi_gradient_f64(const double* m, size_t nx, size_t ny, size_t nz, double dx,
                               double dy, double dz, double dt, int n_steps, size_t src,
                               size_t rec_z, const double* wavelet, const double* d_obs,
                               double* grad, double* src_trace);
/* One shot of the batched modelling: p^{n_steps} of the scheme above from
 * zero fields (wavelet NULL: no source). */
void oracle_shot_run_f32(const float* m, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                         float dt, int n_steps, size_t src, const float* wavelet, float* out);

#ifdef __cplusplus
}
#endif
#endif
