/*
 * oracle/stencil_oracle.c -- CPU restatement of the reference stencil path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library; the product never
 * does.  Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 *
 * Each function restates the reference function cited in its comment, in the
 * same loop order and floating-point association, so that with contraction
 * off it is bit-identical to the reference CPU path built with
 * -ffp-contract=off (checked against oracle/_ref and the committed fixtures
 * in tests/golden/).
 */
#include "stencil_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define R ORACLE_R

/* include/stencil/types.hpp:175-181 -- static_cast<real_t> of the double
 * rationals. */
void oracle_fd_coeffs_f32(float c[5]) {
    c[0] = (float)(-205.0 / 72.0);
    c[1] = (float)(8.0 / 5.0);
    c[2] = (float)(-1.0 / 5.0);
    c[3] = (float)(8.0 / 315.0);
    c[4] = (float)(-1.0 / 560.0);
}

void oracle_fd_coeffs_f64(double c[5]) {
    c[0] = -205.0 / 72.0;
    c[1] = 8.0 / 5.0;
    c[2] = -1.0 / 5.0;
    c[3] = 8.0 / 315.0;
    c[4] = -1.0 / 560.0;
}

/* src/kernels/stencil_naive.cu:136-170 (CPU launch_naive).  The scalars are
 * formed exactly as at :145-148: inv = 1/(h*h), dt2 = dt*dt, in real_t. */
#define ORACLE_WAVE_STEP(T, NAME, COEFFS)                                                     \
    void NAME(const T* p_cur, const T* p_prev, const T* vel2, T* p_next, size_t nx, size_t ny, \
              size_t nz, T dx, T dy, T dz, T dt) {                                            \
        T c_fd[5];                                                                            \
        COEFFS(c_fd);                                                                         \
        const int NX = (int)nx, NY = (int)ny, NZ = (int)nz;                                   \
        const T inv_dx2 = (T)1 / (dx * dx);                                                   \
        const T inv_dy2 = (T)1 / (dy * dy);                                                   \
        const T inv_dz2 = (T)1 / (dz * dz);                                                   \
        const T dt2 = dt * dt;                                                                \
        const ptrdiff_t sy = NX;                                                              \
        const ptrdiff_t sz = (ptrdiff_t)NX * NY;                                              \
        _Pragma("omp parallel for collapse(3) schedule(static)")                              \
        for (int iz = R; iz < NZ - R; ++iz)                                                   \
            for (int iy = R; iy < NY - R; ++iy)                                               \
                for (int ix = R; ix < NX - R; ++ix) {                                         \
                    const ptrdiff_t c = iz * sz + iy * sy + ix;                               \
                    T lap_x = c_fd[0] * p_cur[c] * inv_dx2;                                   \
                    T lap_y = c_fd[0] * p_cur[c] * inv_dy2;                                   \
                    T lap_z = c_fd[0] * p_cur[c] * inv_dz2;                                   \
                    for (int r = 1; r <= R; ++r) {                                            \
                        lap_x += c_fd[r] * (p_cur[c + r] + p_cur[c - r]) * inv_dx2;           \
                        lap_y += c_fd[r] * (p_cur[c + r * sy] + p_cur[c - r * sy]) * inv_dy2; \
                        lap_z += c_fd[r] * (p_cur[c + r * sz] + p_cur[c - r * sz]) * inv_dz2; \
                    }                                                                         \
                    p_next[c] = (T)2 * p_cur[c] - p_prev[c] +                                 \
                                dt2 * vel2[c] * (lap_x + lap_y + lap_z);                      \
                }                                                                             \
    }

ORACLE_WAVE_STEP(float, oracle_wave_step_f32, oracle_fd_coeffs_f32)
ORACLE_WAVE_STEP(double, oracle_wave_step_f64, oracle_fd_coeffs_f64)

/* include/stencil/config.hpp:75-87 -- dt = CFL_LIMIT * min(dx,dy,dz) / v,
 * CFL_LIMIT = static_cast<real_t>(0.45) (types.hpp:139). */
float oracle_make_grid_dt_f32(float dx, float dy, float dz, float velocity) {
    float min_h = dx;
    if (dy < min_h) min_h = dy;
    if (dz < min_h) min_h = dz;
    const float cfl = (float)0.45;
    return cfl * min_h / velocity;
}

/* src/solver/heat_diffusion.cpp:86-87 -- heat_dims_.dt =
 * std::sqrt(static_cast<double>(dt)) narrowed to real_t. */
float oracle_heat_dt_eff_f32(float dt) { return (float)sqrt((double)dt); }

/* src/solver/heat_diffusion.cpp:103-122: the wave kernel with vel2 == alpha
 * everywhere (:79-81) and p_prev == p_cur (the copy at :116-118), using the
 * dims with dt_eff.  Written as the wave step with p_prev = t and a constant
 * vel2, which is the same arithmetic. */
void oracle_heat_step_f32(const float* t, float* t_next, size_t nx, size_t ny, size_t nz,
                          float dx, float dy, float dz, float dt_eff, float alpha) {
    float c_fd[5];
    oracle_fd_coeffs_f32(c_fd);
    const int NX = (int)nx, NY = (int)ny, NZ = (int)nz;
    const float inv_dx2 = 1.0f / (dx * dx);
    const float inv_dy2 = 1.0f / (dy * dy);
    const float inv_dz2 = 1.0f / (dz * dz);
    const float dt2 = dt_eff * dt_eff;
    const ptrdiff_t sy = NX;
    const ptrdiff_t sz = (ptrdiff_t)NX * NY;
#pragma omp parallel for collapse(3) schedule(static)
    for (int iz = R; iz < NZ - R; ++iz)
        for (int iy = R; iy < NY - R; ++iy)
            for (int ix = R; ix < NX - R; ++ix) {
                const ptrdiff_t c = iz * sz + iy * sy + ix;
                float lap_x = c_fd[0] * t[c] * inv_dx2;
                float lap_y = c_fd[0] * t[c] * inv_dy2;
                float lap_z = c_fd[0] * t[c] * inv_dz2;
                for (int r = 1; r <= R; ++r) {
                    lap_x += c_fd[r] * (t[c + r] + t[c - r]) * inv_dx2;
                    lap_y += c_fd[r] * (t[c + r * sy] + t[c - r * sy]) * inv_dy2;
                    lap_z += c_fd[r] * (t[c + r * sz] + t[c - r * sz]) * inv_dz2;
                }
                t_next[c] = 2.0f * t[c] - t[c] + dt2 * alpha * (lap_x + lap_y + lap_z);
            }
}

/* include/stencil/grid.hpp:57-78. */
void oracle_init_gaussian_f32(float* g, size_t nx, size_t ny, size_t nz, float cx, float cy,
                              float cz, float sigma) {
    const float icx = cx * (float)(nx - 1);
    const float icy = cy * (float)(ny - 1);
    const float icz = cz * (float)(nz - 1);
    const float inv2s2 = 1.0f / (2.0f * sigma * sigma);
    for (size_t iz = R; iz < nz - R; ++iz)
        for (size_t iy = R; iy < ny - R; ++iy)
            for (size_t ix = R; ix < nx - R; ++ix) {
                const float dx = (float)ix - icx;
                const float dy = (float)iy - icy;
                const float dz = (float)iz - icz;
                g[iz * nx * ny + iy * nx + ix] = expf(-(dx * dx + dy * dy + dz * dz) * inv2s2);
            }
}

void oracle_init_gaussian_f64(double* g, size_t nx, size_t ny, size_t nz, double cx, double cy,
                              double cz, double sigma) {
    const double icx = cx * (double)(nx - 1);
    const double icy = cy * (double)(ny - 1);
    const double icz = cz * (double)(nz - 1);
    const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
    for (size_t iz = R; iz < nz - R; ++iz)
        for (size_t iy = R; iy < ny - R; ++iy)
            for (size_t ix = R; ix < nx - R; ++ix) {
                const double dx = (double)ix - icx;
                const double dy = (double)iy - icy;
                const double dz = (double)iz - icz;
                g[iz * nx * ny + iy * nx + ix] = exp(-(dx * dx + dy * dy + dz * dz) * inv2s2);
            }
}

/* include/stencil/grid.hpp:83-85. */
void oracle_init_constant_velocity_f32(float* g, size_t n, float v) {
    const float v2 = v * v;
    for (size_t i = 0; i < n; ++i) g[i] = v2;
}

/* include/stencil/grid.hpp:91-99. */
void oracle_init_layered_velocity_f32(float* g, size_t nx, size_t ny, size_t nz, float v_top,
                                      float v_bot) {
    for (size_t iz = 0; iz < nz; ++iz)
        for (size_t iy = 0; iy < ny; ++iy)
            for (size_t ix = 0; ix < nx; ++ix) {
                const float v = (iz < nz / 2) ? v_top : v_bot;
                g[iz * nx * ny + iy * nx + ix] = v * v;
            }
}

/* src/utils/grid.cpp:86-97 -- RMS over every cell, halos included. */
double oracle_l2_error_f32(const float* a, const float* b, size_t n) {
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double d = (double)a[i] - (double)b[i];
        sum += d * d;
    }
    return sqrt(sum / (double)n);
}

/* src/utils/grid.cpp:102-115. */
double oracle_max_abs_error_f32(const float* a, const float* b, size_t n) {
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double d = fabs((double)a[i] - (double)b[i]);
        if (d > m) m = d;
    }
    return m;
}

/* src/solver/solver_base.cpp:28-62 (both levels start at the IC), :65-108
 * (every step is step_impl), src/solver/wave_prop.cpp:67-75 (launch, then
 * rotate), include/stencil/grid.hpp:144-153 (prev=buf[i], cur=buf[i+1],
 * next=buf[i+2]) and :174-176 (download returns cur).  All three levels are
 * zeroed first, which is what keeps the Dirichlet halo at zero
 * (grid.hpp:117,156-160). */
#define ORACLE_WAVE_RUN(T, NAME, STEP)                                                         \
    void NAME(const T* p0, const T* vel2, T* buf0, T* buf1, T* buf2, T* out, size_t nx,         \
              size_t ny, size_t nz, T dx, T dy, T dz, T dt, int n_steps) {                     \
        const size_t n = nx * ny * nz;                                                         \
        T* buf[3] = {buf0, buf1, buf2};                                                        \
        memset(buf2, 0, n * sizeof(T));                                                        \
        memcpy(buf0, p0, n * sizeof(T));                                                       \
        memcpy(buf1, p0, n * sizeof(T));                                                       \
        int idx = 0;                                                                           \
        for (int s = 0; s < n_steps; ++s) {                                                    \
            STEP(buf[(idx + 1) % 3], buf[idx], vel2, buf[(idx + 2) % 3], nx, ny, nz, dx, dy,   \
                 dz, dt);                                                                      \
            idx = (idx + 1) % 3;                                                               \
        }                                                                                      \
        memcpy(out, buf[(idx + 1) % 3], n * sizeof(T));                                        \
    }

ORACLE_WAVE_RUN(float, oracle_wave_run_f32, oracle_wave_step_f32)
ORACLE_WAVE_RUN(double, oracle_wave_run_f64, oracle_wave_step_f64)

/* src/solver/heat_diffusion.cpp:103-122: step into `next`, rotate, copy
 * cur -> prev.  Only cur is observable (download), so two buffers suffice. */
void oracle_heat_run_f32(const float* t0, float* buf0, float* buf1, float* out, size_t nx,
                         size_t ny, size_t nz, float dx, float dy, float dz, float dt_eff,
                         float alpha, int n_steps) {
    const size_t n = nx * ny * nz;
    float* a = buf0;
    float* b = buf1;
    memcpy(a, t0, n * sizeof(float));
    memset(b, 0, n * sizeof(float));
    for (int s = 0; s < n_steps; ++s) {
        oracle_heat_step_f32(a, b, nx, ny, nz, dx, dy, dz, dt_eff, alpha);
        float* t = a;
        a = b;
        b = t;
    }
    memcpy(out, a, n * sizeof(float));
}

/* ---- Seeded synthetic inputs (no reference counterpart). ---- */

/* SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014). */
uint64_t oracle_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* U[0,1) with 24 random bits, keyed by (seed, stream, idx). */
float oracle_u01(uint64_t seed, uint64_t stream, uint64_t idx) {
    const uint64_t k = oracle_splitmix64(seed ^ oracle_splitmix64(stream)) + idx;
    return (float)(oracle_splitmix64(k) >> 40) * (1.0f / 16777216.0f);
}

void oracle_layered_noise_vel2_f32(float* vel2, size_t nx, size_t ny, size_t nz, int n_layers,
                                   float vmin, float vmax, float noise, uint64_t seed) {
    for (size_t iz = 0; iz < nz; ++iz) {
        size_t layer = iz * (size_t)n_layers / nz;
        const float vl = vmin + (vmax - vmin) * oracle_u01(seed, 1, layer);
        for (size_t iy = 0; iy < ny; ++iy)
            for (size_t ix = 0; ix < nx; ++ix) {
                const size_t i = iz * nx * ny + iy * nx + ix;
                const float f = 1.0f + noise * (2.0f * oracle_u01(seed, 2, i) - 1.0f);
                const float v = vl * f;
                vel2[i] = v * v;
            }
    }
}

void oracle_heat_noise_cube_f32(float* t, size_t nx, size_t ny, size_t nz, size_t side,
                                float hot, uint64_t seed) {
    const size_t x0 = nx / 2 - side / 2, y0 = ny / 2 - side / 2, z0 = nz / 2 - side / 2;
    memset(t, 0, nx * ny * nz * sizeof(float));
    for (size_t iz = R; iz < nz - R; ++iz)
        for (size_t iy = R; iy < ny - R; ++iy)
            for (size_t ix = R; ix < nx - R; ++ix) {
                const size_t i = iz * nx * ny + iy * nx + ix;
                const int in_cube = ix >= x0 && ix < x0 + side && iy >= y0 && iy < y0 + side &&
                                    iz >= z0 && iz < z0 + side;
                t[i] = in_cube ? hot : oracle_u01(seed, 3, i);
            }
}

double oracle_rel_l2_interior_f32(const float* a, const float* b, size_t nx, size_t ny,
                                  size_t nz) {
    double num = 0.0, den = 0.0;
    for (size_t iz = R; iz < nz - R; ++iz)
        for (size_t iy = R; iy < ny - R; ++iy)
            for (size_t ix = R; ix < nx - R; ++ix) {
                const size_t i = iz * nx * ny + iy * nx + ix;
                const double d = (double)a[i] - (double)b[i];
                num += d * d;
                den += (double)b[i] * (double)b[i];
            }
    return den > 0.0 ? sqrt(num / den) : sqrt(num);
}

double oracle_rel_l2_interior_f64f32(const float* a, const double* b, size_t nx, size_t ny,
                                     size_t nz) {
    double num = 0.0, den = 0.0;
    for (size_t iz = R; iz < nz - R; ++iz)
        for (size_t iy = R; iy < ny - R; ++iy)
            for (size_t ix = R; ix < nx - R; ++ix) {
                const size_t i = iz * nx * ny + iy * nx + ix;
                const double d = (double)a[i] - b[i];
                num += d * d;
                den += b[i] * b[i];
            }
    return den > 0.0 ? sqrt(num / den) : sqrt(num);
}

/* ---- 16-bit field storage ---- */

static uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

uint16_t oracle_f32_to_bf16(float f) {
    const uint32_t b = f2u(f);
    return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

float oracle_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

/* IEEE 754 binary32 -> binary16, round to nearest even, subnormals kept,
 * overflow to infinity (IEEE 754-2008 section 4.3.1 / 7.4). */
uint16_t oracle_f32_to_fp16(float f) {
    const uint32_t x = f2u(f);
    const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    const int32_t e = (int32_t)((x >> 23) & 0xFF);
    uint32_t m = x & 0x7FFFFFu;
    if (e == 0xFF) return sign | 0x7C00u | (m ? 0x200u : 0u); /* inf / nan */
    int32_t he = e - 127 + 15;
    if (he >= 0x1F) return sign | 0x7C00u; /* overflow */
    if (he <= 0) {
        /* subnormal half (or zero): value = m' * 2^-24 */
        if (he < -10) return sign; /* rounds to zero */
        m |= 0x800000u;            /* implicit bit */
        const int shift = 14 - he; /* 14..24 */
        uint32_t hm = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (hm & 1u))) hm += 1;
        return sign | (uint16_t)hm; /* a carry into the exponent field is correct */
    }
    uint32_t hm = m >> 13;
    const uint32_t rem = m & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) {
        hm += 1;
        if (hm == 0x400u) {
            hm = 0;
            he += 1;
            if (he >= 0x1F) return sign | 0x7C00u;
        }
    }
    return sign | (uint16_t)((uint32_t)he << 10) | (uint16_t)hm;
}

float oracle_fp16_to_f32(uint16_t h) {
    const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1F;
    const uint32_t m = h & 0x3FFu;
    if (e == 0) {
        const float v = ldexpf((float)m, -24);
        return sign ? -v : v;
    }
    if (e == 0x1F) return u2f(sign | 0x7F800000u | (m << 13));
    return u2f(sign | ((e - 15 + 127) << 23) | (m << 13));
}

/* Block scale k of a block whose max |v| is m: e = frexp exponent of
 * fl(m * (1 + 2^-15)), clamped to [-112, 127] (0 for m == 0), k = e - 15.
 * The 1 + 2^-15 factor moves a max within 2^-16 of 2^e up to exponent e+1, so
 * |v| * 2^-k < 32767.5 and the rounded mantissa never leaves [-32767, 32767]. */
static int bfp16_exponent(float m) {
    if (!(m > 0.0f)) return -15;
    int e;
    frexpf(m * (1.0f + 0x1p-15f), &e);
    e = e < -112 ? -112 : (e > 127 ? 127 : e);
    return e - 15;
}

static uint16_t bfp16_mantissa(float v, int k) {
    float q = rintf(ldexpf(v, -k));
    if (q > 32767.0f) q = 32767.0f; /* unreachable by the choice of k; kept as a guard */
    if (q < -32767.0f) q = -32767.0f;
    return (uint16_t)(int16_t)q;
}

static float bfp16_value(uint16_t m, int k) { return ldexpf((float)(int16_t)m, k); }

void oracle_encode16(int fmt, const float* src, uint16_t* dst, int8_t* exp, size_t n, float scale) {
    if (fmt == 3) {
        for (size_t b = 0; b < n / 16; ++b) {
            float m = 0.0f;
            for (int k = 0; k < 16; ++k) m = fmaxf(m, fabsf(src[16 * b + k]));
            const int kb = bfp16_exponent(m);
            exp[b] = (int8_t)kb;
            for (int k = 0; k < 16; ++k) dst[16 * b + k] = bfp16_mantissa(src[16 * b + k], kb);
        }
        return;
    }
    for (size_t i = 0; i < n; ++i)
        dst[i] = fmt == 1 ? oracle_f32_to_bf16(src[i]) : oracle_f32_to_fp16(src[i] * scale);
}

void oracle_decode16(int fmt, const uint16_t* src, const int8_t* exp, float* dst, size_t n,
                     float scale) {
    const float inv = 1.0f / scale;
    for (size_t i = 0; i < n; ++i) {
        if (fmt == 1) dst[i] = oracle_bf16_to_f32(src[i]);
        else if (fmt == 2) dst[i] = oracle_fp16_to_f32(src[i]) * inv;
        else dst[i] = bfp16_value(src[i], exp[i / 16]);
    }
}

void oracle_wave_step16(int fmt, const uint16_t* cur, const int8_t* cur_e, const uint16_t* prev,
                        const int8_t* prev_e, const uint16_t* vel2, const int8_t* vel2_e,
                        uint16_t* next, int8_t* next_e, size_t nx, size_t ny, size_t nz, float dx,
                        float dy, float dz, float dt, float p_scale, float v_scale) {
    const size_t n = nx * ny * nz;
    float* fc = (float*)malloc(n * sizeof(float));
    float* fp = (float*)malloc(n * sizeof(float));
    float* fv = (float*)malloc(n * sizeof(float));
    float* fn = (float*)calloc(n, sizeof(float));
    oracle_decode16(fmt, cur, cur_e, fc, n, p_scale);
    oracle_decode16(fmt, prev, prev_e, fp, n, p_scale);
    oracle_decode16(fmt, vel2, vel2_e, fv, n, v_scale);
    oracle_wave_step_f32(fc, fp, fv, fn, nx, ny, nz, dx, dy, dz, dt);
    for (size_t z = R; z < nz - R; ++z)
        for (size_t y = R; y < ny - R; ++y) {
            const size_t row = (z * ny + y) * nx;
            if (fmt == 3) {
                for (size_t b = 0; b < nx / 16; ++b) {
                    if (16 * b >= nx - R) continue; /* no interior cell in this block */
                    float m = 0.0f;
                    for (size_t x = 16 * b; x < 16 * b + 16; ++x)
                        if (x >= R && x < nx - R) m = fmaxf(m, fabsf(fn[row + x]));
                    const int e = bfp16_exponent(m);
                    next_e[(row >> 4) + b] = (int8_t)e;
                    for (size_t x = 16 * b; x < 16 * b + 16; ++x)
                        if (x >= R && x < nx - R) next[row + x] = bfp16_mantissa(fn[row + x], e);
                }
            } else {
                for (size_t x = R; x < nx - R; ++x)
                    next[row + x] = fmt == 1 ? oracle_f32_to_bf16(fn[row + x])
                                             : oracle_f32_to_fp16(fn[row + x] * p_scale);
            }
        }
    free(fc);
    free(fp);
    free(fv);
    free(fn);
}

/* ---- Adjoint-state FWI (no reference counterpart).  Restates the scheme of
 * the library's FWI path with every time level kept (no checkpointing), in the
 * same association as its kernels:
 *   forward  p^{n+1} = 2p^n - p^{n-1} + (dt2*m)*L p^n  (+ wavelet[n] at src)
 *   L p      = (lx + ly) + lz,  la = c0*p*inv + sum_r c_r*(p[+r]+p[-r])*inv
 *   adjoint  lambda^k = 2 lambda^{k+1} - lambda^{k+2} + dt2*L(m lambda^{k+1})
 *                       (+ res^k on plane rec_z), run in psi = m lambda:
 *            psi^k = 2 psi^{k+1} - psi^{k+2} + (dt2*m)*L psi^{k+1}
 *                    (+ m*res^k on plane rec_z)
 *   gradient gt += (dt2 * L p^k) * psi^{k+1}, k = N .. 1;  g = gt / m
 *   adjoint field at the source: lambda^k = psi^k / m
 * Pinned by tests/test_oracle_fwi.py: central finite differences of the misfit
 * (FP64), the adjoint dot-product identity, and the forward step against the
 * pinned oracle_wave_step. ---- */
#define ORACLE_FWI(T, SUF, COEFFS)                                                                       \
    static void fwi_lap_##SUF(const T* p, T* out, int NX, int NY, int NZ, T ix2, T iy2, T iz2) {       \
        T c[5];                                                                                         \
        COEFFS(c);                                                                                      \
        const ptrdiff_t sy = NX, sz = (ptrdiff_t)NX * NY;                                               \
        for (int iz = R; iz < NZ - R; ++iz)                                                             \
            for (int iy = R; iy < NY - R; ++iy)                                                         \
                for (int ix = R; ix < NX - R; ++ix) {                                                   \
                    const ptrdiff_t q = iz * sz + iy * sy + ix;                                         \
                    const T ctr = p[q];                                                                 \
                    T lx = c[0] * ctr * ix2, ly = c[0] * ctr * iy2, lz = c[0] * ctr * iz2;              \
                    for (int r = 1; r <= R; ++r) {                                                      \
                        const ptrdiff_t o[3] = {r, r * sy, r * sz};                                     \
                        T a[3], b[3];                                                                   \
                        for (int k = 0; k < 3; ++k) {                                                   \
                            a[k] = p[q + o[k]];                                                         \
                            b[k] = p[q - o[k]];                                                         \
                        }                                                                               \
                        lx += c[r] * (a[0] + b[0]) * ix2;                                               \
                        ly += c[r] * (a[1] + b[1]) * iy2;                                               \
                        lz += c[r] * (a[2] + b[2]) * iz2;                                               \
                    }                                                                                   \
                    out[q] = lx + ly + lz;                                                              \
                }                                                                                       \
    }                                                                                                   \
    /* levels: (n_steps+1) fields p^0..p^N (p^0 = 0). */                                               \
    static void fwi_levels_##SUF(const T* m, int NX, int NY, int NZ, T dx, T dy, T dz, T dt, int N,      \
                                 size_t src, const T* wavelet, T* levels, T* laps) {                    \
        const size_t n = (size_t)NX * NY * NZ;                                                          \
        const T ix2 = (T)1 / (dx * dx), iy2 = (T)1 / (dy * dy), iz2 = (T)1 / (dz * dz), dt2 = dt * dt; \
        T* zero = (T*)calloc(n, sizeof(T));                                                             \
        memset(levels, 0, n * sizeof(T));                                                               \
        for (int s = 0; s < N; ++s) {                                                                   \
            const T* p = levels + (size_t)s * n;                                                        \
            const T* pp = s > 0 ? levels + (size_t)(s - 1) * n : zero;                                  \
            T* pn = levels + (size_t)(s + 1) * n;                                                       \
            T* lap = laps + (size_t)s * n;                                                              \
            memset(pn, 0, n * sizeof(T));                                                               \
            fwi_lap_##SUF(p, lap, NX, NY, NZ, ix2, iy2, iz2);                                     \
            for (int iz = R; iz < NZ - R; ++iz)                                                         \
                for (int iy = R; iy < NY - R; ++iy)                                                     \
                    for (int ix = R; ix < NX - R; ++ix) {                                               \
                        const size_t q = ((size_t)iz * NY + iy) * NX + ix;                              \
                        T v = (T)2 * p[q] - pp[q] + dt2 * m[q] * lap[q];                                \
                        if (q == src) v += wavelet[s];                                                  \
                        pn[q] = v;                                                                      \
                    }                                                                                   \
        }                                                                                               \
        free(zero);                                                                                     \
    }                                                                                                   \
    void oracle_fwi_forward_##SUF(const T* m, size_t nx, size_t ny, size_t nz, T dx, T dy, T dz, T dt,  \
                                  int n_steps, size_t src, size_t rec_z, const T* wavelet, T* traces) { \
        const size_t n = nx * ny * nz, plane = nx * ny;                                                 \
        T* levels = (T*)malloc((size_t)(n_steps + 1) * n * sizeof(T));                                  \
        T* laps = (T*)malloc((size_t)(n_steps + 1) * n * sizeof(T));                                    \
        fwi_levels_##SUF(m, (int)nx, (int)ny, (int)nz, dx, dy, dz, dt, n_steps, src, wavelet, levels,   \
                         laps);                                                                         \
        for (int k = 1; k <= n_steps; ++k)                                                              \
            memcpy(traces + (size_t)(k - 1) * plane, levels + (size_t)k * n + rec_z * plane,            \
                   plane * sizeof(T));                                                                  \
        free(levels);                                                                                   \
        free(laps);                                                                                     \
    }                                                                                                   \
    double oracle_fwi_gradient_##SUF(const T* m, size_t nx, size_t ny, size_t nz, T dx, T dy, T dz,     \
                                     T dt, int n_steps, size_t src, size_t rec_z, const T* wavelet,     \
                                     const T* d_obs, T* grad, T* src_trace) {                           \
        const int NX = (int)nx, NY = (int)ny, NZ = (int)nz, N = n_steps;                                \
        const size_t n = nx * ny * nz, plane = nx * ny;                                                 \
        const T ix2 = (T)1 / (dx * dx), iy2 = (T)1 / (dy * dy), iz2 = (T)1 / (dz * dz), dt2 = dt * dt; \
        T* levels = (T*)malloc((size_t)(N + 1) * n * sizeof(T));                                        \
        T* laps = (T*)malloc((size_t)(N + 1) * n * sizeof(T));                                          \
        fwi_levels_##SUF(m, NX, NY, NZ, dx, dy, dz, dt, N, src, wavelet, levels, laps);                 \
        /* L p^N (the forward loop stops at p^N) */                                                     \
        fwi_lap_##SUF(levels + (size_t)N * n, laps + (size_t)N * n, NX, NY, NZ, ix2, iy2, iz2);   \
        T* res = (T*)malloc((size_t)N * plane * sizeof(T));                                             \
        double J = 0.0;                                                                                 \
        for (int k = 1; k <= N; ++k)                                                                    \
            for (size_t i = 0; i < plane; ++i) {                                                        \
                const T r = levels[(size_t)k * n + rec_z * plane + i] - d_obs[(size_t)(k - 1) * plane + i]; \
                res[(size_t)(k - 1) * plane + i] = r;                                                   \
                J += (double)r * (double)r;                                                             \
            }                                                                                           \
        J *= 0.5;                                                                                       \
        T* l1 = (T*)calloc(n, sizeof(T));                                                               \
        T* l2 = (T*)calloc(n, sizeof(T));                                                               \
        T* lo = (T*)calloc(n, sizeof(T));                                                               \
        T* lq = (T*)calloc(n, sizeof(T));                                                               \
        memset(grad, 0, n * sizeof(T));                                                                 \
        /* l1, l2, lo: psi^{k+1}, psi^{k+2}, psi^k */                                                   \
        for (int k = N; k >= 1; --k) {                                                                  \
            fwi_lap_##SUF(l1, lq, NX, NY, NZ, ix2, iy2, iz2);                                     \
            const T* lp = laps + (size_t)k * n;                                                         \
            for (int iz = R; iz < NZ - R; ++iz)                                                         \
                for (int iy = R; iy < NY - R; ++iy)                                                     \
                    for (int ix = R; ix < NX - R; ++ix) {                                               \
                        const size_t q = ((size_t)iz * NY + iy) * NX + ix;                              \
                        T v = (T)2 * l1[q] - l2[q] + dt2 * m[q] * lq[q];                                \
                        if ((size_t)iz == rec_z)                                                        \
                            v += m[q] * res[(size_t)(k - 1) * plane + (size_t)iy * nx + ix];            \
                        lo[q] = v;                                                                      \
                        grad[q] += dt2 * lp[q] * l1[q];                                                 \
                    }                                                                                   \
            if (src_trace) src_trace[k - 1] = lo[src] / m[src];                                         \
            T* t = l2;                                                                                  \
            l2 = l1;                                                                                    \
            l1 = lo;                                                                                    \
            lo = t;                                                                                     \
        }                                                                                               \
        for (int iz = R; iz < NZ - R; ++iz)                                                             \
            for (int iy = R; iy < NY - R; ++iy)                                                         \
                for (int ix = R; ix < NX - R; ++ix) {                                                   \
                    const size_t q = ((size_t)iz * NY + iy) * NX + ix;                                  \
                    grad[q] = grad[q] / m[q];                                                           \
                }                                                                                       \
        free(levels);                                                                                   \
        free(laps);                                                                                     \
        free(res);                                                                                      \
        free(l1);                                                                                       \
        free(l2);                                                                                       \
        free(lo);                                                                                       \
        free(lq);                                                                                       \
        return J;                                                                                       \
    }

ORACLE_FWI(float, f32, oracle_fd_coeffs_f32)
ORACLE_FWI(double, f64, oracle_fd_coeffs_f64)

/* One shot of the batched modelling: the FWI forward scheme from zero fields,
 * out = p^{n_steps} (the same levels as oracle_fwi_forward_f32). */
void oracle_shot_run_f32(const float* m, size_t nx, size_t ny, size_t nz, float dx, float dy, float dz,
                         float dt, int n_steps, size_t src, const float* wavelet, float* out) {
    const size_t n = nx * ny * nz;
    float* levels = (float*)malloc((size_t)(n_steps + 1) * n * sizeof(float));
    float* laps = (float*)malloc((size_t)(n_steps + 1) * n * sizeof(float));
    float* zero_w = NULL;
    if (!wavelet) {
        zero_w = (float*)calloc((size_t)n_steps + 1, sizeof(float));
        src = (size_t)-1;
    }
    fwi_levels_f32(m, (int)nx, (int)ny, (int)nz, dx, dy, dz, dt, n_steps, src, wavelet ? wavelet : zero_w,
                   levels, laps);
    memcpy(out, levels + (size_t)n_steps * n, n * sizeof(float));
    free(levels);
    free(laps);
    free(zero_w);
}
