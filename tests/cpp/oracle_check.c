/* tests/cpp/oracle_check.c -- drives every function of the oracle
 * (oracle/stencil_oracle.c) on small ragged grids, each buffer allocated at
 * exactly its size, for the AddressSanitizer / UBSan build
 * (tests/cpp/Makefile, target `sanitize`).  Beyond running clean it checks
 * a few facts each function guarantees: finite results, untouched halos,
 * round trips of the 16-bit formats, a zero gradient on a zero residual. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "stencil_oracle.h"

static int failures = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        if (!(cond)) {                                                     \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                    \
        }                                                                  \
    } while (0)

static void* xalloc(size_t bytes) {
    void* p = malloc(bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

static int all_finite_f(const float* a, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(a[i])) return 0;
    return 1;
}

static int halo_equals_f(const float* a, size_t nx, size_t ny, size_t nz, float v) {
    const size_t R = ORACLE_R;
    for (size_t z = 0; z < nz; ++z)
        for (size_t y = 0; y < ny; ++y)
            for (size_t x = 0; x < nx; ++x) {
                const int in = x >= R && x < nx - R && y >= R && y < ny - R && z >= R && z < nz - R;
                if (!in && a[(z * ny + y) * nx + x] != v) return 0;
            }
    return 1;
}

static void check_wave_and_heat(size_t nx, size_t ny, size_t nz) {
    const size_t n = nx * ny * nz;
    float* p = xalloc(n * 4);
    float* pp = xalloc(n * 4);
    float* v = xalloc(n * 4);
    float* nx_ = xalloc(n * 4);
    float* out = xalloc(n * 4);
    float* b[3] = {xalloc(n * 4), xalloc(n * 4), xalloc(n * 4)};
    oracle_init_gaussian_f32(p, nx, ny, nz, 0.5f, 0.4f, 0.6f, 2.0f);
    memcpy(pp, p, n * 4);
    oracle_layered_noise_vel2_f32(v, nx, ny, nz, 3, 1500.f, 3000.f, 0.05f, 9);
    for (size_t i = 0; i < n; ++i) nx_[i] = 7.5f;
    oracle_wave_step_f32(p, pp, v, nx_, nx, ny, nz, 10.f, 12.f, 8.f, 1e-3f);
    CHECK(all_finite_f(nx_, n));
    CHECK(halo_equals_f(nx_, nx, ny, nz, 7.5f));
    oracle_wave_run_f32(p, v, b[0], b[1], b[2], out, nx, ny, nz, 10.f, 10.f, 10.f, 1e-3f, 5);
    CHECK(all_finite_f(out, n));
    for (size_t i = 0; i < n; ++i) nx_[i] = -1.f;
    oracle_heat_step_f32(p, nx_, nx, ny, nz, 1.f, 1.f, 1.f, 0.05f, 1.f);
    CHECK(halo_equals_f(nx_, nx, ny, nz, -1.f));
    oracle_heat_run_f32(p, b[0], b[1], out, nx, ny, nz, 1.f, 1.f, 1.f, 0.05f, 1.f, 4);
    CHECK(all_finite_f(out, n));
    oracle_heat_noise_cube_f32(out, nx, ny, nz, 2, 10.f, 3);
    CHECK(halo_equals_f(out, nx, ny, nz, 0.f));
    oracle_init_constant_velocity_f32(v, n, 2000.f);
    oracle_init_layered_velocity_f32(v, nx, ny, nz, 1500.f, 2500.f);
    CHECK(v[0] == 1500.f * 1500.f && v[n - 1] == 2500.f * 2500.f);
    CHECK(oracle_l2_error_f32(p, p, n) == 0.0 && oracle_max_abs_error_f32(p, p, n) == 0.0);
    CHECK(oracle_rel_l2_interior_f32(p, p, nx, ny, nz) == 0.0);
    double* p64 = xalloc(n * 8);
    double* v64 = xalloc(n * 8);
    double* c64[4] = {xalloc(n * 8), xalloc(n * 8), xalloc(n * 8), xalloc(n * 8)};
    oracle_init_gaussian_f64(p64, nx, ny, nz, 0.5, 0.5, 0.5, 2.0);
    for (size_t i = 0; i < n; ++i) v64[i] = 2.25e6;
    oracle_wave_step_f64(p64, p64, v64, c64[3], nx, ny, nz, 10.0, 10.0, 10.0, 1e-3);
    oracle_wave_run_f64(p64, v64, c64[0], c64[1], c64[2], c64[3], nx, ny, nz, 10.0, 10.0, 10.0, 1e-3, 3);
    CHECK(oracle_rel_l2_interior_f64f32(out, c64[3], nx, ny, nz) >= 0.0);
    free(p64);
    free(v64);
    for (int k = 0; k < 4; ++k) free(c64[k]);
    free(p);
    free(pp);
    free(v);
    free(nx_);
    free(out);
    for (int k = 0; k < 3; ++k) free(b[k]);
}

static void check_16bit(size_t nx, size_t ny, size_t nz) {
    const size_t n = nx * ny * nz;  /* nx % 16 == 0 */
    float* src = xalloc(n * 4);
    float* back = xalloc(n * 4);
    float* v = xalloc(n * 4);
    oracle_init_gaussian_f32(src, nx, ny, nz, 0.5f, 0.5f, 0.5f, 2.0f);
    oracle_init_constant_velocity_f32(v, n, 1500.f);
    for (int fmt = 1; fmt <= 3; ++fmt) {
        uint16_t* d[4];
        int8_t* e[4];
        for (int k = 0; k < 4; ++k) {
            d[k] = xalloc(n * 2);
            e[k] = fmt == 3 ? xalloc(n / 16) : NULL;
        }
        const float scale = fmt == 2 ? 1024.f : 1.f;
        oracle_encode16(fmt, src, d[0], e[0], n, scale);
        oracle_decode16(fmt, d[0], e[0], back, n, scale);
        double worst = 0.0;
        for (size_t i = 0; i < n; ++i) worst = fmax(worst, fabs((double)back[i] - src[i]));
        CHECK(worst < 1e-2);
        oracle_encode16(fmt, src, d[1], e[1], n, scale);
        oracle_encode16(fmt, v, d[2], e[2], n, fmt == 2 ? 1.0f / 4096.f : 1.f);
        oracle_wave_step16(fmt, d[0], e[0], d[1], e[1], d[2], e[2], d[3], e[3], nx, ny, nz, 10.f, 10.f, 10.f,
                           1e-3f, scale, fmt == 2 ? 1.0f / 4096.f : 1.f);
        for (int k = 0; k < 4; ++k) {
            free(d[k]);
            free(e[k]);
        }
    }
    CHECK(oracle_fp16_to_f32(oracle_f32_to_fp16(1.5f)) == 1.5f);
    CHECK(oracle_bf16_to_f32(oracle_f32_to_bf16(-2.0f)) == -2.0f);
    free(src);
    free(back);
    free(v);
}

static void check_fwi(size_t nx, size_t ny, size_t nz, int steps) {
    const size_t n = nx * ny * nz, plane = nx * ny;
    const size_t src = (5 * ny + 6) * nx + 7, rec_z = 5;
    float* m = xalloc(n * 4);
    float* w = xalloc(steps * 4);
    float* tr = xalloc((size_t)steps * plane * 4);
    float* g = xalloc(n * 4);
    float* st = xalloc(steps * 4);
    float* out = xalloc(n * 4);
    oracle_init_constant_velocity_f32(m, n, 1500.f);
    for (int k = 0; k < steps; ++k) w[k] = (float)sin(0.3 * k);
    oracle_fwi_forward_f32(m, nx, ny, nz, 10.f, 10.f, 10.f, 1e-3f, steps, src, rec_z, w, tr);
    CHECK(all_finite_f(tr, (size_t)steps * plane));
    const double J = oracle_fwi_gradient_f32(m, nx, ny, nz, 10.f, 10.f, 10.f, 1e-3f, steps, src, rec_z, w, tr, g, st);
    CHECK(J == 0.0);
    int zero = 1;
    for (size_t i = 0; i < n; ++i) zero = zero && g[i] == 0.f;
    CHECK(zero);
    oracle_shot_run_f32(m, nx, ny, nz, 10.f, 10.f, 10.f, 1e-3f, steps, src, w, out);
    CHECK(all_finite_f(out, n));
    double* m64 = xalloc(n * 8);
    double* w64 = xalloc(steps * 8);
    double* t64 = xalloc((size_t)steps * plane * 8);
    double* g64 = xalloc(n * 8);
    for (size_t i = 0; i < n; ++i) m64[i] = 2.25e6;
    for (int k = 0; k < steps; ++k) w64[k] = sin(0.3 * k);
    oracle_fwi_forward_f64(m64, nx, ny, nz, 10.0, 10.0, 10.0, 1e-3, steps, src, rec_z, w64, t64);
    for (size_t i = 0; i < (size_t)steps * plane; ++i) t64[i] *= 0.5;
    CHECK(oracle_fwi_gradient_f64(m64, nx, ny, nz, 10.0, 10.0, 10.0, 1e-3, steps, src, rec_z, w64, t64, g64,
                                  NULL) > 0.0);
    free(m64);
    free(w64);
    free(t64);
    free(g64);
    free(m);
    free(w);
    free(tr);
    free(g);
    free(st);
    free(out);
}

int main(void) {
    float c[5];
    oracle_fd_coeffs_f32(c);
    CHECK(c[0] < 0.f);
    CHECK(oracle_make_grid_dt_f32(10.f, 10.f, 10.f, 1500.f) > 0.f);
    CHECK(oracle_heat_dt_eff_f32(4.f) == 2.f);
    CHECK(oracle_u01(1, 2, 3) >= 0.f && oracle_u01(1, 2, 3) < 1.f);
    (void)oracle_splitmix64(42);
    check_wave_and_heat(9, 10, 11);
    check_wave_and_heat(13, 9, 17);
    check_16bit(16, 11, 10);
    check_16bit(32, 9, 12);
    check_fwi(15, 14, 13, 6);
    printf("oracle_check: %d failed\n", failures);
    return failures ? 1 : 0;
}
