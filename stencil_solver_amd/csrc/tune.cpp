// Launch-shape autotuner for the tiled kernels, and the per-grid shape cache
// the launchers consult.  Replaces the reference's hand-run TILE_Y x PENCIL_Z
// sweeps (tuning_results/sweep*.csv, docs/CASE_STUDY.md section 3.2): every
// built tile shape x z-chunk count x streaming mode is timed with hipEvents on
// the caller's stream and buffers, the fastest is cached for that grid size,
// and the sweep can be written as CSV in the same spirit as the reference's
// (one row per configuration, naive and tiled ms, speedup).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "launchers.hpp"
#include "stencil_capi.h"

namespace stencil::hipk {
namespace {

using Key = std::tuple<int64_t, int64_t, int64_t, int>;  // nx, ny, nz, kind

std::mutex g_mu;
std::map<Key, ZqShape>& cache() {
    static std::map<Key, ZqShape> c;
    return c;
}

}  // namespace

bool zq_cached_shape(const Dims& d, int kind, ZqShape* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = cache().find(Key{d.nx, d.ny, d.nz, kind});
    if (it == cache().end()) return false;
    *out = it->second;
    return true;
}

void zq_cache_put(const Dims& d, int kind, const ZqShape& s) {
    std::lock_guard<std::mutex> lk(g_mu);
    cache()[Key{d.nx, d.ny, d.nz, kind}] = s;
}

void zq_cache_clear() {
    std::lock_guard<std::mutex> lk(g_mu);
    cache().clear();
}

float zq_time_launches(const std::function<hipError_t()>& fn, int reps, hipStream_t st) {
    for (int i = 0; i < 2; ++i)
        if (fn() != hipSuccess) return -1.f;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1.f;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1.f;
    }
    float ms = -1.f;
    bool ok = hipEventRecord(a, st) == hipSuccess;
    for (int i = 0; ok && i < reps; ++i) ok = fn() == hipSuccess;
    ok = ok && hipEventRecord(b, st) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
         hipEventElapsedTime(&ms, a, b) == hipSuccess;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ok ? ms / static_cast<float>(reps) : -1.f;
}

std::vector<ZqShape> zq_tune_candidates(const Dims& d) {
    static const int shapes[][2] = {{16, 16}, {16, 32}, {32, 8}, {32, 16},
                                    {64, 4},  {64, 8},  {64, 12}, {64, 16}};
    const int64_t inz = d.nz - 8;
    std::vector<int> chunks;
    for (int c : {1, 2, 4, 8, 16})
        if (c == 1 || inz / c >= 9) chunks.push_back(c);
    std::vector<ZqShape> out;
    // xcd_remap 0 / 1: workgroup order as dispatched, or each XCD a
    // contiguous range of tiles.
    for (const auto& sh : shapes)
        for (int c : chunks)
            for (int nt : {1, 0})
                for (int remap : {1, 0}) {
                    ZqShape s;
                    s.txv = sh[0];
                    s.ty = sh[1];
                    s.n_chunks = c;
                    s.xcd_remap = remap;
                    s.nt = nt;
                    out.push_back(s);
                }
    return out;
}

}  // namespace stencil::hipk

using namespace stencil::hipk;

namespace {

struct Row {
    ZqShape shape;
    float ms;
};

// A sweep's timings carry a few per cent of launch-to-launch noise, enough to
// reorder shapes that are within a few per cent of each other.  The fastest
// kTop shapes are timed again, kRounds times each in turn with at least
// kRefineReps launches per timing, and each keeps its best time.
constexpr int kTop = 4;
constexpr int kRounds = 2;
constexpr int kRefineReps = 10;

void refine(std::vector<Row>& rows, int reps, const std::function<float(const ZqShape&, int)>& time_shape) {
    std::vector<Row*> order;
    for (auto& r : rows)
        if (r.ms > 0.f) order.push_back(&r);
    std::sort(order.begin(), order.end(), [](const Row* a, const Row* b) { return a->ms < b->ms; });
    if (order.size() > static_cast<size_t>(kTop)) order.resize(kTop);
    const int n = std::max(reps * 2, kRefineReps);
    for (int round = 0; round < kRounds; ++round)
        for (Row* r : order) {
            const float ms = time_shape(r->shape, n);
            if (ms > 0.f && ms < r->ms) r->ms = ms;
        }
}

int finish(const Dims& d, int kind, const std::vector<Row>& rows, float naive_ms,
           stencil_tile_shape* best, const char* csv_path) {
    const Row* top = nullptr;
    for (const auto& r : rows)
        if (r.ms > 0.f && (!top || r.ms < top->ms)) top = &r;
    if (!top) return STENCIL_ERR_HIP;
    zq_cache_put(d, kind, top->shape);
    if (best) {
        best->tile_x4 = top->shape.txv;
        best->tile_y = top->shape.ty;
        best->n_chunks = top->shape.n_chunks;
        best->xcd_remap = top->shape.xcd_remap;
        best->nontemporal = top->shape.nt;
    }
    if (csv_path && *csv_path) {
        FILE* f = std::fopen(csv_path, "w");
        if (!f) return STENCIL_ERR_INVALID;
        std::fprintf(f, "nx,ny,nz,tile_x,tile_y,n_chunks,nontemporal,xcd_remap,naive_ms,tiled_ms,speedup\n");
        for (const auto& r : rows)
            std::fprintf(f, "%lld,%lld,%lld,%d,%d,%d,%d,%d,%.5f,%.5f,%.3f\n", (long long)d.nx,
                         (long long)d.ny, (long long)d.nz, 4 * r.shape.txv, r.shape.ty,
                         r.shape.n_chunks, r.shape.nt, r.shape.xcd_remap, naive_ms, r.ms,
                         r.ms > 0.f ? naive_ms / r.ms : 0.f);
        std::fclose(f);
    }
    return STENCIL_OK;
}

bool dims_ok(const stencil_grid* g) {
    return g && g->nx > 8 && g->ny > 8 && g->nz > 8 && g->nx <= (1u << 30) && g->ny <= (1u << 30) &&
           g->nz <= (1u << 30);
}

}  // namespace

extern "C" {

int stencil_tune_tiled_f32(const float* p_cur, const float* p_prev, const float* vel2,
                           float* p_next, const stencil_grid* grid, int reps,
                           stencil_tile_shape* best, const char* csv_path, void* stream) {
    if (!p_cur || !p_prev || !vel2 || !p_next || !dims_ok(grid) || reps <= 0)
        return STENCIL_ERR_INVALID;
    if (p_next == p_cur || p_next == p_prev || p_next == vel2) return STENCIL_ERR_INVALID;
    Dims d;
    d.nx = int64_t(grid->nx);
    d.ny = int64_t(grid->ny);
    d.nz = int64_t(grid->nz);
    const void* ptrs[4] = {p_cur, p_prev, vel2, p_next};
    if (!zq_supported(d, ptrs, 4)) return STENCIL_ERR_UNSUPPORTED;
    const auto s = make_scalars<float>(float(grid->dx), float(grid->dy), float(grid->dz),
                                       float(grid->dt));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const float naive_ms = zq_time_launches(
        [&] { return wave_naive_f32(p_cur, p_prev, vel2, p_next, d, s, st); }, reps, st);
    auto time_shape = [&](const ZqShape& c, int n) {
        return zq_time_launches([&] { return wave_zq_f32(p_cur, p_prev, vel2, p_next, d, s, true, c, st); }, n, st);
    };
    std::vector<Row> rows;
    for (const ZqShape& c : zq_tune_candidates(d)) rows.push_back({c, time_shape(c, reps)});
    refine(rows, reps, time_shape);
    return finish(d, 0, rows, naive_ms, best, csv_path);
}

int stencil_tune_heat_f32(const float* t, float* t_next, const stencil_grid* grid, int reps,
                          stencil_tile_shape* best, const char* csv_path, void* stream) {
    if (!t || !t_next || t == t_next || !dims_ok(grid) || reps <= 0) return STENCIL_ERR_INVALID;
    Dims d;
    d.nx = int64_t(grid->nx);
    d.ny = int64_t(grid->ny);
    d.nz = int64_t(grid->nz);
    const void* ptrs[2] = {t, t_next};
    if (!zq_supported(d, ptrs, 2)) return STENCIL_ERR_UNSUPPORTED;
    const auto s = make_scalars<float>(float(grid->dx), float(grid->dy), float(grid->dz),
                                       float(grid->dt));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const float naive_ms = zq_time_launches(
        [&] { return heat_naive_f32(t, t_next, d, s, 0.f, st); }, reps, st);
    auto time_shape = [&](const ZqShape& c, int n) {
        return zq_time_launches([&] { return heat_zq_f32(t, t_next, d, s, 0.f, true, c, st); }, n, st);
    };
    std::vector<Row> rows;
    for (const ZqShape& c : zq_tune_candidates(d)) rows.push_back({c, time_shape(c, reps)});
    refine(rows, reps, time_shape);
    return finish(d, 1, rows, naive_ms, best, csv_path);
}

int stencil_tune_tiled_16(int storage, const stencil_field16* p_cur, const stencil_field16* p_prev,
                          const stencil_field16* vel2, const stencil_field16* p_next,
                          const stencil_grid* grid, float p_scale, float v_scale, int reps,
                          stencil_tile_shape* best, const char* csv_path, void* stream) {
    if (storage < STENCIL_STORE_BF16 || storage > STENCIL_STORE_BFP16 || !dims_ok(grid) || reps <= 0)
        return STENCIL_ERR_INVALID;
    const stencil_field16* f[4] = {p_cur, p_prev, vel2, p_next};
    for (const auto* x : f)
        if (!x || !x->data || (storage == STENCIL_STORE_BFP16) != (x->exp != nullptr))
            return STENCIL_ERR_INVALID;
    if (p_next->data == p_cur->data || p_next->data == p_prev->data || p_next->data == vel2->data)
        return STENCIL_ERR_INVALID;
    Dims d;
    d.nx = int64_t(grid->nx);
    d.ny = int64_t(grid->ny);
    d.nz = int64_t(grid->nz);
    if (d.nx % 16 != 0 || d.nx * d.ny * 2 >= (int64_t(1) << 32)) return STENCIL_ERR_UNSUPPORTED;
    const auto s = make_scalars<float>(float(grid->dx), float(grid->dy), float(grid->dz),
                                       float(grid->dt));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const Field16 c{p_cur->data, p_cur->exp}, p{p_prev->data, p_prev->exp};
    const Field16 v{vel2->data, vel2->exp}, n{p_next->data, p_next->exp};
    auto time_shape = [&](const ZqShape& sh, int k) {
        return zq_time_launches(
            [&] { return wave_zq_16(storage, c, p, v, n, d, s, p_scale, v_scale, false, sh, st); }, k, st);
    };
    std::vector<Row> rows;
    for (const ZqShape& sh : zq_tune_candidates(d)) rows.push_back({sh, time_shape(sh, reps)});
    refine(rows, reps, time_shape);
    return finish(d, 2 + storage, rows, 0.f, best, csv_path);
}

int stencil_tuned_shape(const stencil_grid* grid, int kind, stencil_tile_shape* out) {
    if (!dims_ok(grid) || !out || kind < 0 || kind > kKindFwiPair) return STENCIL_ERR_INVALID;
    Dims d;
    d.nx = int64_t(grid->nx);
    d.ny = int64_t(grid->ny);
    d.nz = int64_t(grid->nz);
    ZqShape s;
    const bool hit = zq_cached_shape(d, kind, &s);
    if (!hit) s = zq_resolve(d, ZqShape{}, kind);
    out->tile_x4 = s.txv;
    out->tile_y = s.ty;
    out->n_chunks = s.n_chunks;
    out->xcd_remap = s.xcd_remap;
    out->nontemporal = s.nt;
    return hit ? 1 : 0;
}

int stencil_set_tuned_shape(const stencil_grid* grid, int kind, const stencil_tile_shape* shape) {
    if (!dims_ok(grid) || !shape || kind < 0 || kind > kKindFwiPair || kind == 2) return STENCIL_ERR_INVALID;
    const bool built = kind == kKindFwiPair ? zq_pair_shape_valid(shape->tile_x4, shape->tile_y)
                                            : zq_shape_valid(shape->tile_x4, shape->tile_y);
    if (!built || shape->n_chunks < 0 || shape->xcd_remap < -1 ||
        shape->xcd_remap > 1 || shape->nontemporal < -1 || shape->nontemporal > 1)
        return STENCIL_ERR_INVALID;
    Dims d;
    d.nx = int64_t(grid->nx);
    d.ny = int64_t(grid->ny);
    d.nz = int64_t(grid->nz);
    ZqShape s;
    s.txv = shape->tile_x4;
    s.ty = shape->tile_y;
    s.n_chunks = shape->n_chunks;
    s.xcd_remap = shape->xcd_remap;
    s.nt = shape->nontemporal;
    zq_cache_put(d, kind, s);
    return STENCIL_OK;
}

void stencil_clear_tuning(void) { zq_cache_clear(); }

}  // extern "C"
