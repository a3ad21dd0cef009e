// Batched multi-shot forward modelling (no reference counterpart: the
// reference runs one grid per process).
//
// Shots are grouped by grid (dims, spacing, dt).  For each group the launch
// shape is the wave autotuner's (run on the group's first shot when asked, or
// taken from its cache / the default), and one launch of the tiled forward
// kernel steps every shot of the group (workgroup row = shot, pointers from a
// device table).  The whole run -- zeroing the three time levels of every
// shot, then n_steps steps of every group -- is captured once into a hipGraph
// whose group chains are independent branches (forked capture streams), so a
// run is one graph launch and small groups overlap on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "launchers.hpp"
#include "stencil_capi.h"

namespace stencil::hipk {
hipError_t fill_f32(float*, size_t, float, hipStream_t);
}

using namespace stencil::hipk;

struct stencil_batch {
    struct Group {
        Dims d;
        Scalars<float> s;
        ZqShape shape;
        std::vector<int> shots;
        BatchShot* table = nullptr;  // device
    };
    std::vector<Group> groups;
    std::vector<stencil_shot> shots;
    int n_steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

bool shot_ok(const stencil_shot& s) {
    const stencil_grid& g = s.grid;
    if (g.nx <= 8 || g.ny <= 8 || g.nz <= 8 || g.nx > (1u << 20) || g.ny > (1u << 20) || g.nz > (1u << 20))
        return false;
    if (!(g.dx > 0) || !(g.dy > 0) || !(g.dz > 0) || !(g.dt > 0)) return false;
    if (!s.ring[0] || !s.ring[1] || !s.ring[2] || !s.vel2) return false;
    if (s.ring[0] == s.ring[1] || s.ring[1] == s.ring[2] || s.ring[0] == s.ring[2]) return false;
    auto inside = [](uint64_t v, uint64_t n) { return v >= 4 && v + 4 < n; };
    if (s.wavelet && (!inside(s.src_x, g.nx) || !inside(s.src_y, g.ny) || !inside(s.src_z, g.nz))) return false;
    Dims d;
    d.nx = int64_t(g.nx);
    d.ny = int64_t(g.ny);
    d.nz = int64_t(g.nz);
    const void* ptrs[4] = {s.ring[0], s.ring[1], s.ring[2], s.vel2};
    return zq_supported(d, ptrs, 4);
}

void release(stencil_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto& g : b->groups)
        if (g.table) (void)hipFree(g.table);
    delete b;
}

#define B_TRY(x)                        \
    do {                                \
        if ((x) != hipSuccess) {        \
            ok = false;                 \
            goto capture_end;           \
        }                               \
    } while (0)

int build_graph(stencil_batch* b) {
    // Captured on a stream of its own (the caller's may be the legacy default
    // stream, which cannot capture); one forked stream per group joins it.
    std::vector<hipStream_t> streams(b->groups.size(), nullptr);
    std::vector<hipEvent_t> done(b->groups.size(), nullptr);
    hipEvent_t fork = nullptr;
    hipStream_t st = nullptr;
    bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
    for (size_t i = 0; ok && i < streams.size(); ++i) {
        if (hipStreamCreateWithFlags(&streams[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess)
            ok = false;
    }
    if (ok && hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess) ok = false;
    if (ok && hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) ok = false;
    if (ok) {
        B_TRY(hipEventRecord(fork, st));
        for (size_t gi = 0; gi < b->groups.size(); ++gi) {
            auto& g = b->groups[gi];
            hipStream_t gs = streams[gi];
            B_TRY(hipStreamWaitEvent(gs, fork, 0));
            const size_t cells = size_t(g.d.nx) * size_t(g.d.ny) * size_t(g.d.nz);
            // p^{-1} = p^0 = 0 (and the third level, so halos stay zero): a fill
            // kernel, not a memset node
            for (int si : g.shots)
                for (float* r : b->shots[si].ring) B_TRY(fill_f32(r, cells, 0.f, gs));
            for (int n = 0; n < b->n_steps; ++n)
                B_TRY(batch_zq_fwi_f32(g.table, int(g.shots.size()), n % 3, n, g.d, g.s, g.shape, gs));
            B_TRY(hipEventRecord(done[gi], gs));
            B_TRY(hipStreamWaitEvent(st, done[gi], 0));
        }
    capture_end:
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess || !graph) ok = false;
        if (ok) {
            b->graph = graph;
            ok = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0) == hipSuccess;
        } else if (graph) {
            (void)hipGraphDestroy(graph);
        }
    }
    for (auto s : streams)
        if (s) (void)hipStreamDestroy(s);
    for (auto e : done)
        if (e) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (st) (void)hipStreamDestroy(st);
    return ok ? STENCIL_OK : STENCIL_ERR_HIP;
}

}  // namespace

extern "C" {

int stencil_batch_create(const stencil_shot* shots, int n_shots, int n_steps, int tune, stencil_batch** out,
                         void* stream) {
    if (!out) return STENCIL_ERR_INVALID;
    *out = nullptr;
    if (!shots || n_shots <= 0 || n_steps <= 0 || n_steps > (1 << 30)) return STENCIL_ERR_INVALID;
    for (int i = 0; i < n_shots; ++i)
        if (!shot_ok(shots[i])) return STENCIL_ERR_INVALID;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    auto* b = new (std::nothrow) stencil_batch;
    if (!b) return STENCIL_ERR_HIP;
    b->shots.assign(shots, shots + n_shots);
    b->n_steps = n_steps;
    // group by grid
    using Key = std::tuple<uint64_t, uint64_t, uint64_t, double, double, double, double>;
    std::map<Key, int> index;
    for (int i = 0; i < n_shots; ++i) {
        const stencil_grid& g = shots[i].grid;
        const Key k{g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.dt};
        auto it = index.find(k);
        if (it == index.end()) {
            stencil_batch::Group grp;
            grp.d.nx = int64_t(g.nx);
            grp.d.ny = int64_t(g.ny);
            grp.d.nz = int64_t(g.nz);
            grp.s = make_scalars<float>(float(g.dx), float(g.dy), float(g.dz), float(g.dt));
            it = index.emplace(k, int(b->groups.size())).first;
            b->groups.push_back(grp);
        }
        b->groups[it->second].shots.push_back(i);
    }
    for (auto& g : b->groups) {
        const stencil_shot& s0 = b->shots[g.shots.front()];
        if (tune) {
            // times candidate shapes on the first shot's buffers (overwrites ring[2])
            stencil_tile_shape best;
            const int rc = stencil_tune_tiled_f32(s0.ring[1], s0.ring[0], s0.vel2, s0.ring[2], &s0.grid, 3, &best,
                                                  nullptr, stream);
            if (rc != STENCIL_OK) {
                release(b);
                return rc;
            }
        }
        if (!zq_cached_shape(g.d, 0, &g.shape)) g.shape = zq_resolve(g.d, ZqShape{}, 0);
        std::vector<BatchShot> host(g.shots.size());
        for (size_t i = 0; i < g.shots.size(); ++i) {
            const stencil_shot& s = b->shots[g.shots[i]];
            for (int r = 0; r < 3; ++r) host[i].ring[r] = s.ring[r];
            host[i].vel2 = s.vel2;
            host[i].wavelet = s.wavelet;
            host[i].src_z = int(s.src_z);
            host[i].src_y = int(s.src_y);
            host[i].src_x = int(s.src_x);
            host[i].pad = 0;
        }
        if (hipMalloc(&g.table, host.size() * sizeof(BatchShot)) != hipSuccess ||
            hipMemcpyAsync(g.table, host.data(), host.size() * sizeof(BatchShot), hipMemcpyHostToDevice, st) !=
                hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            release(b);
            return STENCIL_ERR_HIP;
        }
    }
    const int rc = build_graph(b);
    if (rc != STENCIL_OK) {
        release(b);
        return rc;
    }
    *out = b;
    return STENCIL_OK;
}

int stencil_batch_run(stencil_batch* b, void* stream) {
    if (!b || !b->exec) return STENCIL_ERR_INVALID;
    return hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)) == hipSuccess ? STENCIL_OK : STENCIL_ERR_HIP;
}

int stencil_batch_newest(const stencil_batch* b, int shot) {
    if (!b || shot < 0 || shot >= int(b->shots.size())) return STENCIL_ERR_INVALID;
    return (b->n_steps + 1) % 3;
}

int stencil_batch_groups(const stencil_batch* b) { return b ? int(b->groups.size()) : STENCIL_ERR_INVALID; }

int stencil_batch_group(const stencil_batch* b, int group, stencil_grid* grid, stencil_tile_shape* shape,
                        int* n_shots) {
    if (!b || group < 0 || group >= int(b->groups.size())) return STENCIL_ERR_INVALID;
    const auto& g = b->groups[group];
    if (grid) *grid = b->shots[g.shots.front()].grid;
    if (shape) {
        const ZqShape r = zq_resolve(g.d, g.shape, 0);
        shape->tile_x4 = r.txv;
        shape->tile_y = r.ty;
        shape->n_chunks = r.n_chunks;
        shape->xcd_remap = r.xcd_remap;
        shape->nontemporal = 1;
    }
    if (n_shots) *n_shots = int(g.shots.size());
    return STENCIL_OK;
}

void stencil_batch_destroy(stencil_batch* b) { release(b); }

}  // extern "C"
