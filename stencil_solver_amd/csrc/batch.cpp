// Batched multi-shot forward modelling (no reference counterpart: the
// reference runs one grid per process).
//
// Ragged mode (the default): every time step of every shot, whatever its
// grid, is ONE launch of the tiled forward kernel over a table of work items
// (shot, tile, z-chunk), so the small grids fill the CUs together with the
// large ones instead of running as separate, under-filled launches.  The
// table is built on the host:
//   * each shot is tiled with the batch's tile shape and cut into z-chunks of
//     about `chunk_target` planes;
//   * the items (shots in order of decreasing size, each shot's tiles in row
//     order) are cut into eight runs of equal estimated cost, one per XCD, so
//     every XCD has the same work and a shot's neighbouring tiles share an L2;
//   * each run is ordered longest item first (the short ones fill the tail),
//     and the runs are interleaved so that workgroup b, which the dispatcher
//     sends to XCD b % 8, is the next item of that XCD's run (a shorter run is
//     padded with empty items).
// The batch autotuner times this launch itself on the batch's buffers for
// every built tile shape and a few chunk lengths, with and without the XCD
// runs, and keeps the fastest.  The whole run -- zeroing the three time levels
// of every shot, then n_steps launches -- is captured once into a hipGraph;
// a run is one graph launch.
//
// Persistent mode (STENCIL_BATCH_MODE=persistent, the same plan): the n_steps
// launches are replaced by ONE launch of as many workgroups as stay resident
// together, which take each step's items from per-XCD queues and meet at a
// grid barrier between steps (device counters the graph zeroes first; a
// barrier that times out aborts the run and stencil_batch_check reports it).
//
// Grouped mode (STENCIL_BATCH_MODE=grouped, kept for comparison): shots of
// equal grids form a group; each group has its own tuned shape and one launch
// per step (workgroup row = shot), the group chains being independent graph
// branches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <numeric>
#include <tuple>
#include <vector>

#include "launchers.hpp"
#include "stencil_capi.h"

namespace stencil::hipk {
hipError_t fill_f32(float*, size_t, float, hipStream_t);
}

using namespace stencil::hipk;

namespace {

// A ragged launch plan: tile shape, z-chunk length target, XCD runs on/off.
// Words of the persistent kernel's counters: abort flag, barrier, 8 queues.
constexpr size_t kSyncWords = 10;

struct RaggedPlan {
    int txv = 16, ty = 16;
    int chunk_target = 90;
    int xcd_runs = 1;
};

}  // namespace

struct stencil_batch {
    struct Group {
        Dims d;
        Scalars<float> s;
        ZqShape shape;
        std::vector<int> shots;
        BatchShot* table = nullptr;  // device (grouped mode)
        int n_chunks = 0;            // ragged mode: z-chunks per column of this grid
    };
    std::vector<Group> groups;
    std::vector<stencil_shot> shots;
    int n_steps = 0;
    bool ragged = true;
    bool persistent = false;  // ragged plan, one cooperative launch for all steps
    int persist_grid = 0;     // its workgroups
    unsigned barrier_polls = kBarrierPolls;  // polls before a barrier times out
    unsigned* d_sync = nullptr;  // device: abort flag, barrier counter, 8 queues
    RaggedPlan plan;
    RaggedShot* d_shots = nullptr;  // device, one per shot
    RaggedItem* d_items = nullptr;  // device, n_items
    int n_items = 0;
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

bool batch_mode_is(const char* name) {
    const char* e = std::getenv("STENCIL_BATCH_MODE");
    return e && std::strcmp(e, name) == 0;
}

void release(stencil_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto& g : b->groups)
        if (g.table) (void)hipFree(g.table);
    if (b->d_shots) (void)hipFree(b->d_shots);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_sync) (void)hipFree(b->d_sync);
    delete b;
}

// ---- ragged plan -------------------------------------------------------

// Per-shot tiling for this plan, and the work items in dispatch order.
void build_plan(const std::vector<stencil_shot>& shots, const RaggedPlan& p, std::vector<RaggedShot>& rs,
                std::vector<RaggedItem>& out, std::vector<int>* chunks_per_shot = nullptr) {
    const int n = int(shots.size());
    rs.assign(n, RaggedShot{});
    struct Item {
        RaggedItem it;
        double cost;
    };
    std::vector<std::vector<Item>> per_shot(n);
    if (chunks_per_shot) chunks_per_shot->assign(n, 0);
    const double threads = double(p.txv) * double(p.ty);
    for (int i = 0; i < n; ++i) {
        const stencil_shot& sh = shots[i];
        RaggedShot& r = rs[i];
        for (int k = 0; k < 3; ++k) r.ring[k] = sh.ring[k];
        r.vel2 = sh.vel2;
        r.wavelet = sh.wavelet;
        r.src_z = int(sh.src_z);
        r.src_y = int(sh.src_y);
        r.src_x = int(sh.src_x);
        r.nx = int(sh.grid.nx);
        r.ny = int(sh.grid.ny);
        r.nz = int(sh.grid.nz);
        r.tiles_x = int((sh.grid.nx + 4 * p.txv - 1) / (4 * p.txv));
        r.tiles_y = int((sh.grid.ny - 8 + p.ty - 1) / p.ty);
        const int64_t inz = int64_t(sh.grid.nz) - 8;
        int64_t want = (inz + p.chunk_target / 2) / std::max(1, p.chunk_target);
        if (want < 1) want = 1;
        r.chunk = zq_chunk_planes(inz, int(want));
        // a workgroup's buffer resources span its chunk plus 21 planes: keep
        // them below 2^32 bytes (zq_supported guarantees a 9-plane chunk fits)
        const int64_t max_chunk = (int64_t(0xFFFFFFFF) / (int64_t(sh.grid.nx) * int64_t(sh.grid.ny) * 4) - 21) / 9 * 9;
        if (r.chunk > max_chunk) r.chunk = int(max_chunk);
        r.s = make_scalars<float>(float(sh.grid.dx), float(sh.grid.dy), float(sh.grid.dz), float(sh.grid.dt));
        const int n_ch = int((inz + r.chunk - 1) / r.chunk);
        if (chunks_per_shot) (*chunks_per_shot)[i] = n_ch;
        for (int c = 0; c < n_ch; ++c) {
            const int64_t planes = std::min<int64_t>(r.chunk, inz - int64_t(c) * r.chunk);
            // steps run in whole groups of 9, after an 8-plane queue fill
            const double cost = threads * (double((planes + 8) / 9 * 9) + 4.0);
            for (int ty = 0; ty < r.tiles_y; ++ty)
                for (int tx = 0; tx < r.tiles_x; ++tx) per_shot[i].push_back({{i, tx, ty, c}, cost});
        }
    }
    // shots in order of decreasing size (ties: input order)
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const auto sa = shots[a].grid.nx * shots[a].grid.ny * shots[a].grid.nz;
        const auto sb = shots[b].grid.nx * shots[b].grid.ny * shots[b].grid.nz;
        return sa > sb;
    });
    std::vector<Item> seq;
    for (int i : order) seq.insert(seq.end(), per_shot[i].begin(), per_shot[i].end());
    auto by_cost = [](const Item& a, const Item& b) { return a.cost > b.cost; };
    out.clear();
    if (!p.xcd_runs) {
        std::stable_sort(seq.begin(), seq.end(), by_cost);
        for (const Item& x : seq) out.push_back(x.it);
        return;
    }
    constexpr int kXcds = 8;
    double total = 0.0;
    for (const Item& x : seq) total += x.cost;
    std::vector<std::vector<Item>> runs(kXcds);
    double before = 0.0;
    for (const Item& x : seq) {
        // the run whose cost interval holds this item's midpoint
        int k = int((before + 0.5 * x.cost) / total * kXcds);
        k = std::min(std::max(k, 0), kXcds - 1);
        runs[k].push_back(x);
        before += x.cost;
    }
    size_t longest = 0;
    for (auto& r : runs) {
        std::stable_sort(r.begin(), r.end(), by_cost);
        longest = std::max(longest, r.size());
    }
    for (size_t j = 0; j < longest; ++j)
        for (int k = 0; k < kXcds; ++k) out.push_back(j < runs[k].size() ? runs[k][j].it : RaggedItem{-1, 0, 0, 0});
}

// Uploads the plan's shot table and items (replacing the batch's).
hipError_t upload_plan(stencil_batch* b, const RaggedPlan& p, hipStream_t st) {
    std::vector<RaggedShot> rs;
    std::vector<RaggedItem> items;
    std::vector<int> chunks;
    build_plan(b->shots, p, rs, items, &chunks);
    if (b->d_shots) (void)hipFree(b->d_shots);
    if (b->d_items) (void)hipFree(b->d_items);
    b->d_shots = nullptr;
    b->d_items = nullptr;
    b->n_items = 0;
    hipError_t e = hipMalloc(&b->d_shots, rs.size() * sizeof(RaggedShot));
    if (e == hipSuccess) e = hipMalloc(&b->d_items, items.size() * sizeof(RaggedItem));
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_shots, rs.data(), rs.size() * sizeof(RaggedShot), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_items, items.data(), items.size() * sizeof(RaggedItem), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    b->n_items = int(items.size());
    b->plan = p;
    for (auto& g : b->groups) g.n_chunks = chunks[g.shots.front()];
    return hipSuccess;
}

hipError_t zero_rings(const stencil_batch* b, hipStream_t st) {
    for (const stencil_shot& s : b->shots) {
        const size_t cells = size_t(s.grid.nx) * size_t(s.grid.ny) * size_t(s.grid.nz);
        for (float* r : s.ring) {
            const hipError_t e = fill_f32(r, cells, 0.f, st);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// The batch autotuner: every built tile shape x chunk target x XCD runs on /
// off, each timed over `reps` steps run as the batch runs them (per-step
// launches, or persistent launches for a persistent batch) on the batch's own
// buffers (zeroed first; the run starts from zero again anyway).  A sweep
// timing carries a few per cent of noise, enough to reorder plans within a
// few per cent of each other: the four fastest plans are timed again, twice
// each with at least 32 launches, and each keeps its best time.
hipError_t tune_plan(stencil_batch* b, int reps, hipStream_t st) {
    static const int shapes[][2] = {{16, 16}, {16, 32}, {32, 8}, {32, 16}, {64, 4}, {64, 8}, {64, 12}, {64, 16}};
    hipError_t e = zero_rings(b, st);
    if (e != hipSuccess) return e;
    struct Row {
        RaggedPlan plan;
        float ms;
    };
    std::vector<Row> rows;
    // ms per step of this plan, run as the batch will run it: one launch per
    // step, or (persistent batches) one persistent launch of kSteps steps per
    // timed call, counters zeroed first
    constexpr int kSteps = 8;
    auto time_plan = [&](const RaggedPlan& p, int n, float* ms) -> hipError_t {
        const hipError_t ue = upload_plan(b, p, st);
        if (ue != hipSuccess) return ue;
        if (b->persistent) {
            const int grid = std::min(ragged_persist_grid(p.txv, p.ty), (b->n_items + 7) / 8 * 8);
            if (grid < 8) {  // the plan cannot run persistently: not a candidate
                *ms = 1e30f;
                return hipSuccess;
            }
            const int steps = std::min(kSteps, b->n_steps);
            const float call = zq_time_launches(
                [&] {
                    const hipError_t fe = fill_f32(reinterpret_cast<float*>(b->d_sync), kSyncWords, 0.f, st);
                    if (fe != hipSuccess) return fe;
                    return ragged_persist_fwi_f32(b->d_shots, b->d_items, b->n_items, steps, p.txv, p.ty, grid,
                                                  b->d_sync, b->barrier_polls, st);
                },
                std::max(1, n / steps), st);
            *ms = call < 0.f ? -1.f : call / float(steps);
        } else {
            int step = 0;
            *ms = zq_time_launches(
                [&] {
                    const int k = step++;
                    return ragged_zq_fwi_f32(b->d_shots, b->d_items, b->n_items, k % 3, k % b->n_steps, p.txv,
                                             p.ty, true, st);
                },
                n, st);
        }
        return *ms < 0.f ? hipErrorLaunchFailure : hipSuccess;
    };
    for (const auto& sh : shapes)
        for (int target : {18, 27, 54, 90, 144, 4096})
            for (int runs : {1, 0}) {
                RaggedPlan p;
                p.txv = sh[0];
                p.ty = sh[1];
                p.chunk_target = target;
                p.xcd_runs = runs;
                float ms = -1.f;
                e = time_plan(p, reps, &ms);
                if (e != hipSuccess) return e;
                rows.push_back({p, ms});
            }
    std::sort(rows.begin(), rows.end(), [](const Row& x, const Row& y) { return x.ms < y.ms; });
    if (rows.size() > 4) rows.resize(4);
    for (int round = 0; round < 2; ++round)
        for (Row& r : rows) {
            float ms = -1.f;
            e = time_plan(r.plan, std::max(8 * reps, 32), &ms);
            if (e != hipSuccess) return e;
            if (ms < r.ms) r.ms = ms;
        }
    const Row* best = &rows.front();
    for (const Row& r : rows)
        if (r.ms < best->ms) best = &r;
    return upload_plan(b, best->plan, st);
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
    // stream, which cannot capture); in grouped mode one forked stream per
    // group joins it.
    const size_t n_branches = b->ragged ? 0 : b->groups.size();
    std::vector<hipStream_t> streams(n_branches, nullptr);
    std::vector<hipEvent_t> done(n_branches, nullptr);
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
        if (b->ragged) {
            // p^{-1} = p^0 = 0 (and the third level, so halos stay zero): fill
            // kernels, not memset nodes
            B_TRY(zero_rings(b, st));
            if (b->persistent) {
                const int cap = ragged_persist_grid(b->plan.txv, b->plan.ty);
                b->persist_grid = std::min(cap, (b->n_items + 7) / 8 * 8);
                if (b->persist_grid < 8) B_TRY(hipErrorNotSupported);
                B_TRY(fill_f32(reinterpret_cast<float*>(b->d_sync), kSyncWords, 0.f, st));  // zero bits
                B_TRY(ragged_persist_fwi_f32(b->d_shots, b->d_items, b->n_items, b->n_steps, b->plan.txv, b->plan.ty,
                                             b->persist_grid, b->d_sync, b->barrier_polls, st));
            } else {
                for (int n = 0; n < b->n_steps; ++n)
                    B_TRY(ragged_zq_fwi_f32(b->d_shots, b->d_items, b->n_items, n % 3, n, b->plan.txv, b->plan.ty,
                                            true, st));
            }
        } else {
            B_TRY(hipEventRecord(fork, st));
            for (size_t gi = 0; gi < b->groups.size(); ++gi) {
                auto& g = b->groups[gi];
                hipStream_t gs = streams[gi];
                B_TRY(hipStreamWaitEvent(gs, fork, 0));
                const size_t cells = size_t(g.d.nx) * size_t(g.d.ny) * size_t(g.d.nz);
                for (int si : g.shots)
                    for (float* r : b->shots[si].ring) B_TRY(fill_f32(r, cells, 0.f, gs));
                for (int n = 0; n < b->n_steps; ++n)
                    B_TRY(batch_zq_fwi_f32(g.table, int(g.shots.size()), n % 3, n, g.d, g.s, g.shape, gs));
                B_TRY(hipEventRecord(done[gi], gs));
                B_TRY(hipStreamWaitEvent(st, done[gi], 0));
            }
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

// Grouped mode: each group's shape and device table.
int setup_groups(stencil_batch* b, int tune, void* stream) {
    const hipStream_t st = static_cast<hipStream_t>(stream);
    for (auto& g : b->groups) {
        const stencil_shot& s0 = b->shots[g.shots.front()];
        if (tune) {
            // times candidate shapes on the first shot's buffers (overwrites ring[2])
            stencil_tile_shape best;
            const int rc = stencil_tune_tiled_f32(s0.ring[1], s0.ring[0], s0.vel2, s0.ring[2], &s0.grid, 3, &best,
                                                  nullptr, stream);
            if (rc != STENCIL_OK) return rc;
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
            hipStreamSynchronize(st) != hipSuccess)
            return STENCIL_ERR_HIP;
    }
    return STENCIL_OK;
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
    b->ragged = !batch_mode_is("grouped");
    b->persistent = b->ragged && batch_mode_is("persistent");
    // STENCIL_BATCH_BARRIER_POLLS: a test's way to make a persistent run's
    // barriers time out (every workgroup still leaves at its next barrier)
    if (const char* e = std::getenv("STENCIL_BATCH_BARRIER_POLLS")) {
        char* end = nullptr;
        const unsigned long v = std::strtoul(e, &end, 10);
        if (end != e && *end == '\0' && v < kBarrierPolls) b->barrier_polls = unsigned(v);
    }
    // group by grid (the grouped mode's launches; reported in either mode)
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
    int rc = STENCIL_OK;
    if (b->persistent && hipMalloc(&b->d_sync, kSyncWords * sizeof(unsigned)) != hipSuccess) rc = STENCIL_ERR_HIP;
    if (rc != STENCIL_OK) {
    } else if (b->ragged) {
        const hipError_t e = tune ? tune_plan(b, 4, st) : upload_plan(b, RaggedPlan{}, st);
        if (e != hipSuccess) rc = STENCIL_ERR_HIP;
    } else {
        rc = setup_groups(b, tune, stream);
    }
    if (rc == STENCIL_OK) rc = build_graph(b);
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
        if (b->ragged) {
            // one tile shape for the whole batch; this grid's z-chunk count;
            // xcd_remap reports the per-XCD runs
            shape->tile_x4 = b->plan.txv;
            shape->tile_y = b->plan.ty;
            shape->n_chunks = g.n_chunks;
            shape->xcd_remap = b->plan.xcd_runs;
        } else {
            const ZqShape r = zq_resolve(g.d, g.shape, 0);
            shape->tile_x4 = r.txv;
            shape->tile_y = r.ty;
            shape->n_chunks = r.n_chunks;
            shape->xcd_remap = r.xcd_remap;
        }
        shape->nontemporal = 1;
    }
    if (n_shots) *n_shots = int(g.shots.size());
    return STENCIL_OK;
}

int stencil_batch_set_plan(stencil_batch* b, const stencil_tile_shape* shape, int chunk_planes, void* stream) {
    if (!b || !shape || chunk_planes < 1 || !zq_shape_valid(shape->tile_x4, shape->tile_y)) return STENCIL_ERR_INVALID;
    if (!b->ragged) return STENCIL_ERR_UNSUPPORTED;
    if (hipDeviceSynchronize() != hipSuccess) return STENCIL_ERR_HIP;  // no run of the old graph in flight
    RaggedPlan p;
    p.txv = shape->tile_x4;
    p.ty = shape->tile_y;
    p.chunk_target = chunk_planes;
    p.xcd_runs = shape->xcd_remap != 0 ? 1 : 0;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    if (upload_plan(b, p, static_cast<hipStream_t>(stream)) != hipSuccess) return STENCIL_ERR_HIP;
    return build_graph(b);
}

int stencil_batch_is_ragged(const stencil_batch* b) { return b ? (b->ragged ? 1 : 0) : STENCIL_ERR_INVALID; }

int stencil_batch_check(const stencil_batch* b, void* stream) {
    if (!b) return STENCIL_ERR_INVALID;
    if (hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess) return STENCIL_ERR_HIP;
    if (!b->d_sync) return STENCIL_OK;
    unsigned flag = 0;
    if (hipMemcpy(&flag, b->d_sync, sizeof(flag), hipMemcpyDeviceToHost) != hipSuccess) return STENCIL_ERR_HIP;
    return flag == 0 ? STENCIL_OK : STENCIL_ERR_TIMEOUT;
}

int stencil_batch_persistent_grid(const stencil_batch* b) {
    return b ? (b->persistent ? b->persist_grid : 0) : STENCIL_ERR_INVALID;
}

void stencil_batch_destroy(stencil_batch* b) { release(b); }

}  // extern "C"
