// Adjoint-state FWI gradient with sqrt(n) checkpointing (no reference
// counterpart).  The discrete scheme and its adjoint are stated in fwi.hip.
//
// Forward pass: n_steps acoustic steps from p^0 = p^-1 = 0 with a point source,
// the receiver plane of every level p^1..p^N copied into the trace array, and
// the level pair (p^{sC}, p^{sC+1}) kept for every segment s (C = ckpt_every,
// default ceil(sqrt(N))).  With observed data the traces become the residual
// R p^k - d^k in place and J = 1/2 |residual|^2 (FP64 reduction).
// Backward pass, segments last to first: the forward steps of segment s are
// re-run from its checkpoint, writing L p^k for k = sC+1 .. min((s+1)C, N-1);
// then the adjoint steps k = min((s+1)C, N) .. sC+1 consume them, in the
// variable psi = m lambda:
//   psi^k = 2 psi^{k+1} - psi^{k+2} + dt^2 m L psi^{k+1} + m res^k
//   grad += dt^2 (L p^k) psi^{k+1},   and at the end grad /= m.
// The recompute of segment s-1 runs on a second stream, overlapped with the
// adjoint steps of segment s (at small grids neither kernel fills the GPU).
// Device memory: all of it is the caller's workspace (stencil_fwi_workspace_bytes):
// 3 wavefield levels, 3 adjoint levels, 2 * ceil(N/C) checkpoint fields, C
// Laplacian fields (two banks of C with more than one segment), N receiver
// planes and the reduction scratch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "launchers.hpp"
#include "stencil_capi.h"

namespace stencil::hipk {

hipError_t fwi_fwd_step_f32(const float*, const float*, const float*, float*, float*, const Dims&,
                            const Scalars<float>&, long long, const float*, int, hipStream_t);
hipError_t fwi_fwd_step_f64(const double*, const double*, const double*, double*, double*, const Dims&,
                            const Scalars<double>&, long long, const double*, int, hipStream_t);
hipError_t fwi_adj_step_f32(const float*, const float*, const float*, const float*, const float*, float*, float*,
                            const float*, int, const Dims&, const Scalars<float>&, long long, float*, hipStream_t);
hipError_t fwi_adj_step_f64(const double*, const double*, const double*, const double*, const double*, double*,
                            double*, const double*, int, const Dims&, const Scalars<double>&, long long, double*,
                            hipStream_t);
hipError_t fwi_residual_f32(const float*, const float*, float*, size_t, double*, double*, hipStream_t);
hipError_t fwi_residual_f64(const double*, const double*, double*, size_t, double*, double*, hipStream_t);
hipError_t fill_f32(float*, size_t, float, hipStream_t);
hipError_t fill_f64(double*, size_t, double, hipStream_t);
hipError_t fwi_divide_f32(float*, const float*, const Dims&, hipStream_t);
hipError_t fwi_divide_f64(double*, const double*, const Dims&, hipStream_t);

}  // namespace stencil::hipk

using namespace stencil::hipk;

namespace {

constexpr size_t kAlign = 256;
constexpr size_t kPartials = 1024;  // residual reduction blocks

size_t up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// Source cell and launch choice shared by the step calls of one run.
struct StepCtx {
    Dims d;
    long long src = -1;  // linear cell, -1 = none
    int src_z = -1, src_y = -1, src_x = -1;
    bool tiled = false;  // FP32 on the LDS z-queue kernel
    ZqShape shape{};      // forward steps
    ZqShape adj_shape{};  // adjoint steps
    ZqShape rec_shape{};  // forward steps recomputed beside the adjoint steps
    ZqShape pair_shape{};  // the backward pass's step pairs (recompute + adjoint in one launch)
};

StepCtx step_ctx(const Dims& d, long long src) {
    StepCtx c;
    c.d = d;
    c.src = src;
    if (src >= 0) {
        const long long plane = d.nx * d.ny;
        c.src_z = int(src / plane);
        c.src_y = int((src % plane) / d.nx);
        c.src_x = int(src % d.nx);
    }
    // The FWI tuner's shapes, else the wave step's, else defaults.
    ZqShape tuned;
    if (zq_cached_shape(d, kKindFwiForward, &tuned) || zq_cached_shape(d, 0, &tuned)) c.shape = tuned;
    if (zq_cached_shape(d, kKindFwiAdjoint, &tuned) || zq_cached_shape(d, 0, &tuned)) c.adj_shape = tuned;
    c.rec_shape = zq_cached_shape(d, kKindFwiRecompute, &tuned) ? tuned : c.shape;
    if (zq_cached_shape(d, kKindFwiPair, &tuned)) c.pair_shape = tuned;
    return c;
}

bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

struct Plan {
    Dims d;
    size_t cells = 0, plane = 0;
    int64_t n = 0, C = 0, S = 0, banks = 1;
    long long src = 0;
    int rec_z = 0;
    // workspace offsets (bytes)
    size_t ring = 0, lam = 0, ckpt = 0, lap = 0, traces = 0, partial = 0, misfit = 0, total = 0;
    size_t field = 0;  // bytes from one field of the workspace to the next
};

bool make_plan(const stencil_grid* g, const stencil_fwi_config* c, size_t esz, Plan* p) {
    if (!g || !c || g->nx <= 8 || g->ny <= 8 || g->nz <= 8 || c->n_steps == 0) return false;
    if (g->nx > (1u << 20) || g->ny > (1u << 20) || g->nz > (1u << 20)) return false;
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0) || !(g->dt > 0)) return false;
    if (c->n_steps > (1u << 30) || c->variant > STENCIL_VARIANT_MULTIGPU) return false;
    auto inside = [](uint64_t v, uint64_t n) { return v >= 4 && v + 4 < n; };
    if (!inside(c->src_x, g->nx) || !inside(c->src_y, g->ny) || !inside(c->src_z, g->nz)) return false;
    if (!inside(c->rec_z, g->nz)) return false;
    p->d.nx = int64_t(g->nx);
    p->d.ny = int64_t(g->ny);
    p->d.nz = int64_t(g->nz);
    p->plane = size_t(g->nx) * size_t(g->ny);
    p->cells = p->plane * size_t(g->nz);
    p->n = int64_t(c->n_steps);
    p->C = c->ckpt_every ? int64_t(c->ckpt_every) : int64_t(std::ceil(std::sqrt(double(p->n))));
    if (p->C < 1) p->C = 1;
    if (p->C > p->n) p->C = p->n;
    p->S = (p->n + p->C - 1) / p->C;
    p->src = (long long)((c->src_z * g->ny + c->src_y) * g->nx + c->src_x);
    p->rec_z = int(c->rec_z);
    const size_t field = up(p->cells * esz);
    p->field = field;
    size_t o = 0;
    p->ring = o;
    o += 3 * field;
    p->lam = o;
    o += 3 * field;
    p->ckpt = o;
    o += size_t(2 * p->S) * field;
    // With more than one segment, two banks of C Laplacian fields: the next
    // segment's recompute fills one while this segment's adjoint reads the other.
    p->banks = p->S >= 2 ? 2 : 1;
    p->lap = o;
    o += size_t(p->banks * p->C) * field;
    p->traces = o;
    o += up(size_t(p->n) * p->plane * esz);
    p->partial = o;
    o += up(kPartials * sizeof(double));
    p->misfit = o;
    o += up(sizeof(double));
    p->total = o;
    return true;
}

template <class T>
struct Ops;
// FP32 steps run on the tiled kernel when it takes the dims and pointers
// (bit-identical to the global-memory kernels), else on the latter.
template <>
struct Ops<float> {
    // rec (nullable): plane rec_z of p_next into this nx*ny plane (halo pre-zeroed)
    static hipError_t fwd(const float* p, const float* pp, const float* m, float* pn, float* lap, const StepCtx& c,
                          const Scalars<float>& s, const float* w, int n, float* rec, int rec_z, hipStream_t st) {
        const void* ptrs[4] = {p, pp, m, pn};
        if (c.tiled && zq_supported(c.d, ptrs, 4) && (!lap || aligned16(lap)) && (!rec || aligned16(rec)))
            return wave_zq_fwi_f32(p, pp, m, pn, lap, c.src_z, c.src_y, c.src_x, c.src >= 0 ? w + n : nullptr, rec,
                                   rec_z, c.d, s, c.shape, st);
        const hipError_t e = fwi_fwd_step_f32(p, pp, m, pn, lap, c.d, s, c.src, w, n, st);
        if (e != hipSuccess || !rec) return e;
        const size_t plane = size_t(c.d.nx) * size_t(c.d.ny);
        return hipMemcpyAsync(rec, pn + size_t(rec_z) * plane, plane * sizeof(float), hipMemcpyDeviceToDevice, st);
    }
    // lap2 (nullable, with lap): L p^{k+1}, step k+1's gradient term added first
    static hipError_t adj(const float* l1, const float* l2, const float* m, const float* lap, const float* lap2,
                          float* g, float* lo, const float* res, int rz, const StepCtx& c, const Scalars<float>& s,
                          float* tr, hipStream_t st) {
        const void* ptrs[4] = {l1, l2, m, lo};
        if (c.tiled && zq_supported(c.d, ptrs, 4) && (!lap || (aligned16(lap) && aligned16(g))) &&
            (!lap2 || aligned16(lap2)) && (!res || aligned16(res)))
            return adjoint_zq_f32(l1, l2, m, lap, lap2, g, lo, res, rz, c.src_z, c.src_y, c.src_x, tr, c.d, s,
                                  c.adj_shape, st);
        return fwi_adj_step_f32(l1, l2, m, lap, lap2, g, lo, res, rz, c.d, s, tr ? c.src : -1LL, tr, st);
    }
    // the step pair: adj(l1 .. tr) and fwd(p, pp -> pn, L p into lap_out,
    // wavelet[n] at the source) in one launch on the tiled kernel
    static hipError_t pair(const float* l1, const float* l2, const float* m, const float* lap, const float* lap2,
                           float* g, float* lo, const float* res, int rz, float* tr, const float* p, const float* pp,
                           float* pn, float* lap_out, const float* w, int n, const StepCtx& c,
                           const Scalars<float>& s, hipStream_t st) {
        return pair_zq_f32(l1, l2, m, lap, lap2, g, lo, res, rz, c.src_z, c.src_y, c.src_x, tr, p, pp, pn, lap_out,
                           c.src >= 0 ? w + n : nullptr, c.d, s, c.pair_shape, st);
    }
    static hipError_t residual(const float* a, const float* b, float* r, size_t n, double* part, double* out,
                               hipStream_t st) {
        return fwi_residual_f32(a, b, r, n, part, out, st);
    }
    static hipError_t fill(float* p, size_t n, hipStream_t st) { return fill_f32(p, n, 0.f, st); }
    static hipError_t divide(float* g, const float* m, const Dims& d, hipStream_t st) {
        return fwi_divide_f32(g, m, d, st);
    }
};
template <>
struct Ops<double> {
    static hipError_t fwd(const double* p, const double* pp, const double* m, double* pn, double* lap,
                          const StepCtx& c, const Scalars<double>& s, const double* w, int n, double* rec, int rec_z,
                          hipStream_t st) {
        const hipError_t e = fwi_fwd_step_f64(p, pp, m, pn, lap, c.d, s, c.src, w, n, st);
        if (e != hipSuccess || !rec) return e;
        const size_t plane = size_t(c.d.nx) * size_t(c.d.ny);
        return hipMemcpyAsync(rec, pn + size_t(rec_z) * plane, plane * sizeof(double), hipMemcpyDeviceToDevice, st);
    }
    static hipError_t adj(const double* l1, const double* l2, const double* m, const double* lap,
                          const double* lap2, double* g, double* lo, const double* res, int rz, const StepCtx& c,
                          const Scalars<double>& s, double* tr, hipStream_t st) {
        return fwi_adj_step_f64(l1, l2, m, lap, lap2, g, lo, res, rz, c.d, s, tr ? c.src : -1LL, tr, st);
    }
    // FP64 runs the global-memory kernels: no step pair (pair_ok is false)
    static hipError_t pair(const double*, const double*, const double*, const double*, const double*, double*,
                           double*, const double*, int, double*, const double*, const double*, double*, double*,
                           const double*, int, const StepCtx&, const Scalars<double>&, hipStream_t) {
        return hipErrorInvalidValue;
    }
    static hipError_t residual(const double* a, const double* b, double* r, size_t n, double* part, double* out,
                               hipStream_t st) {
        return fwi_residual_f64(a, b, r, n, part, out, st);
    }
    static hipError_t fill(double* p, size_t n, hipStream_t st) { return fill_f64(p, n, 0.0, st); }
    static hipError_t divide(double* g, const double* m, const Dims& d, hipStream_t st) {
        return fwi_divide_f64(g, m, d, st);
    }
};

#define FWI_TRY(x)                                      \
    do {                                                \
        const hipError_t e_ = (x);                      \
        if (e_ != hipSuccess) return STENCIL_ERR_HIP;   \
    } while (0)

template <class T>
T* at(void* work, size_t off) {
    return reinterpret_cast<T*>(static_cast<char*>(work) + off);
}

template <class T>
Scalars<T> scalars(const stencil_grid* g) {
    return make_scalars<T>(T(g->dx), T(g->dy), T(g->dz), T(g->dt));
}

template <class T>
int forward(const T* m, const stencil_grid* g, const stencil_fwi_config* c, const T* wavelet, const T* d_obs,
            T* traces_out, double* misfit_out, void* work, hipStream_t st, bool tiled) {
    Plan p;
    if (!m || !wavelet || !work || !make_plan(g, c, sizeof(T), &p)) return STENCIL_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(work) % kAlign != 0) return STENCIL_ERR_INVALID;
    const Scalars<T> s = scalars<T>(g);
    StepCtx cx = step_ctx(p.d, p.src);
    cx.tiled = tiled;
    const size_t field = p.field;
    T* ring[3] = {at<T>(work, p.ring), at<T>(work, p.ring + field), at<T>(work, p.ring + 2 * field)};
    T* traces = at<T>(work, p.traces);
    FWI_TRY(Ops<T>::fill(ring[0], p.cells, st));
    FWI_TRY(Ops<T>::fill(ring[1], p.cells, st));
    FWI_TRY(Ops<T>::fill(ring[2], p.cells, st));
    FWI_TRY(Ops<T>::fill(traces, size_t(p.n) * p.plane, st));  // the steps write the plane interiors
    // One segment (ckpt_every = n_steps): the workspace holds L p^k of every
    // step, so a misfit forward writes them now and the backward needs no
    // recompute (lap slot k-1 holds L p^k, as the segment recompute fills it).
    const bool keep_lap = p.C == p.n && d_obs != nullptr;
    int prev = 0, cur = 1, next = 2;
    for (int64_t n = 0; n < p.n; ++n) {
        T* lap = keep_lap && n >= 1 ? at<T>(work, p.lap + size_t(n - 1) * field) : nullptr;
        FWI_TRY(Ops<T>::fwd(ring[cur], ring[prev], m, ring[next], lap, cx, s, wavelet, int(n),
                            traces + size_t(n) * p.plane, p.rec_z, st));
        if (n % p.C == 0) {
            T* ck = at<T>(work, p.ckpt + size_t(2 * (n / p.C)) * field);
            FWI_TRY(hipMemcpyAsync(ck, ring[cur], p.cells * sizeof(T), hipMemcpyDeviceToDevice, st));
            FWI_TRY(hipMemcpyAsync(reinterpret_cast<char*>(ck) + field, ring[next], p.cells * sizeof(T),
                                   hipMemcpyDeviceToDevice, st));
        }
        const int t = prev;
        prev = cur;
        cur = next;
        next = t;
    }
    const size_t nt = size_t(p.n) * p.plane;
    if (traces_out)
        FWI_TRY(hipMemcpyAsync(traces_out, traces, nt * sizeof(T), hipMemcpyDeviceToDevice, st));
    if (d_obs) {
        double* mis = at<double>(work, p.misfit);
        FWI_TRY(Ops<T>::residual(traces, d_obs, traces, nt, at<double>(work, p.partial), mis, st));
        if (misfit_out)
            FWI_TRY(hipMemcpyAsync(misfit_out, mis, sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    return STENCIL_OK;
}

// How the backward pass of the tiled FP32 path runs the recompute of segment
// s-1 beside the adjoint steps of segment s (more than one segment):
//   overlap  the recompute on a second stream, joined once per segment (the
//            default: the fastest at 256^3);
//   pair     one launch per step pair (a recompute step and an adjoint step on
//            the same tiles, m read once: 40 instead of 44 B per point; its two
//            register queues take 207-235 VGPRs, two waves per SIMD, and it
//            runs 5-10 % slower than the overlap);
//   serial   each segment's recompute, then its adjoint steps, one stream.
// Read from STENCIL_FWI_SCHEDULE on every backward call (STENCIL_FWI_OVERLAP=0,
// the older switch, means serial).  All three give the same bits.
enum class Schedule { kSerial, kOverlap, kPair };

Schedule fwi_schedule() {
    const char* e = std::getenv("STENCIL_FWI_SCHEDULE");
    if (e && std::strcmp(e, "serial") == 0) return Schedule::kSerial;
    if (e && std::strcmp(e, "overlap") == 0) return Schedule::kOverlap;
    if (e && std::strcmp(e, "pair") == 0) return Schedule::kPair;
    const char* o = std::getenv("STENCIL_FWI_OVERLAP");
    if (o && o[0] == '0') return Schedule::kSerial;
    return Schedule::kOverlap;
}

// Whether the step pair can take this backward call: FP32 on the tiled kernel
// with dims and buffers it accepts (the workspace's fields are 256-byte
// aligned offsets of an aligned base, and its receiver planes 16-byte
// aligned since nx % 4 == 0).
template <class T>
bool pair_ok(bool tiled, const Plan& p, const T* m, const T* grad, const void* work) {
    if (!std::is_same<T, float>::value || !tiled) return false;
    const void* ptrs[3] = {m, grad, work};
    return zq_supported(p.d, ptrs, 3);
}

// The recompute stream of the current device: created once (default
// priority, non-blocking) and kept for the process, so a backward call
// neither creates nor destroys streams; calls on one device share it.
hipError_t side_stream(hipStream_t* out) {
    static std::mutex mu;
    static std::map<int, hipStream_t> streams;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(mu);
    auto it = streams.find(dev);
    if (it == streams.end()) {
        hipStream_t st = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) return e;
        it = streams.emplace(dev, st).first;
    }
    *out = it->second;
    return hipSuccess;
}

// Events ordering the recompute stream against the caller's, destroyed when
// the backward call returns (the runtime frees them once their work is done).
struct Events {
    std::vector<hipEvent_t> ev;
    ~Events() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    hipError_t init(size_t n) {
        ev.reserve(n);
        for (size_t i = 0; i < n; ++i) {
            hipEvent_t x = nullptr;
            const hipError_t e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
            if (e != hipSuccess) return e;
            ev.push_back(x);
        }
        return hipSuccess;
    }
};

template <class T>
int backward(const T* m, const stencil_grid* g, const stencil_fwi_config* c, const T* wavelet, T* grad,
             T* src_trace, void* work, hipStream_t st, bool tiled) {
    Plan p;
    if (!m || !wavelet || !grad || !work || !make_plan(g, c, sizeof(T), &p)) return STENCIL_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(work) % kAlign != 0) return STENCIL_ERR_INVALID;
    const Scalars<T> s = scalars<T>(g);
    StepCtx cx = step_ctx(p.d, p.src);
    cx.tiled = tiled;
    const size_t field = p.field;
    T* ring[3] = {at<T>(work, p.ring), at<T>(work, p.ring + field), at<T>(work, p.ring + 2 * field)};
    T* lam[3] = {at<T>(work, p.lam), at<T>(work, p.lam + field), at<T>(work, p.lam + 2 * field)};
    const T* res = at<T>(work, p.traces);
    FWI_TRY(Ops<T>::fill(grad, p.cells, st));
    for (auto* l : lam) FWI_TRY(Ops<T>::fill(l, p.cells, st));
    // L p^k of segment sg, k = sg*C + 1 + i, in bank sg % banks
    auto lap_slot = [&](int64_t sg, int64_t i) {
        return at<T>(work, p.lap + size_t((sg % p.banks) * p.C + i) * field);
    };
    auto k_lo = [&](int64_t sg) { return sg * p.C + 1; };
    auto k_hi = [&](int64_t sg) { return std::min((sg + 1) * p.C, p.n); };
    // Re-run the forward steps of segment sg from its checkpoint, writing
    // L p^k; L p^N would only meet lambda^{N+1} = 0, so the last segment
    // stops at step N-1 (the wavelet has N samples).
    StepCtx rcx = cx;
    rcx.shape = cx.rec_shape;
    auto rec_last = [&](int64_t sg) { return std::min(k_hi(sg), p.n - 1); };
    int rp = 0, rc = 1, rn = 2;  // p^{j-1}, p^j, p^{j+1} of the segment being recomputed
    auto rec_start = [&](int64_t sg, hipStream_t rs) -> hipError_t {
        const T* ck = at<T>(work, p.ckpt + size_t(2 * sg) * field);
        rp = 0;
        rc = 1;
        rn = 2;
        hipError_t e = hipMemcpyAsync(ring[0], ck, p.cells * sizeof(T), hipMemcpyDeviceToDevice, rs);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ring[1], reinterpret_cast<const char*>(ck) + field, p.cells * sizeof(T),
                               hipMemcpyDeviceToDevice, rs);
        return e;
    };
    auto rec_rotate = [&] {
        const int t = rp;
        rp = rc;
        rc = rn;
        rn = t;
    };
    auto rec_step = [&](int64_t sg, int64_t j, hipStream_t rs) -> hipError_t {
        const hipError_t e = Ops<T>::fwd(ring[rc], ring[rp], m, ring[rn], lap_slot(sg, j - k_lo(sg)), rcx, s, wavelet,
                                         int(j), nullptr, 0, rs);
        rec_rotate();
        return e;
    };
    auto recompute = [&](int64_t sg, hipStream_t rs) -> hipError_t {
        hipError_t e = rec_start(sg, rs);
        for (int64_t j = k_lo(sg); e == hipSuccess && j <= rec_last(sg); ++j) e = rec_step(sg, j, rs);
        return e;
    };
    // The adjoint steps k = k_hi .. k_lo of segment sg.  Gradient terms
    // T_k = dt^2 (L p^k) psi^{k+1} exist for k in [a, b] = [k_lo, min(k_hi,
    // N-1)] (L p^N would meet lambda^{N+1} = 0).  Counting from the top of the
    // segment, every other step defers its term to the step below, which adds
    // T_{k+1} (it holds psi^{k+2} as its psi2) and then T_k: one gradient read
    // and write per two steps, the same additions in the same order, and
    // never a pair across segments (whose L p live in different banks).
    int l1 = 0, l2 = 1, lo = 2;  // psi^{k+1}, psi^{k+2}, psi^k (psi = m lambda)
    auto adj_terms = [&](int64_t sg, int64_t k, const T** lap, const T** lap2) {
        const int64_t a = k_lo(sg), b = std::min(k_hi(sg), p.n - 1);
        *lap = nullptr;
        *lap2 = nullptr;
        if (k <= b) {
            const bool defer = (b - k) % 2 == 0 && k > a;
            if (!defer) *lap = lap_slot(sg, k - a);
            if ((b - k) % 2 == 1) *lap2 = lap_slot(sg, k + 1 - a);
        }
    };
    auto adj_rotate = [&] {
        const int t = l2;
        l2 = l1;
        l1 = lo;
        lo = t;
    };
    auto adj_step = [&](int64_t sg, int64_t k) -> hipError_t {
        const T* lap = nullptr;
        const T* lap2 = nullptr;
        adj_terms(sg, k, &lap, &lap2);
        const hipError_t e = Ops<T>::adj(lam[l1], lam[l2], m, lap, lap2, grad, lam[lo], res + size_t(k - 1) * p.plane,
                                         p.rec_z, cx, s, src_trace ? src_trace + (k - 1) : nullptr, st);
        adj_rotate();
        return e;
    };
    auto adjoint = [&](int64_t sg) -> hipError_t {
        hipError_t e = hipSuccess;
        for (int64_t k = k_hi(sg); e == hipSuccess && k >= k_lo(sg); --k) e = adj_step(sg, k);
        return e;
    };
    Schedule sched = fwi_schedule();
    if (sched == Schedule::kPair && !pair_ok<T>(tiled, p, m, grad, work)) sched = Schedule::kOverlap;
    if (p.S == 1) {
        // one segment: the misfit forward wrote every L p^k (bank 0)
        FWI_TRY(adjoint(0));
    } else if (sched == Schedule::kSerial) {
        FWI_TRY(Ops<T>::fill(ring[2], p.cells, st));  // the zero walls, as below
        for (int64_t sg = p.S - 1; sg >= 0; --sg) {
            FWI_TRY(recompute(sg, st));
            FWI_TRY(adjoint(sg));
        }
    } else if (sched == Schedule::kPair) {
        // One stream.  The last segment is recomputed alone; then adjoint step
        // i of segment sg runs in one launch with recompute step i of segment
        // sg-1 (bank (sg-1) % 2, last read by the adjoint steps of sg+1, which
        // are done), the longer of the two finishing with single steps; the
        // adjoint steps of segment 0 run alone.
        FWI_TRY(Ops<T>::fill(ring[2], p.cells, st));  // the zero walls, as below
        FWI_TRY(recompute(p.S - 1, st));
        for (int64_t sg = p.S - 1; sg >= 1; --sg) {
            const int64_t sr = sg - 1;
            FWI_TRY(rec_start(sr, st));
            int64_t k = k_hi(sg), j = k_lo(sr);
            for (; k >= k_lo(sg) && j <= rec_last(sr); --k, ++j) {
                const T* lap = nullptr;
                const T* lap2 = nullptr;
                adj_terms(sg, k, &lap, &lap2);
                FWI_TRY(Ops<T>::pair(lam[l1], lam[l2], m, lap, lap2, grad, lam[lo], res + size_t(k - 1) * p.plane,
                                     p.rec_z, src_trace ? src_trace + (k - 1) : nullptr, ring[rc], ring[rp],
                                     ring[rn], lap_slot(sr, j - k_lo(sr)), wavelet, int(j), cx, s, st));
                adj_rotate();
                rec_rotate();
            }
            for (; k >= k_lo(sg); --k) FWI_TRY(adj_step(sg, k));
            for (; j <= rec_last(sr); ++j) FWI_TRY(rec_step(sr, j, st));
        }
        FWI_TRY(adjoint(0));
    } else {
        // Segments last to first; the recompute of segment sg-1 runs on a
        // second stream while the adjoint steps of segment sg run on the
        // caller's (the two read and write disjoint buffers; at small grids
        // neither kernel fills the GPU alone).  Events, each recorded once:
        // start, then "recomputed" and "adjoint done" per segment.
        hipStream_t rs = nullptr;
        FWI_TRY(side_stream(&rs));
        Events evs;
        FWI_TRY(evs.init(size_t(1 + 2 * p.S)));
        hipEvent_t start = evs.ev[0];
        auto recomputed = [&](int64_t sg) { return evs.ev[size_t(1 + sg)]; };
        auto adjoint_done = [&](int64_t sg) { return evs.ev[size_t(1 + p.S + sg)]; };
        // The steps write interiors only: the third level's halo (the zero
        // walls) is cleared here rather than left to the buffer's last user.
        FWI_TRY(Ops<T>::fill(ring[2], p.cells, st));
        FWI_TRY(hipEventRecord(start, st));
        FWI_TRY(hipStreamWaitEvent(rs, start, 0));
        FWI_TRY(recompute(p.S - 1, rs));
        FWI_TRY(hipEventRecord(recomputed(p.S - 1), rs));
        for (int64_t sg = p.S - 1; sg >= 0; --sg) {
            if (sg >= 1) {
                // bank (sg-1) % 2 was last read by the adjoint steps of sg+1
                if (sg + 1 <= p.S - 1) FWI_TRY(hipStreamWaitEvent(rs, adjoint_done(sg + 1), 0));
                FWI_TRY(recompute(sg - 1, rs));
                FWI_TRY(hipEventRecord(recomputed(sg - 1), rs));
            }
            FWI_TRY(hipStreamWaitEvent(st, recomputed(sg), 0));
            FWI_TRY(adjoint(sg));
            FWI_TRY(hipEventRecord(adjoint_done(sg), st));
        }
    }
    // grad held sum_k dt^2 (L p^k) psi^{k+1}; lambda = psi / m
    FWI_TRY(Ops<T>::divide(grad, m, p.d, st));
    return STENCIL_OK;
}

void put_shape(const ZqShape& z, stencil_tile_shape* out) {
    if (!out) return;
    out->tile_x4 = z.txv;
    out->tile_y = z.ty;
    out->n_chunks = z.n_chunks;
    out->xcd_remap = z.xcd_remap;
    out->nontemporal = z.nt;
}

// Launch-shape sweep of the FP32 FWI kernels on the workspace's wavefield,
// adjoint and (with more than one segment) Laplacian buffers, over the
// tuner's shapes and both streaming modes (cached accesses win where a
// step's fields fit the memory-side cache).  Each timed chain rotates three
// buffers as the leapfrog loops do, so
// the sweep-direction alternation between consecutive steps is part of what
// is timed: the forward step recording the receiver plane (the forward pass),
// the adjoint step alone, and the forward step writing L p while the fastest
// adjoint step of either streaming mode runs beside it on the recompute
// stream (the overlapped backward pass, which takes the best such pair); with
// more than one segment, also the step pair (both steps in one launch) over
// the shapes it is built for.
int tune(const stencil_grid* g, const stencil_fwi_config* c, void* work, int reps, stencil_tile_shape* best_fwd,
         stencil_tile_shape* best_adj, stencil_tile_shape* best_rec, const char* csv_path, hipStream_t st) {
    Plan p;
    if (!work || reps <= 0 || !make_plan(g, c, sizeof(float), &p)) return STENCIL_ERR_INVALID;
    if (reinterpret_cast<uintptr_t>(work) % kAlign != 0) return STENCIL_ERR_INVALID;
    const size_t field = p.field;
    // Scratch fields: 3 wavefield and 3 adjoint levels, then the Laplacian
    // slots, which hold nothing between calls when there are two banks.
    std::vector<float*> f;
    for (int i = 0; i < 3; ++i) f.push_back(at<float>(work, p.ring + size_t(i) * field));
    for (int i = 0; i < 3; ++i) f.push_back(at<float>(work, p.lam + size_t(i) * field));
    if (p.banks == 2)
        for (int64_t i = 0; i < std::min<int64_t>(p.banks * p.C, 5); ++i)
            f.push_back(at<float>(work, p.lap + size_t(i) * field));
    const bool pair = f.size() >= 10;
    // the adjoint's second L p field: a field of its own where there is one
    const float* lap2_solo = f.size() >= 7 ? f[6] : f[0];
    const float* lap2_pair = f.size() >= 11 ? f[10] : f[0];
    float* trace = at<float>(work, p.partial);  // reduction scratch
    const void* ptrs[4] = {f[0], f[1], f[2], f[3]};
    if (!zq_supported(p.d, ptrs, 4)) return STENCIL_ERR_UNSUPPORTED;
    float* m = f[0];  // (2000 m/s)^2; every other field ones
    FWI_TRY(fill_f32(m, p.cells, 4.0e6f, st));
    for (size_t i = 1; i < f.size(); ++i) FWI_TRY(fill_f32(f[i], p.cells, 1.0f, st));
    const Scalars<float> s = scalars<float>(g);
    const StepCtx cx = step_ctx(p.d, p.src);
    hipStream_t rs = nullptr;
    if (pair) FWI_TRY(side_stream(&rs));
    // step i of a chain over b[0..2]: prev b[i % 3], cur b[(i+1) % 3], next b[(i+2) % 3]
    struct Chain {
        float* b[3];
        int i = 0;
        float* prev() const { return b[i % 3]; }
        float* cur() const { return b[(i + 1) % 3]; }
        float* next() const { return b[(i + 2) % 3]; }
    };
    auto fwd = [&](Chain& ch, const ZqShape& sh, float* lap, float* rec, hipStream_t q) {
        const hipError_t e = wave_zq_fwi_f32(ch.cur(), ch.prev(), m, ch.next(), lap, cx.src_z, cx.src_y, cx.src_x, m,
                                             rec, p.rec_z, p.d, s, sh, q);
        ++ch.i;
        return e;
    };
    // adjoint steps alternate as the backward pass runs them: one adding two
    // gradient terms (L p^k, L p^{k+1}), the next none
    auto adj = [&](Chain& ch, const ZqShape& sh, const float* lap, const float* lap2, float* grad,
                   const float* res) {
        const bool terms = ch.i % 2 == 0;
        const hipError_t e = adjoint_zq_f32(ch.cur(), ch.prev(), m, terms ? lap : nullptr, terms ? lap2 : nullptr,
                                            grad, ch.next(), res, p.rec_z, cx.src_z, cx.src_y, cx.src_x, trace, p.d,
                                            s, sh, st);
        ++ch.i;
        return e;
    };
    struct Row {
        ZqShape shape;
        float fwd_ms, adj_ms, rec_ms[2];  // rec_ms[k]: beside the best adjoint shape with nt == k
    };
    std::vector<Row> rows;
    for (const ZqShape& sh : zq_tune_candidates(p.d)) {
        Row r{sh, -1.f, -1.f, {-1.f, -1.f}};
        Chain a{{f[1], f[2], f[3]}};
        r.fwd_ms = zq_time_launches([&] { return fwd(a, sh, nullptr, f[5], st); }, reps, st);
        // the adjoint runs 64 x 4 and 64 x 16 as 64 x 8: not timed twice
        if (!(sh.txv == 64 && (sh.ty == 4 || sh.ty == 16))) {
            Chain b{{f[1], f[2], f[3]}};
            r.adj_ms = zq_time_launches([&] { return adj(b, sh, f[4], lap2_solo, f[5], f[4]); }, reps, st);
        }
        rows.push_back(r);
    }
    const Row* bf = nullptr;
    const Row* ba = nullptr;
    const Row* ba_nt[2] = {nullptr, nullptr};  // the fastest adjoint in each streaming mode
    for (const Row& r : rows) {
        if (r.fwd_ms > 0.f && (!bf || r.fwd_ms < bf->fwd_ms)) bf = &r;
        if (r.adj_ms > 0.f && (!ba || r.adj_ms < ba->adj_ms)) ba = &r;
        const int k = r.shape.nt ? 1 : 0;
        if (r.adj_ms > 0.f && (!ba_nt[k] || r.adj_ms < ba_nt[k]->adj_ms)) ba_nt[k] = &r;
    }
    if (!bf || !ba) return STENCIL_ERR_HIP;
    ZqShape adj_best = ba->shape;  // alone; with the overlap, the best pair's
    ZqShape rec_best = bf->shape;
    bool have_rec = false;
    struct Beside {
        ZqShape adj, rec;
        float ms;
    };
    std::vector<Beside> refined;
    if (pair) {
        // recompute chain f[1..3] (L p into f[4]) on the recompute stream, the
        // adjoint chain f[6..8] (L p from f[5] and lap2_pair, gradient f[9])
        // on st, joined every kSteps steps (the backward pass joins once per
        // segment); a fork and a join event per timed call, each recorded once.
        constexpr int kSteps = 8;
        auto beside = [&](const ZqShape& adj_sh, const ZqShape& rec_sh, int n) -> float {
            Chain a{{f[1], f[2], f[3]}};
            Chain b{{f[6], f[7], f[8]}};
            Events evs;
            if (evs.init(size_t(2 * (n + 2))) != hipSuccess) return -1.f;
            size_t call = 0;
            const float ms = zq_time_launches(
                [&] {
                    if (2 * call + 1 >= evs.ev.size()) return hipErrorInvalidValue;  // more calls than events
                    hipEvent_t fork = evs.ev[2 * call], join = evs.ev[2 * call + 1];
                    ++call;
                    hipError_t e = hipEventRecord(fork, st);
                    if (e == hipSuccess) e = hipStreamWaitEvent(rs, fork, 0);
                    for (int j = 0; e == hipSuccess && j < kSteps; ++j) e = fwd(a, rec_sh, f[4], nullptr, rs);
                    if (e == hipSuccess) e = hipEventRecord(join, rs);
                    for (int j = 0; e == hipSuccess && j < kSteps; ++j) e = adj(b, adj_sh, f[5], lap2_pair, f[9], f[5]);
                    if (e == hipSuccess) e = hipStreamWaitEvent(st, join, 0);
                    return e;
                },
                n, st);
            return ms > 0.f ? ms / kSteps : -1.f;
        };
        // Sweep: every recompute shape beside the fastest adjoint of each
        // streaming mode (a cached adjoint that wins alone can lose beside a
        // recompute competing for the same memory-side cache).
        for (int k = 0; k < 2; ++k) {
            if (!ba_nt[k]) continue;
            for (Row& r : rows) r.rec_ms[k] = beside(ba_nt[k]->shape, r.shape, reps);
        }
        // Refinement: the pairs that run beside each other also differ by a
        // few per cent of noise per timing, and the adjoint that is fastest
        // alone need not be fastest beside a recompute.  The four fastest
        // adjoint shapes alone (plus the best of each streaming mode) are
        // timed beside the four fastest recompute shapes of the sweep, twice
        // each with at least 10 timed calls, each pair keeping its best time.
        constexpr int kTopPair = 4;
        std::vector<const Row*> ta, tr;
        for (const Row& r : rows)
            if (r.adj_ms > 0.f) ta.push_back(&r);
        std::sort(ta.begin(), ta.end(), [](const Row* x, const Row* y) { return x->adj_ms < y->adj_ms; });
        if (ta.size() > size_t(kTopPair)) ta.resize(kTopPair);
        for (const Row* r : ba_nt)
            if (r && std::find(ta.begin(), ta.end(), r) == ta.end()) ta.push_back(r);
        auto rec_of = [](const Row* r) {
            const float a0 = r->rec_ms[0], a1 = r->rec_ms[1];
            return a0 > 0.f && (a1 <= 0.f || a0 < a1) ? a0 : a1;
        };
        for (const Row& r : rows)
            if (rec_of(&r) > 0.f) tr.push_back(&r);
        std::sort(tr.begin(), tr.end(), [&](const Row* x, const Row* y) { return rec_of(x) < rec_of(y); });
        if (tr.size() > size_t(kTopPair)) tr.resize(kTopPair);
        const int n = std::max(2 * reps, 10);
        for (const Row* x : ta)
            for (const Row* y : tr) refined.push_back({x->shape, y->shape, -1.f});
        for (int round = 0; round < 2; ++round)
            for (Beside& bs : refined) {
                const float ms = beside(bs.adj, bs.rec, n);
                if (ms > 0.f && (bs.ms <= 0.f || ms < bs.ms)) bs.ms = ms;
            }
        const Beside* best = nullptr;
        for (const Beside& bs : refined)
            if (bs.ms > 0.f && (!best || bs.ms < best->ms)) best = &bs;
        if (best) {
            adj_best = best->adj;
            rec_best = best->rec;
            have_rec = true;
        }
    }
    // The step pair (recompute chain f[1..3] writing L p into f[4], adjoint
    // chain f[6..8] as above) over the shapes it is built for, one launch per
    // step pair as the backward pass runs it.
    std::vector<std::pair<ZqShape, float>> pair_rows;
    if (pair) {
        for (const ZqShape& sh : zq_tune_candidates(p.d)) {
            if (!zq_pair_shape_valid(sh.txv, sh.ty)) continue;
            Chain a{{f[1], f[2], f[3]}};
            Chain b{{f[6], f[7], f[8]}};
            const float ms = zq_time_launches(
                [&] {
                    const bool terms = b.i % 2 == 0;
                    const hipError_t e = pair_zq_f32(
                        b.cur(), b.prev(), m, terms ? f[5] : nullptr, terms ? lap2_pair : nullptr, f[9], b.next(),
                        f[5], p.rec_z, cx.src_z, cx.src_y, cx.src_x, trace, a.cur(), a.prev(), a.next(), f[4], m, p.d,
                        s, sh, st);
                    ++a.i;
                    ++b.i;
                    return e;
                },
                reps, st);
            if (ms > 0.f) pair_rows.emplace_back(sh, ms);
        }
        const std::pair<ZqShape, float>* bp = nullptr;
        for (const auto& r : pair_rows)
            if (!bp || r.second < bp->second) bp = &r;
        if (bp) zq_cache_put(p.d, kKindFwiPair, bp->first);
    }
    zq_cache_put(p.d, kKindFwiForward, bf->shape);
    zq_cache_put(p.d, kKindFwiAdjoint, adj_best);
    put_shape(bf->shape, best_fwd);
    put_shape(adj_best, best_adj);
    if (have_rec) zq_cache_put(p.d, kKindFwiRecompute, rec_best);
    put_shape(rec_best, best_rec);
    if (csv_path && *csv_path) {
        FILE* out = std::fopen(csv_path, "w");
        if (!out) return STENCIL_ERR_INVALID;
        std::fprintf(out, "kernel,nx,ny,nz,tile_x,tile_y,n_chunks,xcd_remap,nontemporal,ms,adjoint_beside\n");
        static const char* const names[4] = {"forward", "adjoint", "recompute_beside_cached_adjoint",
                                             "recompute_beside_nontemporal_adjoint"};
        for (const Row& r : rows)
            for (int k = 0; k < 4; ++k) {
                const float ms = k == 0 ? r.fwd_ms : k == 1 ? r.adj_ms : r.rec_ms[k - 2];
                if (ms > 0.f)
                    std::fprintf(out, "%s,%lld,%lld,%lld,%d,%d,%d,%d,%d,%.5f,\n", names[k], (long long)p.d.nx,
                                 (long long)p.d.ny, (long long)p.d.nz, 4 * r.shape.txv, r.shape.ty, r.shape.n_chunks,
                                 r.shape.xcd_remap, r.shape.nt, ms);
            }
        for (const Beside& bs : refined)
            if (bs.ms > 0.f)
                std::fprintf(out, "recompute_beside_adjoint_refined,%lld,%lld,%lld,%d,%d,%d,%d,%d,%.5f,"
                             "%dx%d/%d/%d/%d\n",
                             (long long)p.d.nx, (long long)p.d.ny, (long long)p.d.nz, 4 * bs.rec.txv, bs.rec.ty,
                             bs.rec.n_chunks, bs.rec.xcd_remap, bs.rec.nt, bs.ms, 4 * bs.adj.txv, bs.adj.ty,
                             bs.adj.n_chunks, bs.adj.xcd_remap, bs.adj.nt);
        for (const auto& r : pair_rows)
            std::fprintf(out, "step_pair,%lld,%lld,%lld,%d,%d,%d,%d,%d,%.5f,\n", (long long)p.d.nx, (long long)p.d.ny,
                         (long long)p.d.nz, 4 * r.first.txv, r.first.ty, r.first.n_chunks, r.first.xcd_remap,
                         r.first.nt, r.second);
        std::fclose(out);
    }
    return STENCIL_OK;
}

}  // namespace

extern "C" {

uint64_t stencil_fwi_workspace_bytes(const stencil_grid* grid, const stencil_fwi_config* cfg, int bytes_per_value) {
    if (bytes_per_value != 4 && bytes_per_value != 8) return 0;
    Plan p;
    if (!make_plan(grid, cfg, size_t(bytes_per_value), &p)) return 0;
    return p.total;
}

int stencil_fwi_forward_f32(const float* vel2, const stencil_grid* grid, const stencil_fwi_config* cfg,
                            const float* wavelet, const float* d_obs, float* traces_out, double* misfit_out,
                            void* work, void* stream) {
    return forward<float>(vel2, grid, cfg, wavelet, d_obs, traces_out, misfit_out, work,
                          static_cast<hipStream_t>(stream), cfg && cfg->variant != STENCIL_VARIANT_NAIVE);
}
int stencil_fwi_forward_f64(const double* vel2, const stencil_grid* grid, const stencil_fwi_config* cfg,
                            const double* wavelet, const double* d_obs, double* traces_out, double* misfit_out,
                            void* work, void* stream) {
    return forward<double>(vel2, grid, cfg, wavelet, d_obs, traces_out, misfit_out, work,
                           static_cast<hipStream_t>(stream), false);
}
int stencil_fwi_backward_f32(const float* vel2, const stencil_grid* grid, const stencil_fwi_config* cfg,
                             const float* wavelet, float* grad, float* src_trace, void* work, void* stream) {
    return backward<float>(vel2, grid, cfg, wavelet, grad, src_trace, work, static_cast<hipStream_t>(stream),
                           cfg && cfg->variant != STENCIL_VARIANT_NAIVE);
}
int stencil_fwi_backward_f64(const double* vel2, const stencil_grid* grid, const stencil_fwi_config* cfg,
                             const double* wavelet, double* grad, double* src_trace, void* work, void* stream) {
    return backward<double>(vel2, grid, cfg, wavelet, grad, src_trace, work, static_cast<hipStream_t>(stream),
                            false);
}
int stencil_tune_fwi_f32(const stencil_grid* grid, const stencil_fwi_config* cfg, void* work, int reps,
                         stencil_tile_shape* best_forward, stencil_tile_shape* best_adjoint,
                         stencil_tile_shape* best_recompute, const char* csv_path, void* stream) {
    return tune(grid, cfg, work, reps, best_forward, best_adjoint, best_recompute, csv_path,
                static_cast<hipStream_t>(stream));
}

}  // extern "C"

namespace {
bool step_grid_ok(const stencil_grid* g) {
    return g && g->nx > 8 && g->ny > 8 && g->nz > 8 && g->nx <= (1u << 20) && g->ny <= (1u << 20) &&
           g->nz <= (1u << 20) && g->dx > 0 && g->dy > 0 && g->dz > 0 && g->dt > 0;
}
Dims dims_of_grid(const stencil_grid* g) {
    Dims d;
    d.nx = int64_t(g->nx);
    d.ny = int64_t(g->ny);
    d.nz = int64_t(g->nz);
    return d;
}
bool variant_ok(int v) { return v >= STENCIL_VARIANT_NAIVE && v <= STENCIL_VARIANT_MULTIGPU; }
template <class T>
int fwd_step_api(const T* p, const T* pp, const T* m, T* pn, T* lap_out, const stencil_grid* g, int64_t src,
                 const T* wavelet, int64_t n, int variant, void* stream) {
    if (!p || !pp || !m || !pn || !step_grid_ok(g) || !variant_ok(variant)) return STENCIL_ERR_INVALID;
    if (pn == p || pn == pp || pn == m || (lap_out && (lap_out == p || lap_out == pn))) return STENCIL_ERR_INVALID;
    const int64_t cells = int64_t(g->nx * g->ny * g->nz);
    if (src >= cells || (src >= 0 && (!wavelet || n < 0))) return STENCIL_ERR_INVALID;
    StepCtx cx = step_ctx(dims_of_grid(g), src >= 0 ? (long long)src : -1LL);
    cx.tiled = variant != STENCIL_VARIANT_NAIVE;
    return Ops<T>::fwd(p, pp, m, pn, lap_out, cx, scalars<T>(g), wavelet, int(n), nullptr, 0,
                       static_cast<hipStream_t>(stream)) ==
                   hipSuccess
               ? STENCIL_OK
               : STENCIL_ERR_HIP;
}
template <class T>
int adj_step_api(const T* l1, const T* l2, const T* m, const T* lap, T* grad, T* lo, const T* res, int64_t rec_z,
                 const stencil_grid* g, int64_t src, T* trace, int variant, void* stream) {
    if (!l1 || !l2 || !m || !lo || !step_grid_ok(g) || !variant_ok(variant)) return STENCIL_ERR_INVALID;
    if (lo == l1 || lo == l2 || lo == m || (lap && !grad)) return STENCIL_ERR_INVALID;
    if (res && (rec_z < 4 || rec_z + 4 >= int64_t(g->nz))) return STENCIL_ERR_INVALID;
    const int64_t cells = int64_t(g->nx * g->ny * g->nz);
    if (trace && (src < 0 || src >= cells)) return STENCIL_ERR_INVALID;
    StepCtx cx = step_ctx(dims_of_grid(g), trace ? (long long)src : -1LL);
    cx.tiled = variant != STENCIL_VARIANT_NAIVE;
    return Ops<T>::adj(l1, l2, m, lap, nullptr, grad, lo, res, int(rec_z), cx, scalars<T>(g), trace,
                       static_cast<hipStream_t>(stream)) == hipSuccess
               ? STENCIL_OK
               : STENCIL_ERR_HIP;
}
}  // namespace

extern "C" {
int stencil_forward_step_f32(const float* p_cur, const float* p_prev, const float* vel2, float* p_next,
                             float* lap_out, const stencil_grid* grid, int64_t src_index, const float* wavelet,
                             int64_t n, int variant, void* stream) {
    return fwd_step_api(p_cur, p_prev, vel2, p_next, lap_out, grid, src_index, wavelet, n, variant, stream);
}
int stencil_forward_step_f64(const double* p_cur, const double* p_prev, const double* vel2, double* p_next,
                             double* lap_out, const stencil_grid* grid, int64_t src_index, const double* wavelet,
                             int64_t n, int variant, void* stream) {
    return fwd_step_api(p_cur, p_prev, vel2, p_next, lap_out, grid, src_index, wavelet, n, variant, stream);
}
int stencil_adjoint_step_f32(const float* psi1, const float* psi2, const float* vel2, const float* lap, float* grad,
                             float* psi_out, const float* res_plane, int64_t rec_z, const stencil_grid* grid,
                             int64_t src_index, float* src_value, int variant, void* stream) {
    return adj_step_api(psi1, psi2, vel2, lap, grad, psi_out, res_plane, rec_z, grid, src_index, src_value, variant, stream);
}
int stencil_adjoint_step_f64(const double* psi1, const double* psi2, const double* vel2, const double* lap,
                             double* grad, double* psi_out, const double* res_plane, int64_t rec_z,
                             const stencil_grid* grid, int64_t src_index, double* src_value, int variant, void* stream) {
    return adj_step_api(psi1, psi2, vel2, lap, grad, psi_out, res_plane, rec_z, grid, src_index, src_value, variant, stream);
}
}  // extern "C"
