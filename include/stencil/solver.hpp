#pragma once

/**
 * include/stencil/solver.hpp -- the solver API of the reference
 * (include/stencil/solver.hpp:13-24): ISolver{init, step, run, download,
 * dims, name} and make_solver.  Solvers own a DeviceGrid ring and a HIP stream
 * and call the launchers in include/stencil/kernels.hpp.
 *
 * Templated on the field type; `ISolver`/`make_solver` are the real_t
 * instantiation (float, or double under STENCIL_FP64).  The library provides
 * both.
 */

#include <algorithm>
#include <memory>
#include <string>

#include "stencil/config.hpp"
#include "stencil/grid.hpp"
#include "stencil/timer.hpp"
#include "stencil/types.hpp"

namespace stencil {

template <class T>
class BasicISolver {
public:
    virtual ~BasicISolver() = default;
    virtual void init(const BasicConfig<T>& cfg) = 0;
    virtual void step() = 0;
    virtual PerfResult run(int n_steps, int warmup = 0) = 0;
    virtual void download(BasicHostGrid<T>& out) const = 0;
    virtual const BasicGridDims<T>& dims() const = 0;
    virtual std::string name() const = 0;
    /// Addition: z-plane z of the newest level (nx*ny values, x fastest) into
    /// out, synchronised.  This default downloads the whole field; the
    /// library's solvers copy the one contiguous plane instead.
    virtual void download_plane_z(std::size_t z, T* out) const {
        const BasicGridDims<T>& d = dims();
        BasicHostGrid<T> h(d);
        download(h);
        std::copy_n(h.ptr() + d.idx(0, 0, z), d.nx * d.ny, out);
    }
};
using ISolver = BasicISolver<real_t>;

/// Physics -> solver factory, then init(cfg) (reference solver_base.cpp:119-133).
template <class T>
std::unique_ptr<BasicISolver<T>> make_basic_solver(const BasicConfig<T>& cfg);

inline std::unique_ptr<ISolver> make_solver(const Config& cfg) {
    return make_basic_solver<real_t>(cfg);
}

}  // namespace stencil
