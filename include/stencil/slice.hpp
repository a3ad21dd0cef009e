#pragma once

/**
 * include/stencil/slice.hpp -- z-plane snapshots of a run on disk.
 *
 * The on-disk format is the reference's slice format (src/main_slice.cpp:106-152),
 * which its scripts/make_slice_gif.py and scripts/reference_slice.py read:
 *   PREFIX.f32   float32 frames, frame after frame, each ny rows of nx values
 *                (x fastest);
 *   PREFIX.json  {nx, ny, nz, plane: "xy", plane_index, frames, every, steps,
 *                dt, dx, velocity, physics, kernel, dtype: "float32"}.
 * Frame 0 is the state before the first step, then one frame every `every`
 * steps.
 */

#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "stencil/config.hpp"
#include "stencil/solver.hpp"

namespace stencil {

/// Metadata of a slice recording (the JSON side of the format).
struct SliceMeta {
    std::size_t nx = 0, ny = 0, nz = 0;
    std::size_t plane_index = 0;
    int every = 1;
    int steps = 0;
    double dt = 0.0, dx = 0.0, velocity = 0.0;
    std::string physics = "wave";  ///< "wave" or "heat"
    std::string kernel = "naive";  ///< "naive", "tiled" or "multigpu"
};

/// Appends float32 frames of nx*ny values to PREFIX.f32 and writes
/// PREFIX.json when finished.  Throws std::runtime_error when a file cannot
/// be written.
class SliceWriter {
public:
    SliceWriter(std::string prefix, std::size_t nx, std::size_t ny);
    SliceWriter(const SliceWriter&) = delete;
    SliceWriter& operator=(const SliceWriter&) = delete;
    ~SliceWriter();

    void append(const float* frame);
    int frames() const noexcept { return frames_; }
    /// Closes the frame file and writes the metadata (frames counted here).
    void finish(SliceMeta meta);
    const std::string& prefix() const noexcept { return prefix_; }

private:
    std::string prefix_;
    std::size_t values_;
    std::FILE* bin_ = nullptr;
    int frames_ = 0;
};

/// Runs `steps` steps of `solver`, recording its middle z-plane (nz / 2)
/// before the first step and after every `every`-th step; returns the
/// metadata written.  cfg supplies velocity, physics and kernel names.
template <class T>
SliceMeta record_middle_plane(BasicISolver<T>& solver, const BasicConfig<T>& cfg, int steps, int every,
                              const std::string& prefix);

}  // namespace stencil
