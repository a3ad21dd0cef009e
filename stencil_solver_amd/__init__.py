"""MI355X-native 8th-order 3-D stencil time-step (gfx950 HIP kernels).

The compute path is libstencil_hip.so (stencil_solver_amd/csrc): hand-written
HIP kernels behind the C-ABI in include/stencil_capi.h and the C++ API in
include/stencil/.  This package is the Python binding of that C-ABI.
"""
from .api import (  # noqa: F401
    STORAGE,
    Field16,
    Reducer,
    decode16,
    encode16,
    launch_tiled_16,
    tune_tiled_16,
    pow2_scale_for,
    VARIANTS,
    Config,
    GridDims,
    PerfResult,
    Solver,
    StencilError,
    heat_bytes_per_step,
    heat_step,
    host_gaussian,
    host_layered_noise_vel2,
    host_noise_cube,
    launch,
    launch_naive,
    launch_tiled,
    launch_tiled_ex,
    make_solver,
    tile_shape,
    tune_heat,
    tune_tiled,
    tuned_shape,
    set_tuned_shape,
    clear_tuning,
    upload_fd_coefficients,
    wave_bytes_per_step,
)
from ._lib import LIB_PATH, lib  # noqa: F401
from .batch import ShotBatch  # noqa: F401
from .fwi import FwiProblem, adjoint_step, forward_step, fwi_misfit, ricker  # noqa: F401
