"""Loader and ctypes signatures for libstencil_hip.so (include/stencil_capi.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C stencil_solver_amd/csrc``).  There is no fallback: if the library is
missing, every entry point raises.

torch is imported before the library is opened so that the process holds a
single HIP runtime (torch's), which makes torch device pointers and streams
valid arguments.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

try:  # one HIP runtime per process: torch's, when torch is present
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is part of the image
    torch = None

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("STENCIL_HIP_LIB", PKG_DIR / "lib" / "libstencil_hip.so"))


class stencil_grid(C.Structure):
    _fields_ = [("nx", C.c_uint64), ("ny", C.c_uint64), ("nz", C.c_uint64),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double)]


class stencil_config(C.Structure):
    _fields_ = [("nx", C.c_uint64), ("ny", C.c_uint64), ("nz", C.c_uint64),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("steps", C.c_int), ("warmup_steps", C.c_int),
                ("velocity", C.c_double), ("diffusivity", C.c_double),
                ("pulse_cx", C.c_double), ("pulse_cy", C.c_double), ("pulse_cz", C.c_double),
                ("pulse_sigma", C.c_double),
                ("physics", C.c_int), ("variant", C.c_int), ("device", C.c_int),
                ("vmodel", C.c_int), ("n_layers", C.c_int),
                ("vmin", C.c_double), ("vmax", C.c_double), ("vnoise", C.c_double),
                ("ic", C.c_int), ("cube_side", C.c_uint64), ("cube_value", C.c_double),
                ("seed", C.c_uint64), ("fp64", C.c_int)]


class stencil_tile_shape(C.Structure):
    _fields_ = [("tile_x4", C.c_int), ("tile_y", C.c_int), ("n_chunks", C.c_int), ("xcd_remap", C.c_int),
                ("nontemporal", C.c_int)]


class stencil_field16(C.Structure):
    _fields_ = [("data", C.c_void_p), ("exp", C.c_void_p)]


class stencil_fwi_config(C.Structure):
    _fields_ = [("n_steps", C.c_uint64), ("src_x", C.c_uint64), ("src_y", C.c_uint64), ("src_z", C.c_uint64),
                ("rec_z", C.c_uint64), ("ckpt_every", C.c_uint64), ("variant", C.c_uint64)]


class stencil_shot(C.Structure):
    _fields_ = [("grid", stencil_grid), ("ring", C.c_void_p * 3), ("vel2", C.c_void_p), ("wavelet", C.c_void_p),
                ("src_x", C.c_uint64), ("src_y", C.c_uint64), ("src_z", C.c_uint64)]


class stencil_perf(C.Structure):
    _fields_ = [("time_ms", C.c_double), ("bandwidth_gbs", C.c_double), ("gflops", C.c_double),
                ("model_bandwidth_gbs", C.c_double), ("grid_size", C.c_uint64),
                ("label", C.c_char * 32)]


_P = C.c_void_p
_G = C.POINTER(stencil_grid)
_I = C.c_int

# name -> (restype, argtypes); every symbol declared in include/stencil_capi.h.
SIGNATURES = {
    "stencil_strerror": (C.c_char_p, [_I]),
    "stencil_last_hip_error": (_I, []),
    "stencil_upload_fd_coefficients": (_I, []),
    "stencil_launch_naive_f32": (_I, [_P, _P, _P, _P, _G, _P]),
    "stencil_launch_tiled_f32": (_I, [_P, _P, _P, _P, _G, _P]),
    "stencil_launch_f32": (_I, [_I, _P, _P, _P, _P, _G, _P]),
    "stencil_launch_naive_f64": (_I, [_P, _P, _P, _P, _G, _P]),
    "stencil_launch_tiled_f64": (_I, [_P, _P, _P, _P, _G, _P]),
    "stencil_launch_f64": (_I, [_I, _P, _P, _P, _P, _G, _P]),
    "stencil_heat_step_f32": (_I, [_P, _P, C.c_float, _G, _I, _P]),
    "stencil_heat_step_f64": (_I, [_P, _P, C.c_double, _G, _I, _P]),
    "stencil_wave_bytes_per_step": (C.c_double, [_G, _I]),
    "stencil_heat_bytes_per_step": (C.c_double, [_G, _I]),
    "stencil_launch_tiled_ex_f32": (_I, [_P, _P, _P, _P, _G, _I, C.POINTER(stencil_tile_shape), _P]),
    "stencil_heat_step_ex_f32": (_I, [_P, _P, C.c_float, _G, _I, C.POINTER(stencil_tile_shape), _P]),
    "stencil_stream_f32": (_I, [_I, _P, _P, C.c_uint64, _P]),
    "stencil_tune_tiled_f32": (_I, [_P, _P, _P, _P, _G, _I, C.POINTER(stencil_tile_shape), C.c_char_p, _P]),
    "stencil_tune_heat_f32": (_I, [_P, _P, _G, _I, C.POINTER(stencil_tile_shape), C.c_char_p, _P]),
    "stencil_tuned_shape": (_I, [_G, _I, C.POINTER(stencil_tile_shape)]),
    "stencil_set_tuned_shape": (_I, [_G, _I, C.POINTER(stencil_tile_shape)]),
    "stencil_clear_tuning": (None, []),
    "stencil_field16_bytes": (_I, [_I, _G, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "stencil_encode16": (_I, [_I, _P, C.POINTER(stencil_field16), C.c_uint64, C.c_float, _P]),
    "stencil_decode16": (_I, [_I, C.POINTER(stencil_field16), _P, C.c_uint64, C.c_float, _P]),
    "stencil_launch_tiled_16": (_I, [_I] + [C.POINTER(stencil_field16)] * 4 + [_G, C.c_float, C.c_float, _I,
                                                                              C.POINTER(stencil_tile_shape), _P]),
    "stencil_tune_tiled_16": (_I, [_I] + [C.POINTER(stencil_field16)] * 4 + [_G, C.c_float, C.c_float, _I,
                                                                            C.POINTER(stencil_tile_shape), C.c_char_p, _P]),
    "stencil_reduce_workspace_bytes": (C.c_uint64, []),
    "stencil_energy_f32": (_I, [_P, _P, _P, _G, _P, _P, _P]),
    "stencil_energy_f32_noalloc": (_I, [_P, _P, _P, _G, _P, _P]),
    "stencil_energy_f64": (_I, [_P, _P, _P, _G, _P, _P, _P]),
    "stencil_diff_norms_f32_f64": (_I, [_P, _P, _G, _P, _P, _P]),
    "stencil_dot_f64": (_I, [_P, _P, _G, _P, _P, _P]),
    "stencil_dot_f32": (_I, [_P, _P, _G, _P, _P, _P]),
    "stencil_host_init_gaussian_f32": (_I, [_P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_float,
                                            C.c_float, C.c_float, C.c_float]),
    "stencil_host_layered_noise_vel2_f32": (_I, [_P, C.c_uint64, C.c_uint64, C.c_uint64, _I,
                                                 C.c_float, C.c_float, C.c_float, C.c_uint64]),
    "stencil_host_noise_cube_f32": (_I, [_P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                         C.c_float, C.c_uint64]),
    "stencil_fwi_workspace_bytes": (C.c_uint64, [_G, C.POINTER(stencil_fwi_config), _I]),
    "stencil_fwi_forward_f32": (_I, [_P, _G, C.POINTER(stencil_fwi_config), _P, _P, _P, _P, _P, _P]),
    "stencil_fwi_forward_f64": (_I, [_P, _G, C.POINTER(stencil_fwi_config), _P, _P, _P, _P, _P, _P]),
    "stencil_fwi_backward_f32": (_I, [_P, _G, C.POINTER(stencil_fwi_config), _P, _P, _P, _P, _P]),
    "stencil_fwi_backward_f64": (_I, [_P, _G, C.POINTER(stencil_fwi_config), _P, _P, _P, _P, _P]),
    "stencil_tune_fwi_f32": (_I, [_G, C.POINTER(stencil_fwi_config), _P, _I] + [C.POINTER(stencil_tile_shape)] * 3
                             + [C.c_char_p, _P]),
    "stencil_forward_step_f32": (_I, [_P, _P, _P, _P, _P, _G, C.c_int64, _P, C.c_int64, _I, _P]),
    "stencil_forward_step_f64": (_I, [_P, _P, _P, _P, _P, _G, C.c_int64, _P, C.c_int64, _I, _P]),
    "stencil_adjoint_step_f32": (_I, [_P, _P, _P, _P, _P, _P, _P, C.c_int64, _G, C.c_int64, _P, _I, _P]),
    "stencil_adjoint_step_f64": (_I, [_P, _P, _P, _P, _P, _P, _P, C.c_int64, _G, C.c_int64, _P, _I, _P]),
    "stencil_batch_create": (_I, [C.POINTER(stencil_shot), _I, _I, _I, C.POINTER(_P), _P]),
    "stencil_batch_run": (_I, [_P, _P]),
    "stencil_batch_newest": (_I, [_P, _I]),
    "stencil_batch_groups": (_I, [_P]),
    "stencil_batch_group": (_I, [_P, _I, _G, C.POINTER(stencil_tile_shape), C.POINTER(_I)]),
    "stencil_batch_set_plan": (_I, [_P, C.POINTER(stencil_tile_shape), _I, _P]),
    "stencil_batch_is_ragged": (_I, [_P]),
    "stencil_batch_persistent_grid": (_I, [_P]),
    "stencil_batch_check": (_I, [_P, _P]),
    "stencil_batch_destroy": (None, [_P]),
    "stencil_config_default": (None, [C.POINTER(stencil_config)]),
    "stencil_solver_create": (_I, [C.POINTER(stencil_config), C.POINTER(_P)]),
    "stencil_solver_step": (_I, [_P]),
    "stencil_solver_run": (_I, [_P, _I, _I, C.POINTER(stencil_perf)]),
    "stencil_solver_download": (_I, [_P, _P, C.c_uint64]),
    "stencil_solver_dims": (_I, [_P, _G]),
    "stencil_solver_name": (_I, [_P, C.c_char_p, C.c_uint64]),
    "stencil_solver_destroy": (None, [_P]),
}

_lib = None


class StencilError(RuntimeError):
    """A non-zero status from the C-ABI."""

    def __init__(self, status: int, what: str):
        self.status = status
        msg = _lib.stencil_strerror(status).decode() if _lib is not None else str(status)
        extra = ""
        if status == 2 and _lib is not None:
            extra = f" (hip error {_lib.stencil_last_hip_error()})"
        super().__init__(f"{what}: {msg}{extra}")


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        raise StencilError(status, what)
