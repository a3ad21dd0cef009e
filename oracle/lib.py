"""ctypes bindings of the oracle libraries (TEST INFRASTRUCTURE ONLY).

Fields are numpy arrays of shape (nz, ny, nx), C order, so that the flat
index is z*nx*ny + y*nx + x as in the reference (include/stencil/types.hpp:126-128).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
_SZ = C.c_size_t
_F = C.c_float
_D = C.c_double
_P = C.c_void_p


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _dims(a: np.ndarray):
    nz, ny, nx = a.shape
    return nx, ny, nz


class Oracle:
    """The C restatement (oracle/stencil_oracle.c)."""

    def __init__(self, path: str | os.PathLike | None = None):
        path = Path(path) if path else ORACLE_DIR / "liboracle.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (run `make -C oracle`)")
        L = self.lib = C.CDLL(str(path))
        L.oracle_wave_step_f32.argtypes = [_P] * 4 + [_SZ] * 3 + [_F] * 4
        L.oracle_wave_step_f64.argtypes = [_P] * 4 + [_SZ] * 3 + [_D] * 4
        L.oracle_make_grid_dt_f32.argtypes = [_F] * 4
        L.oracle_make_grid_dt_f32.restype = _F
        L.oracle_heat_dt_eff_f32.argtypes = [_F]
        L.oracle_heat_dt_eff_f32.restype = _F
        L.oracle_heat_step_f32.argtypes = [_P] * 2 + [_SZ] * 3 + [_F] * 5
        L.oracle_init_gaussian_f32.argtypes = [_P] + [_SZ] * 3 + [_F] * 4
        L.oracle_init_gaussian_f64.argtypes = [_P] + [_SZ] * 3 + [_D] * 4
        L.oracle_init_constant_velocity_f32.argtypes = [_P, _SZ, _F]
        L.oracle_init_layered_velocity_f32.argtypes = [_P] + [_SZ] * 3 + [_F] * 2
        L.oracle_l2_error_f32.argtypes = [_P, _P, _SZ]
        L.oracle_l2_error_f32.restype = _D
        L.oracle_max_abs_error_f32.argtypes = [_P, _P, _SZ]
        L.oracle_max_abs_error_f32.restype = _D
        L.oracle_wave_run_f32.argtypes = [_P] * 6 + [_SZ] * 3 + [_F] * 4 + [C.c_int]
        L.oracle_wave_run_f64.argtypes = [_P] * 6 + [_SZ] * 3 + [_D] * 4 + [C.c_int]
        L.oracle_heat_run_f32.argtypes = [_P] * 4 + [_SZ] * 3 + [_F] * 5 + [C.c_int]
        L.oracle_splitmix64.argtypes = [C.c_uint64]
        L.oracle_splitmix64.restype = C.c_uint64
        L.oracle_u01.argtypes = [C.c_uint64] * 3
        L.oracle_u01.restype = _F
        L.oracle_layered_noise_vel2_f32.argtypes = [_P] + [_SZ] * 3 + [C.c_int] + [_F] * 3 + [C.c_uint64]
        L.oracle_heat_noise_cube_f32.argtypes = [_P] + [_SZ] * 4 + [_F, C.c_uint64]
        L.oracle_rel_l2_interior_f32.argtypes = [_P, _P] + [_SZ] * 3
        L.oracle_rel_l2_interior_f32.restype = _D
        L.oracle_rel_l2_interior_f64f32.argtypes = [_P, _P] + [_SZ] * 3
        L.oracle_rel_l2_interior_f64f32.restype = _D
        L.oracle_f32_to_bf16.argtypes = [_F]
        L.oracle_f32_to_bf16.restype = C.c_uint16
        L.oracle_bf16_to_f32.argtypes = [C.c_uint16]
        L.oracle_bf16_to_f32.restype = _F
        L.oracle_f32_to_fp16.argtypes = [_F]
        L.oracle_f32_to_fp16.restype = C.c_uint16
        L.oracle_fp16_to_f32.argtypes = [C.c_uint16]
        L.oracle_fp16_to_f32.restype = _F
        L.oracle_encode16.argtypes = [C.c_int, _P, _P, _P, _SZ, _F]
        L.oracle_decode16.argtypes = [C.c_int, _P, _P, _P, _SZ, _F]
        L.oracle_wave_step16.argtypes = [C.c_int] + [_P] * 8 + [_SZ] * 3 + [_F] * 6
        for suf, T in (("f32", _F), ("f64", _D)):
            getattr(L, f"oracle_fwi_forward_{suf}").argtypes = [_P] + [_SZ] * 3 + [T] * 4 + [C.c_int, _SZ, _SZ,
                                                                                              _P, _P]
            g = getattr(L, f"oracle_fwi_gradient_{suf}")
            g.argtypes = [_P] + [_SZ] * 3 + [T] * 4 + [C.c_int, _SZ, _SZ, _P, _P, _P, _P]
            g.restype = _D
        L.oracle_shot_run_f32.argtypes = [_P] + [_SZ] * 3 + [_F] * 4 + [C.c_int, _SZ, _P, _P]

    # -- coefficients ------------------------------------------------------
    def fd_coeffs(self, dtype=np.float32) -> np.ndarray:
        if dtype == np.float32:
            c = (_F * 5)()
            self.lib.oracle_fd_coeffs_f32(c)
        else:
            c = (_D * 5)()
            self.lib.oracle_fd_coeffs_f64(c)
        return np.array(list(c), dtype=dtype)

    # -- one step ----------------------------------------------------------
    def wave_step(self, cur, prev, vel2, nxt, dx, dt, dy=None, dz=None):
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        f = self.lib.oracle_wave_step_f32 if cur.dtype == np.float32 else self.lib.oracle_wave_step_f64
        f(_ptr(cur), _ptr(prev), _ptr(vel2), _ptr(nxt), *_dims(cur), dx, dy, dz, dt)
        return nxt

    def heat_step(self, t, t_next, dx, dt_eff, alpha, dy=None, dz=None):
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        self.lib.oracle_heat_step_f32(_ptr(t), _ptr(t_next), *_dims(t), dx, dy, dz, dt_eff, alpha)
        return t_next

    # -- scalars -----------------------------------------------------------
    def make_grid_dt(self, dx, velocity, dy=None, dz=None) -> float:
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        return float(np.float32(self.lib.oracle_make_grid_dt_f32(dx, dy, dz, velocity)))

    def heat_dt_eff(self, dt) -> float:
        return float(np.float32(self.lib.oracle_heat_dt_eff_f32(dt)))

    # -- initial conditions -----------------------------------------------
    def gaussian(self, n, cx=0.5, cy=0.5, cz=0.5, sigma=5.0, dtype=np.float32, shape=None):
        nz, ny, nx = shape if shape else (n, n, n)
        g = np.zeros((nz, ny, nx), dtype=dtype)
        f = self.lib.oracle_init_gaussian_f32 if dtype == np.float32 else self.lib.oracle_init_gaussian_f64
        f(_ptr(g), nx, ny, nz, cx, cy, cz, sigma)
        return g

    def constant_velocity(self, shape, v):
        g = np.empty(shape, dtype=np.float32)
        self.lib.oracle_init_constant_velocity_f32(_ptr(g), g.size, v)
        return g

    def layered_velocity(self, shape, v_top, v_bot):
        g = np.empty(shape, dtype=np.float32)
        nz, ny, nx = shape
        self.lib.oracle_init_layered_velocity_f32(_ptr(g), nx, ny, nz, v_top, v_bot)
        return g

    def layered_noise_vel2(self, shape, n_layers=16, vmin=1500.0, vmax=4000.0, noise=0.05, seed=2):
        g = np.empty(shape, dtype=np.float32)
        nz, ny, nx = shape
        self.lib.oracle_layered_noise_vel2_f32(_ptr(g), nx, ny, nz, n_layers, vmin, vmax, noise, seed)
        return g

    def heat_noise_cube(self, shape, side=64, hot=10.0, seed=2):
        g = np.empty(shape, dtype=np.float32)
        nz, ny, nx = shape
        self.lib.oracle_heat_noise_cube_f32(_ptr(g), nx, ny, nz, side, hot, seed)
        return g

    def u01(self, seed, stream, idx) -> float:
        return float(self.lib.oracle_u01(seed, stream, idx))

    def splitmix64(self, x) -> int:
        return int(self.lib.oracle_splitmix64(x))

    # -- norms ---------------------------------------------------------------
    def l2_error(self, a, b) -> float:
        return float(self.lib.oracle_l2_error_f32(_ptr(a), _ptr(b), a.size))

    def max_abs_error(self, a, b) -> float:
        return float(self.lib.oracle_max_abs_error_f32(_ptr(a), _ptr(b), a.size))

    def rel_l2_interior(self, a, b) -> float:
        if b.dtype == np.float64:
            return float(self.lib.oracle_rel_l2_interior_f64f32(_ptr(a), _ptr(b), *_dims(a)))
        return float(self.lib.oracle_rel_l2_interior_f32(_ptr(a), _ptr(b), *_dims(a)))

    # -- 16-bit storage ----------------------------------------------------
    FORMATS = {"bf16": 1, "fp16": 2, "bfp16": 3}

    def encode16(self, fmt, x, scale=1.0):
        """float32 array -> (uint16 array, int8 block exponents or None)."""
        f = self.FORMATS[fmt]
        x = np.ascontiguousarray(x, dtype=np.float32)
        d = np.zeros(x.shape, dtype=np.uint16)
        e = np.zeros(x.size // 16, dtype=np.int8)
        self.lib.oracle_encode16(f, _ptr(x), _ptr(d), _ptr(e), x.size, scale)
        return d, (e if f == 3 else None)

    def decode16(self, fmt, d, e=None, scale=1.0):
        f = self.FORMATS[fmt]
        out = np.empty(d.shape, dtype=np.float32)
        ee = e if e is not None else np.zeros(max(d.size // 16, 1), dtype=np.int8)
        self.lib.oracle_decode16(f, _ptr(d), _ptr(ee), _ptr(out), d.size, scale)
        return out

    def wave_step16(self, fmt, cur, prev, vel2, nxt, dx, dt, p_scale=1.0, v_scale=1.0, dy=None, dz=None):
        """One 16-bit step; each field is (uint16 array, int8 exps or None); nxt is updated."""
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        f = self.FORMATS[fmt]
        dummy = np.zeros(max(cur[0].size // 16, 1), dtype=np.int8)
        ex = lambda fe: _ptr(fe[1]) if fe[1] is not None else _ptr(dummy)
        nz, ny, nx = cur[0].shape
        self.lib.oracle_wave_step16(f, _ptr(cur[0]), ex(cur), _ptr(prev[0]), ex(prev), _ptr(vel2[0]), ex(vel2),
                                    _ptr(nxt[0]), ex(nxt), nx, ny, nz, dx, dy, dz, dt, p_scale, v_scale)
        return nxt

    # -- adjoint-state FWI ------------------------------------------------
    def fwi_forward(self, m, h, dt, n_steps, src, rec_z, wavelet):
        """Traces (n_steps, ny, nx): trace n = p^{n+1} on plane rec_z.  src is (z, y, x)."""
        nz, ny, nx = m.shape
        T = m.dtype
        out = np.zeros((n_steps, ny, nx), dtype=T)
        w = np.ascontiguousarray(wavelet, dtype=T)
        f = self.lib.oracle_fwi_forward_f32 if T == np.float32 else self.lib.oracle_fwi_forward_f64
        si = (src[0] * ny + src[1]) * nx + src[2]
        f(_ptr(m), nx, ny, nz, h, h, h, dt, n_steps, si, rec_z, _ptr(w), _ptr(out))
        return out

    def fwi_gradient(self, m, h, dt, n_steps, src, rec_z, wavelet, d_obs):
        """(J, dJ/dm, lambda^{n+1} at src) of J = 1/2 |traces - d_obs|^2."""
        nz, ny, nx = m.shape
        T = m.dtype
        grad = np.zeros_like(m)
        st = np.zeros(n_steps, dtype=T)
        w = np.ascontiguousarray(wavelet, dtype=T)
        d = np.ascontiguousarray(d_obs, dtype=T)
        f = self.lib.oracle_fwi_gradient_f32 if T == np.float32 else self.lib.oracle_fwi_gradient_f64
        si = (src[0] * ny + src[1]) * nx + src[2]
        J = f(_ptr(m), nx, ny, nz, h, h, h, dt, n_steps, si, rec_z, _ptr(w), _ptr(d), _ptr(grad), _ptr(st))
        return J, grad, st

    def shot_run(self, m, h, dt, n_steps, src, wavelet):
        """p^{n_steps} of one shot (FP32) from zero fields; src (z, y, x) or None."""
        nz, ny, nx = m.shape
        out = np.zeros_like(m)
        if wavelet is None:
            self.lib.oracle_shot_run_f32(_ptr(m), nx, ny, nz, h, h, h, dt, n_steps, 0, None, _ptr(out))
        else:
            w = np.ascontiguousarray(wavelet, dtype=np.float32)
            si = (src[0] * ny + src[1]) * nx + src[2]
            self.lib.oracle_shot_run_f32(_ptr(m), nx, ny, nz, h, h, h, dt, n_steps, si, _ptr(w), _ptr(out))
        return out

    # -- solver runs -------------------------------------------------------
    def wave_run(self, p0, vel2, dx, dt, n_steps, dy=None, dz=None):
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        bufs = [np.empty_like(p0) for _ in range(3)]
        out = np.empty_like(p0)
        f = self.lib.oracle_wave_run_f32 if p0.dtype == np.float32 else self.lib.oracle_wave_run_f64
        f(_ptr(p0), _ptr(vel2), *[_ptr(b) for b in bufs], _ptr(out), *_dims(p0), dx, dy, dz, dt, n_steps)
        return out

    def heat_run(self, t0, dx, dt_eff, alpha, n_steps, dy=None, dz=None):
        dy = dx if dy is None else dy
        dz = dx if dz is None else dz
        bufs = [np.empty_like(t0) for _ in range(2)]
        out = np.empty_like(t0)
        self.lib.oracle_heat_run_f32(_ptr(t0), *[_ptr(b) for b in bufs], _ptr(out), *_dims(t0), dx, dy, dz,
                                     dt_eff, alpha, n_steps)
        return out


class RefLib:
    """The reference's own CPU path (oracle/_ref/libref_<flavour>.so)."""

    def __init__(self, flavour: str = "nofma"):
        path = ORACLE_DIR / "_ref" / f"libref_{flavour}.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (run `make -C oracle ref`)")
        L = self.lib = C.CDLL(str(path))
        L.ref_sizeof_real.restype = C.c_int
        self.real = C.c_float if L.ref_sizeof_real() == 4 else C.c_double
        self.dtype = np.float32 if self.real is C.c_float else np.float64
        T = self.real
        L.ref_wave_step.argtypes = [_P] * 4 + [_SZ] * 3 + [T] * 4
        L.ref_solver.argtypes = [C.c_int, C.c_int] + [_SZ] * 3 + [T] * 7 + [C.c_int] * 3 + [_P]
        L.ref_solver.restype = _D
        L.ref_init_gaussian.argtypes = [_P] + [_SZ] * 3 + [T] * 4
        L.ref_init_layered_velocity.argtypes = [_P] + [_SZ] * 3 + [T] * 2
        L.ref_make_grid_dt.argtypes = [T] * 4
        L.ref_make_grid_dt.restype = T
        L.ref_l2_error.argtypes = [_P, _P] + [_SZ] * 3
        L.ref_l2_error.restype = _D
        L.ref_max_abs_error.argtypes = [_P, _P] + [_SZ] * 3
        L.ref_max_abs_error.restype = _D
        L.ref_time_wave_steps.argtypes = [_P] * 4 + [_SZ] * 3 + [T] * 2 + [C.c_int] * 2
        L.ref_time_wave_steps.restype = _D
        L.ref_omp_max_threads.restype = C.c_int
        L.ref_set_ftz_daz.argtypes = [C.c_int]
        L.ref_set_ftz_daz.restype = C.c_int

    def omp_threads(self) -> int:
        return int(self.lib.ref_omp_max_threads())

    def set_ftz_daz(self, on: bool) -> int:
        """FTZ/DAZ in every OpenMP pool thread; returns the threads that took it."""
        return int(self.lib.ref_set_ftz_daz(int(bool(on))))

    def wave_step(self, cur, prev, vel2, nxt, dx, dt):
        self.lib.ref_wave_step(_ptr(cur), _ptr(prev), _ptr(vel2), _ptr(nxt), *_dims(cur), dx, dx, dx, dt)
        return nxt

    def solver(self, physics, variant, n, dx=10.0, velocity=1500.0, diffusivity=1e-4,
               cx=0.5, cy=0.5, cz=0.5, sigma=5.0, n_steps=1, warmup=0, use_run=False):
        out = np.empty((n, n, n), dtype=self.dtype)
        ms = self.lib.ref_solver(physics, variant, n, n, n, dx, velocity, diffusivity, cx, cy, cz, sigma,
                                 n_steps, warmup, int(use_run), _ptr(out))
        return out, ms

    def gaussian(self, n, cx=0.5, cy=0.5, cz=0.5, sigma=5.0):
        out = np.empty((n, n, n), dtype=self.dtype)
        self.lib.ref_init_gaussian(_ptr(out), n, n, n, cx, cy, cz, sigma)
        return out

    def layered_velocity(self, shape, v_top, v_bot):
        nz, ny, nx = shape
        out = np.empty(shape, dtype=self.dtype)
        self.lib.ref_init_layered_velocity(_ptr(out), nx, ny, nz, v_top, v_bot)
        return out

    def make_grid_dt(self, dx, velocity):
        return float(self.lib.ref_make_grid_dt(dx, dx, dx, velocity))

    def l2_error(self, a, b):
        return float(self.lib.ref_l2_error(_ptr(a), _ptr(b), *_dims(a)))

    def max_abs_error(self, a, b):
        return float(self.lib.ref_max_abs_error(_ptr(a), _ptr(b), *_dims(a)))

    def time_wave_steps(self, p0, vel2, dx, dt, warmup, n_steps):
        bufs = [p0.copy(), p0.copy(), np.zeros_like(p0)]
        return float(self.lib.ref_time_wave_steps(*[_ptr(b) for b in bufs], _ptr(vel2), *_dims(p0), dx, dt,
                                                  warmup, n_steps))


def load_oracle() -> Oracle:
    return Oracle()


def ref_available(flavour: str = "nofma") -> bool:
    return (ORACLE_DIR / "_ref" / f"libref_{flavour}.so").exists()


def load_ref(flavour: str = "nofma") -> RefLib:
    return RefLib(flavour)
