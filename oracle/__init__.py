"""CPU oracle for the D2Q9 simulation step — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import
this package, and only to check results (or as the timed CPU baseline).  The
product package ``lbm_amd`` never imports it.

Contents
--------
``RefCPU``    ctypes wrapper of refcpu.c: a faithful f32 restatement of the
              reference's eight WGSL passes (lbm-wgpu src/lbm.rs:1112-1134).
``Oracle64``  ctypes wrapper of oracle64.c: an independent FP64 D2Q9 BGK solver
              with the same boundary semantics.
``np_ref``    a second, independently written f32 restatement in numpy, used to
              cross-check ``RefCPU`` on small grids.

Pins (tests/test_oracle.py): omega = 0 pure-streaming permutation and the
bounce-back known-answer tables, equilibrium identities and the f32 fixed point,
closed-box mass conservation, per-cell collision conservation, RefCPU == np_ref
bit for bit, RefCPU vs Oracle64 at the f32 rounding level, and the summary pass
(curl, display density and speed, velocity planes) against closed forms.  The reference has no
tests, fixtures or golden vectors of its own for this path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile refcpu.c + oracle64.c into liboracle.so (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


def build_native(out_dir: str) -> str:
    """refcpu.c + oracle64.c built for this host (-O3 -march=native, no
    contraction) into out_dir/liboracle_native.so: the CPU baseline's build."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "native", f"NATIVE_OUT={out_dir}"], check=True)
    return os.path.join(out_dir, "liboracle_native.so")


_libs = {}


def _load(path: str = None):
    global _lib
    if path is not None and path != _LIB_PATH:
        if path not in _libs:
            _libs[path] = _bind(ctypes.CDLL(path))
        return _libs[path]
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def _bind(lib):
    P = ctypes.c_void_p
    u32, u64, f32, f64, i32 = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float,
                               ctypes.c_double, ctypes.c_int)
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    up = ctypes.POINTER(ctypes.c_uint32)
    bp = ctypes.POINTER(ctypes.c_uint8)
    sig = {
        "ref_new": (P, [u32, u32, f32]),
        "ref_free": (None, [P]),
        "ref_set_threads": (None, [i32]),
        "ref_get_threads": (i32, []),
        "ref_host_feq": (None, [f32, f32, f32, fp]),
        "ref_reset_barrier": (None, [P]),
        "ref_set_omega": (None, [P, f32]),
        "ref_collide": (None, [P]),
        "ref_stream": (None, [P]),
        "ref_step": (None, [P, u32]),
        "ref_reset_equilibrium": (None, [P, f32, f32, f32]),
        "ref_single_cell": (i32, [P, i32]),
        "ref_paint": (i32, [P, up, u64]),
        "ref_set_mask": (None, [P, up]),
        "ref_get_mask": (None, [P, up]),
        "ref_summary": (None, [P, i32, fp]),
        "ref_get_state": (None, [P, fp]),
        "ref_get_set": (None, [P, i32, fp]),
        "ref_set_set": (None, [P, i32, fp]),
        "ref_get_density": (None, [P, fp, fp, fp]),
        "ref_step_count": (u64, [P]),
        "ref_set_step_count": (None, [P, u64]),
        "o64_new": (P, [i32, i32, f64, f64, i32]),
        "o64_free": (None, [P]),
        "o64_set_omega": (None, [P, f64]),
        "o64_set_solid": (None, [P, bp]),
        "o64_set_equilibrium": (None, [P, f64, f64, f64]),
        "o64_set_state": (None, [P, dp]),
        "o64_get_state": (None, [P, dp]),
        "o64_step": (None, [P, i32]),
        "o64_macro": (None, [P, dp, dp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def host_feq(ux: float, uy: float, rho: float) -> np.ndarray:
    """The reference's host equilibrium (set_equil, src/lbm.rs:611-643), 9 f32."""
    lib = _load()
    out = np.zeros(9, np.float32)
    lib.ref_host_feq(ux, uy, rho, _ptr(out, ctypes.c_float))
    return out


class RefCPU:
    """Faithful f32 restatement of the reference step (see refcpu.c)."""

    def __init__(self, nx: int, ny: int, omega: float, threads: int = 0, lib_path: str = None):
        self.lib = _load(lib_path)
        if threads:
            self.lib.ref_set_threads(threads)
        self.nx, self.ny = int(nx), int(ny)
        self.h = self.lib.ref_new(self.nx, self.ny, omega)
        if not self.h:
            raise ValueError("ref_new failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_free(self.h)
            self.h = None

    @property
    def total(self) -> int:
        return self.nx * self.ny

    def set_omega(self, omega: float):
        self.lib.ref_set_omega(self.h, omega)

    def step(self, n: int = 1):
        self.lib.ref_step(self.h, n)

    def collide(self):
        self.lib.ref_collide(self.h)

    def stream(self):
        self.lib.ref_stream(self.h)

    def reset_equilibrium(self, rho: float = 1.0, ux: float = 0.1, uy: float = 0.0):
        self.lib.ref_reset_equilibrium(self.h, rho, ux, uy)

    def single_cell(self, k: int):
        self.lib.ref_single_cell(self.h, k)

    def reset_barrier(self):
        self.lib.ref_reset_barrier(self.h)

    def paint(self, pairs) -> int:
        a = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).ravel())
        return self.lib.ref_paint(self.h, _ptr(a, ctypes.c_uint32), a.size)

    def set_mask(self, mask):
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint32).reshape(-1))
        assert m.size == self.total
        self.lib.ref_set_mask(self.h, _ptr(m, ctypes.c_uint32))

    def mask(self) -> np.ndarray:
        m = np.zeros(self.total, np.uint32)
        self.lib.ref_get_mask(self.h, _ptr(m, ctypes.c_uint32))
        return m.reshape(self.ny, self.nx)

    def state(self) -> np.ndarray:
        """S_n as (9, ny, nx) f32: planes k != 4 of set n%2, rest plane of set 0."""
        out = np.zeros((9, self.ny, self.nx), np.float32)
        self.lib.ref_get_state(self.h, _ptr(out, ctypes.c_float))
        return out

    def get_set(self, p: int) -> np.ndarray:
        out = np.zeros((9, self.ny, self.nx), np.float32)
        self.lib.ref_get_set(self.h, p, _ptr(out, ctypes.c_float))
        return out

    def set_set(self, p: int, f9):
        a = np.ascontiguousarray(np.asarray(f9, np.float32).reshape(9, self.ny, self.nx))
        self.lib.ref_set_set(self.h, p, _ptr(a, ctypes.c_float))

    def density(self):
        """density_bg: (jx, jy, rho) planes, each (ny, nx)."""
        jx = np.zeros(self.total, np.float32)
        jy = np.zeros(self.total, np.float32)
        rho = np.zeros(self.total, np.float32)
        self.lib.ref_get_density(self.h, _ptr(jx, ctypes.c_float), _ptr(jy, ctypes.c_float),
                                 _ptr(rho, ctypes.c_float))
        s = (self.ny, self.nx)
        return jx.reshape(s), jy.reshape(s), rho.reshape(s)

    def summary(self, kind: int) -> np.ndarray:
        out = np.zeros(self.total, np.float32)
        self.lib.ref_summary(self.h, kind, _ptr(out, ctypes.c_float))
        return out.reshape(self.ny, self.nx)

    @property
    def compute_step(self) -> int:
        return int(self.lib.ref_step_count(self.h))

    @compute_step.setter
    def compute_step(self, n: int):
        self.lib.ref_set_step_count(self.h, n)


class Oracle64:
    """Independent FP64 D2Q9 BGK solver (see oracle64.c).  mode 0 = the
    reference's flattened-index outlet, 1 = zero-gradient outflow."""

    def __init__(self, nx: int, ny: int, omega: float, u_in: float, mode: int = 0):
        self.lib = _load()
        self.nx, self.ny = int(nx), int(ny)
        self.h = self.lib.o64_new(self.nx, self.ny, omega, u_in, mode)
        if not self.h:
            raise ValueError("o64_new failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.o64_free(self.h)
            self.h = None

    def set_solid(self, solid):
        s = np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(-1))
        assert s.size == self.nx * self.ny
        self.lib.o64_set_solid(self.h, _ptr(s, ctypes.c_uint8))

    def set_omega(self, omega: float):
        self.lib.o64_set_omega(self.h, omega)

    def set_equilibrium(self, rho=1.0, ux=0.1, uy=0.0):
        self.lib.o64_set_equilibrium(self.h, rho, ux, uy)

    def set_state(self, f9):
        a = np.ascontiguousarray(np.asarray(f9, np.float64).reshape(9, self.ny, self.nx))
        self.lib.o64_set_state(self.h, _ptr(a, ctypes.c_double))

    def state(self) -> np.ndarray:
        out = np.zeros((9, self.ny, self.nx), np.float64)
        self.lib.o64_get_state(self.h, _ptr(out, ctypes.c_double))
        return out

    def step(self, n: int = 1):
        self.lib.o64_step(self.h, n)

    def macro(self):
        """(rho, ux, uy) with uy north-positive, each (ny, nx) f64."""
        n = self.nx * self.ny
        r, ux, uy = np.zeros(n), np.zeros(n), np.zeros(n)
        self.lib.o64_macro(self.h, _ptr(r, ctypes.c_double), _ptr(ux, ctypes.c_double),
                           _ptr(uy, ctypes.c_double))
        s = (self.ny, self.nx)
        return r.reshape(s), ux.reshape(s), uy.reshape(s)


def moments_ref_order(f9: np.ndarray):
    """(jx, jy, rho) of a state in the reference's f32 summation order
    (corner_pre_collision.wgsl:19-21, cardinal_pre_collision.wgsl:19-21,
    corner_collision.wgsl:24).  f9 is (9, ...) f32."""
    f = np.asarray(f9, np.float32)
    nw, n, ne, w, c, e, sw, s, se = (f[k] for k in range(9))
    jx = ((ne + se) - nw) - sw
    jy = ((ne + nw) - se) - sw
    rho = ((ne + se) + nw) + sw
    jx = jx + (e - w)
    jy = jy + (n - s)
    rho = rho + (((e + n) + s) + w)
    rho = rho + c
    return jx, jy, rho
