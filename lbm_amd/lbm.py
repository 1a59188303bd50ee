"""Host-side mirror of the reference's ``LBM`` type (lbm-wgpu src/lbm.rs:32-98).

Method names, argument meaning and order follow the reference so that code
written against the crate's surface reads the same:

==========================  ===============================================
reference (src/lbm.rs)      here
==========================  ===============================================
``LBM::new(d, omega, x, y)``  ``LBM(omega, x, y)`` (``lbm_create``)
``iterate(d, n)``             ``iterate(n)`` — n steps + the summary field
``draw_shape(d, shape)``      ``draw_shape(shape)`` — a shapes.Shape or points
``curl/chaos/welcome_barrier`` ``curl_barrier()`` etc. (shapes.py presets)
``update_omega_buffer``       ``update_omega_buffer(omega)``
``reset_barrier``             ``reset_barrier()``
``reset_to_equilibrium``      ``reset_to_equilibrium()``
``custom_speed(d, ux)``       ``custom_speed(ux)``
``single_cell(d, k)``         ``single_cell(k)``
``set_summary(stat)``         ``set_summary(stat)``
``rerender`` (summary part)   ``calculate_summary()``
``color_map`` + summary        ``render(stat, color_map)`` — (ny, nx, 3) colours
``get_compute_num``           ``get_compute_num()``
==========================  ===============================================

Everything runs in lbm_amd/liblbm_hip.so on the GPU; the arrays returned here
are host copies.  Errors raise ``LbmError`` (the reference panics instead).
"""
from __future__ import annotations

import ctypes
import os
from enum import IntEnum

import numpy as np

from . import _native as N
from . import shapes


class SummaryStat(IntEnum):
    """Order of the reference enum (src/lbm.rs:10-16)."""
    Curl = 0
    Ux = 1
    Uy = 2
    Rho = 3
    Speed = 4


class ColorMap(IntEnum):
    """Order of the reference enum (src/lbm.rs:20-24)."""
    Inferno = 0
    Viridis = 1
    Jet = 2


def _f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def to_rgb8(rgb) -> np.ndarray:
    """(..., 3) colours in [0, 1] -> uint8 (round to nearest)."""
    return np.clip(np.rint(np.asarray(rgb, np.float32) * 255.0), 0, 255).astype(np.uint8)


def save_ppm(path, rgb) -> None:
    """Write an (H, W, 3) float colour plane (lbm_render's output) as a binary
    PPM image, row 0 at the top as the reference's render pass draws it."""
    img = to_rgb8(rgb)
    h, w, _ = img.shape
    with open(path, "wb") as fh:
        fh.write(f"P6 {w} {h} 255\n".encode())
        fh.write(np.ascontiguousarray(img).tobytes())


def points_to_pairs(points, nx: int) -> np.ndarray:
    """get_points_vector (barrier_shapes/merge_shapes.rs:12-22): (x, y, solid)
    triples -> flat u32 [x + y * nx, 1|0, ...]."""
    pts = list(points)
    out = np.empty(2 * len(pts), np.uint32)
    for j, (x, y, v) in enumerate(pts):
        out[2 * j] = np.uint32(int(x) + int(y) * nx)
        out[2 * j + 1] = 1 if v else 0
    return out


class LBM:
    def __init__(self, omega: float, x: int, y: int, *, precision: int = N.PRECISION_F32,
                 boundary: int = N.BOUNDARY_REFERENCE, device: int = 0, u_init: float = 0.1,
                 row_pad: int = 0):
        self._lib = N.load()
        cfg = N.lbm_config()
        N.check(self._lib.lbm_config_default(ctypes.byref(cfg), x, y))
        cfg.omega = omega
        cfg.u_init = u_init
        cfg.precision = precision
        cfg.boundary = boundary
        cfg.device = device
        cfg.row_pad = row_pad
        h = ctypes.c_void_p()
        N.check(self._lib.lbm_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.x, self.y = int(x), int(y)
        self.precision = precision
        self.boundary = boundary
        self.summary_stat = SummaryStat.Curl
        self.color_map = ColorMap.Jet  # the reference's default (lbm.rs:1024)

    # -- lifecycle ---------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.lbm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc):
        return N.check(rc, self._h)

    @property
    def total(self) -> int:
        return self.x * self.y

    # -- reference surface ----------------------------------------------------------
    def iterate(self, compute_steps: int) -> np.ndarray:
        """n steps, then the current summary field (lbm.rs:1065-1074)."""
        self.step(compute_steps)
        return self.calculate_summary()

    def update_omega_buffer(self, omega: float):
        self._chk(self._lib.lbm_set_omega(self._h, omega))

    def update_viscosity(self, nu: float):
        self._chk(self._lib.lbm_set_viscosity(self._h, nu))

    @property
    def omega(self) -> float:
        v = ctypes.c_float()
        self._chk(self._lib.lbm_get_omega(self._h, ctypes.byref(v)))
        return v.value

    def reset_to_equilibrium(self):
        self.custom_speed(0.1)

    def custom_speed(self, ux: float):
        self.reset_equilibrium(1.0, ux, 0.0)

    def reset_equilibrium(self, rho: float, ux: float, uy: float):
        self._chk(self._lib.lbm_reset_equilibrium(self._h, rho, ux, uy))

    def single_cell(self, index: int):
        self._chk(self._lib.lbm_single_cell(self._h, index))

    def set_pulse(self, k: int, x: int, y: int, v: float):
        self._chk(self._lib.lbm_set_pulse(self._h, k, x, y, v))

    def reset_barrier(self):
        self._chk(self._lib.lbm_reset_barrier(self._h))

    def draw_shape(self, shape):
        """draw_shape (lbm.rs:1337-1339): paint a shapes.Shape, or any iterable
        of (x, y, solid) points; see points_to_pairs."""
        if isinstance(shape, shapes.Shape):
            shape = shapes.ordered_points(shape)
        self.paint(points_to_pairs(shape, self.x))

    def curl_barrier(self):
        """curl_barrier (lbm.rs:1367-1370)."""
        self.draw_shape(shapes.curl_barrier(self.x, self.y))

    def chaos_barrier(self):
        """chaos_barrier (lbm.rs:1372-1386)."""
        self.draw_shape(shapes.chaos_barrier(self.x, self.y))

    def welcome_barrier(self):
        """welcome_barrier (lbm.rs:1388-1480)."""
        self.draw_shape(shapes.welcome_barrier(self.x, self.y))

    def paint(self, pairs):
        a = np.ascontiguousarray(np.asarray(pairs, np.uint32).ravel())
        self._chk(self._lib.lbm_paint(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), a.size))

    def set_summary(self, stat):
        self.summary_stat = SummaryStat(stat)

    def calculate_summary(self, stat=None) -> np.ndarray:
        kind = int(self.summary_stat if stat is None else stat)
        out = np.empty((self.y, self.x), np.float32)
        self._chk(self._lib.lbm_summary(self._h, kind, _f32p(out)))
        return out

    def render(self, stat=None, color_map=None) -> np.ndarray:
        """calculate_summary + color_map (lbm.rs:1104-1110 without the draw):
        (ny, nx, 3) f32 colours of the current summary field."""
        kind = int(self.summary_stat if stat is None else stat)
        cmap = int(self.color_map if color_map is None else color_map)
        out = np.empty((self.y, self.x, 3), np.float32)
        self._chk(self._lib.lbm_render(self._h, kind, cmap, _f32p(out)))
        return out

    def rerender(self) -> np.ndarray:
        return self.render()

    def get_compute_num(self) -> int:
        n = ctypes.c_uint64()
        self._chk(self._lib.lbm_step_count(self._h, ctypes.byref(n)))
        return n.value

    compute_step = property(get_compute_num)

    # -- stepping ----------------------------------------------------------------------
    def step(self, n: int = 1):
        self._chk(self._lib.lbm_step(self._h, n))

    def step_timed(self, n: int) -> float:
        """n steps bracketed by HIP events on the grid's stream; device ms."""
        ms = ctypes.c_float()
        self._chk(self._lib.lbm_step_timed(self._h, n, ctypes.byref(ms)))
        return ms.value

    @property
    def variant(self) -> int:
        v = ctypes.c_int()
        self._chk(self._lib.lbm_get_variant(self._h, ctypes.byref(v)))
        return v.value

    @variant.setter
    def variant(self, v: int):
        self._chk(self._lib.lbm_set_variant(self._h, v))

    def autotune(self, steps: int = 20):
        """Time every step-kernel variant on this grid (state restored), keep the
        fastest; returns (best, ms_per_step list)."""
        best = ctypes.c_int()
        ms = (ctypes.c_float * N.NUM_VARIANTS)()
        self._chk(self._lib.lbm_autotune(self._h, steps, ctypes.byref(best), ms))
        return best.value, list(ms)

    def autotune_cached(self, cache_path: str, steps: int = 20):
        """autotune through a JSON cache keyed by (device, shape class,
        precision, boundary): a cached key selects its variant without timing.
        Returns (best, hit)."""
        best, hit = ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.lbm_autotune_cached(self._h, os.fsencode(cache_path), steps, ctypes.byref(best),
                                                ctypes.byref(hit)))
        return best.value, bool(hit.value)

    def sync(self):
        self._chk(self._lib.lbm_sync(self._h))

    @property
    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._chk(self._lib.lbm_get_stream(self._h, ctypes.byref(s)))
        return s.value or 0

    # -- obstacles -----------------------------------------------------------------------
    @property
    def mask(self) -> np.ndarray:
        m = np.empty((self.y, self.x), np.uint8)
        self._chk(self._lib.lbm_get_mask(self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return m

    @mask.setter
    def mask(self, solid):
        m = np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(self.y, self.x))
        self._chk(self._lib.lbm_set_mask(self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    # -- fields and state --------------------------------------------------------------------
    def moments(self, lagged: bool = False):
        """(jx, jy, rho) planes; lagged=True gives the renderer's one-step-old view."""
        jx = np.empty((self.y, self.x), np.float32)
        jy = np.empty_like(jx)
        rho = np.empty_like(jx)
        self._chk(self._lib.lbm_read_moments(self._h, _f32p(jx), _f32p(jy), _f32p(rho), int(lagged)))
        return jx, jy, rho

    def fields(self, lagged: bool = False):
        """(rho, ux, uy, curl) with u = j / rho."""
        outs = [np.empty((self.y, self.x), np.float32) for _ in range(4)]
        self._chk(self._lib.lbm_read_fields(self._h, *[_f32p(o) for o in outs], int(lagged)))
        return tuple(outs)

    def conservation(self) -> np.ndarray:
        out = np.zeros(3, np.float64)
        self._chk(self._lib.lbm_conservation(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def read_f(self) -> np.ndarray:
        f = np.empty((9, self.y, self.x), np.float32)
        self._chk(self._lib.lbm_read_f(self._h, _f32p(f)))
        return f

    def write_f(self, f9):
        a = np.ascontiguousarray(np.asarray(f9, np.float32).reshape(9, self.y, self.x))
        self._chk(self._lib.lbm_write_f(self._h, _f32p(a)))

    def snapshot(self):
        """Device-side copy of the evolving state (one slot)."""
        self._chk(self._lib.lbm_snapshot(self._h))

    def restore(self):
        self._chk(self._lib.lbm_restore(self._h))

    def velocity_mark(self, slot: int = 0):
        self._chk(self._lib.lbm_velocity_mark_slot(self._h, slot))

    def velocity_skill(self):
        """(|u - m1| / |u - m0|, |u - m1| / |m1|) over fluid cells: u the current
        velocity, m0 / m1 the marks of slots 0 and 1."""
        out = np.zeros(2, np.float64)
        self._chk(self._lib.lbm_velocity_skill(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return float(out[0]), float(out[1])

    def velocity_change(self) -> float:
        r = ctypes.c_double()
        self._chk(self._lib.lbm_velocity_change(self._h, ctypes.byref(r)))
        return r.value

    def surrogate_jump(self, net, steps: int = 64, s_rho: float = 0.01, s_u: float = 0.1, d_rho: float = 0.01,
                       d_u: float = 0.003):
        """Predict the change `steps` ahead with a trained UNet of shape (1, ny, nx)
        or (1, ny / 2, nx / 2) and re-equilibrate the streamed fluid cells from
        the prediction; s_*: input units, d_*: output units."""
        self._chk(self._lib.lbm_surrogate_jump(self._h, net._h, steps, s_rho, s_u, d_rho, d_u))

    def save_state(self) -> bytes:
        n = ctypes.c_size_t()
        self._chk(self._lib.lbm_state_size(self._h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._chk(self._lib.lbm_save_state(self._h, ctypes.cast(buf, ctypes.c_void_p), n.value))
        return buf.raw

    def load_state(self, blob: bytes):
        buf = ctypes.create_string_buffer(blob, len(blob))
        self._chk(self._lib.lbm_load_state(self._h, ctypes.cast(buf, ctypes.c_void_p), len(blob)))


def gen_circles(nx: int, ny: int, seed: int, n: int, r_min: int, r_max: int, walls: bool = True) -> np.ndarray:
    """Seeded circle obstacles (lbm_gen_circles), u8 (ny, nx), 1 = solid."""
    lib = N.load()
    m = np.zeros((ny, nx), np.uint8)
    if walls:
        m[0, :] = 1
        m[-1, :] = 1
    N.check(lib.lbm_gen_circles(m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), nx, ny, seed, n, r_min, r_max))
    return m


def splitmix64(seed: int, count: int):
    lib = N.load()
    s = ctypes.c_uint64(seed)
    return [lib.lbm_splitmix64(ctypes.byref(s)) for _ in range(count)]
