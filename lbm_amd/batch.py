"""Batched domains: many independent D2Q9 grids stepped together on one GPU
(lbm_batch_* in include/lbm_hip.h).  No reference counterpart: the reference
steps one grid; each domain here follows that grid's semantics exactly."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _native as N


def gen_batch(seed: int, n: int):
    """n domain descriptors (nx, ny, omega, u_in) from one seed (lbm_gen_batch)."""
    lib = N.load()
    arr = (N.lbm_domain * n)()
    N.check_batch(lib.lbm_gen_batch(seed, n, arr))
    return [(d.nx, d.ny, d.omega, d.u_in) for d in arr]


def gen_shapes(nx: int, ny: int, seed: int, n_min: int = 1, n_max: int = 12, walls: bool = True) -> np.ndarray:
    """Seeded circles / rectangles (lbm_gen_shapes), u8 (ny, nx), 1 = solid."""
    lib = N.load()
    m = np.zeros((ny, nx), np.uint8)
    if walls:
        m[0, :] = 1
        m[-1, :] = 1
    N.check_batch(lib.lbm_gen_shapes(m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), nx, ny, seed, n_min, n_max))
    return m


class Batch:
    def __init__(self, domains, *, precision: int = N.PRECISION_F32, boundary: int = N.BOUNDARY_REFERENCE,
                 device: int = 0):
        self._lib = N.load()
        doms = list(domains)
        arr = (N.lbm_domain * len(doms))()
        for i, (nx, ny, omega, u_in) in enumerate(doms):
            arr[i].nx, arr[i].ny, arr[i].omega, arr[i].u_in = int(nx), int(ny), float(omega), float(u_in)
        h = ctypes.c_void_p()
        N.check_batch(self._lib.lbm_batch_create(arr, len(doms), precision, boundary, device, ctypes.byref(h)))
        self._h = h
        self.domains = [(int(nx), int(ny), float(om), float(u)) for nx, ny, om, u in doms]
        self.precision = precision

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.lbm_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        return N.check_batch(rc, self._h)

    def __len__(self):
        return len(self.domains)

    def reset(self):
        self._chk(self._lib.lbm_batch_reset(self._h))

    def set_mask(self, d: int, solid):
        nx, ny = self.domains[d][:2]
        m = np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(ny, nx))
        self._chk(self._lib.lbm_batch_set_mask(self._h, d, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def step(self, n: int = 1):
        self._chk(self._lib.lbm_batch_step(self._h, n))

    def step_timed(self, n: int) -> float:
        ms = ctypes.c_float()
        self._chk(self._lib.lbm_batch_step_timed(self._h, n, ctypes.byref(ms)))
        return ms.value

    def sync(self):
        self._chk(self._lib.lbm_batch_sync(self._h))

    @property
    def step_count(self) -> int:
        n = ctypes.c_uint64()
        self._chk(self._lib.lbm_batch_step_count(self._h, ctypes.byref(n)))
        return n.value

    def info(self):
        c = ctypes.c_uint64()
        i = ctypes.c_uint32()
        b = ctypes.c_uint32()
        self._chk(self._lib.lbm_batch_info(self._h, ctypes.byref(c), ctypes.byref(i), ctypes.byref(b)))
        return {"cells": c.value, "interior_tiles": i.value, "boundary_segments": b.value}

    def set_graph(self, enable: bool):
        self._chk(self._lib.lbm_batch_set_graph(self._h, int(enable)))

    @property
    def variant(self) -> int:
        v = ctypes.c_int()
        self._chk(self._lib.lbm_batch_get_variant(self._h, ctypes.byref(v)))
        return v.value

    @variant.setter
    def variant(self, v: int):
        self._chk(self._lib.lbm_batch_set_variant(self._h, v))

    def autotune(self, steps: int = 10):
        best = ctypes.c_int()
        ms = (ctypes.c_float * N.BATCH_NUM_VARIANTS)()
        self._chk(self._lib.lbm_batch_autotune(self._h, steps, ctypes.byref(best), ms))
        return best.value, list(ms)

    def autotune_cached(self, cache_path: str, steps: int = 10):
        """autotune through a JSON cache keyed by (device, domain count and
        total cells rounded to powers of two, precision, boundary).  Returns
        (best, hit)."""
        best, hit = ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.lbm_batch_autotune_cached(self._h, os.fsencode(cache_path), steps, ctypes.byref(best),
                                                      ctypes.byref(hit)))
        return best.value, bool(hit.value)

    def read_f(self, d: int) -> np.ndarray:
        nx, ny = self.domains[d][:2]
        f = np.empty((9, ny, nx), np.float32)
        self._chk(self._lib.lbm_batch_read_f(self._h, d, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return f

    def moments(self, d: int):
        """Renderer (lagged) moments of domain d: (jx, jy, rho)."""
        nx, ny = self.domains[d][:2]
        out = [np.empty((ny, nx), np.float32) for _ in range(3)]
        self._chk(self._lib.lbm_batch_read_moments(
            self._h, d, *[o.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for o in out]))
        return tuple(out)
