"""FP64 D2Q9 solver on the GPU (lbm_f64_* in include/lbm_hip.h): the yardstick
for the error of the f32 and FP16-shifted grids at full size.  No reference
counterpart; its scheme is the textbook one with the reference step's
boundary rules."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N


class LBM64:
    def __init__(self, omega: float, x: int, y: int, u_in: float = 0.1, *, boundary: int = N.BOUNDARY_REFERENCE,
                 device: int = 0):
        self._lib = N.load()
        h = ctypes.c_void_p()
        N.check_f64(self._lib.lbm_f64_create(x, y, omega, u_in, boundary, device, ctypes.byref(h)))
        self._h = h
        self.x, self.y = x, y

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.lbm_f64_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mask(self, solid):
        m = np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(self.y, self.x))
        N.check_f64(self._lib.lbm_f64_set_mask(self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def step(self, n: int = 1):
        N.check_f64(self._lib.lbm_f64_step(self._h, n))

    def sync(self):
        N.check_f64(self._lib.lbm_f64_sync(self._h))

    def get_compute_num(self) -> int:
        n = ctypes.c_uint64()
        N.check_f64(self._lib.lbm_f64_step_count(self._h, ctypes.byref(n)))
        return n.value

    def macro(self):
        """(rho, ux, uy) of the current state, uy north-positive."""
        out = [np.empty((self.y, self.x), np.float64) for _ in range(3)]
        N.check_f64(self._lib.lbm_f64_macro(self._h, *[o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                                                       for o in out]))
        return tuple(out)

    def read_f(self) -> np.ndarray:
        f = np.empty((9, self.y, self.x), np.float64)
        N.check_f64(self._lib.lbm_f64_read_f(self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return f
