"""Learned 64-step surrogate for the D2Q9 step (no reference counterpart).

``UNet`` wraps the library's BF16-MFMA U-Net (lbm_unet_* in include/lbm_hip.h):
forward, masked MSE loss + gradients, Adam training steps.  ``fields_to_input``
and ``target_from`` define the normalised channels; ``Rollout`` (in
rollout.py) uses a trained network to jump a grid 64 steps at a time with
plain LBM corrector steps and a residual-gated fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N

S_RHO = 0.01   # density deviation scale
S_U = 0.1      # velocity scale


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def fields_to_input(mask, rho, ux, uy, s_rho: float = S_RHO, s_u: float = S_U) -> np.ndarray:
    """(..., H, W) fields -> (..., H, W, 4) network input; solid cells read as rest."""
    solid = np.asarray(mask) == 1
    x = np.stack([solid.astype(np.float32),
                  np.where(solid, 0.0, (rho - 1.0) / s_rho),
                  np.where(solid, 0.0, ux / s_u),
                  np.where(solid, 0.0, uy / s_u)], axis=-1)
    return np.ascontiguousarray(x, np.float32)


def target_from(mask, f0, f1, s_rho: float = S_RHO, s_u: float = S_U) -> np.ndarray:
    """Change of (rho, ux, uy) between two field triples, in network units."""
    solid = np.asarray(mask) == 1
    d = [np.where(solid, 0.0, (f1[0] - f0[0]) / s_rho),
         np.where(solid, 0.0, (f1[1] - f0[1]) / s_u),
         np.where(solid, 0.0, (f1[2] - f0[2]) / s_u)]
    return np.ascontiguousarray(np.stack(d, axis=-1), np.float32)


class UNet:
    def __init__(self, batch: int, height: int, width: int, seed: int = 0, device: int = 0):
        self._lib = N.load()
        h = ctypes.c_void_p()
        N.check_unet(self._lib.lbm_unet_create(batch, height, width, seed, device, ctypes.byref(h)))
        self._h = h
        self.batch, self.height, self.width = batch, height, width

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.lbm_unet_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        n = ctypes.c_size_t()
        N.check_unet(self._lib.lbm_unet_num_params(self._h, ctypes.byref(n)))
        return n.value

    def get_params(self) -> np.ndarray:
        p = np.empty(self.num_params, np.float32)
        N.check_unet(self._lib.lbm_unet_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        N.check_unet(self._lib.lbm_unet_set_params(self._h, _fp(p)))

    def _shape(self, a, c):
        a = np.ascontiguousarray(a, np.float32)
        assert a.shape == (self.batch, self.height, self.width, c), a.shape
        return a

    def forward(self, x) -> np.ndarray:
        x = self._shape(x, 4)
        y = np.empty((self.batch, self.height, self.width, 3), np.float32)
        N.check_unet(self._lib.lbm_unet_forward(self._h, _fp(x), _fp(y)))
        return y

    def loss_grad(self, x, target, weights):
        x, t = self._shape(x, 4), self._shape(target, 3)
        w = np.ascontiguousarray(weights, np.float32)
        loss = ctypes.c_float()
        g = np.empty(self.num_params, np.float32)
        N.check_unet(self._lib.lbm_unet_loss_grad(self._h, _fp(x), _fp(t), _fp(w), ctypes.byref(loss), _fp(g)))
        return loss.value, g

    def train_step(self, x, target, weights, lr: float) -> float:
        x, t = self._shape(x, 4), self._shape(target, 3)
        w = np.ascontiguousarray(weights, np.float32)
        loss = ctypes.c_float()
        N.check_unet(self._lib.lbm_unet_train_step(self._h, _fp(x), _fp(t), _fp(w), lr, ctypes.byref(loss)))
        return loss.value

    def time_forward(self, reps: int = 5) -> float:
        ms = ctypes.c_float()
        N.check_unet(self._lib.lbm_unet_time_forward(self._h, reps, ctypes.byref(ms)))
        return ms.value
