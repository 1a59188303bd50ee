"""Learned 64-step surrogate for the D2Q9 step (no reference counterpart).

``UNet`` wraps the library's BF16-MFMA U-Net (lbm_unet_* in include/lbm_hip.h):
forward, masked MSE loss + gradients, Adam training steps.  ``fields_to_input``
and ``target_from`` define the normalised channels, pooled over F x F cells for
a network that runs at 1/F of the grid's resolution (F = 1 or 2);
``Rollout`` (in rollout.py) uses a trained network to jump a grid 64 steps at a
time with plain LBM corrector steps and a residual-gated fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N

S_RHO = 0.01   # input unit of rho - 1
S_U = 0.1      # input unit of the velocity
D_RHO = 0.01   # output unit of the change of rho over one jump
D_U = 0.003    # output unit of the change of the velocity over one jump
NET_IN = 8     # host input channels per pixel


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def pool_fields(mask, rho, ux, uy, factor: int = 1):
    """(..., H, W) cell fields -> (..., H/F, W/F) pixels of F x F cells: solid
    fraction, fluid-cell count and the fluid-cell means of rho, ux, uy (0 where
    a pixel has no fluid), as the grid's surrogate input kernel computes them.
    Returns (solid, nfluid, [rho, ux, uy])."""
    fl = (np.asarray(mask) != 1)
    F = factor

    def blocks(a):
        a = np.asarray(a, np.float32)
        if F == 1:
            return a
        sh = a.shape[:-2] + (a.shape[-2] // F, F, a.shape[-1] // F, F)
        return a.reshape(sh).sum(axis=(-3, -1))
    nf = blocks(fl)
    inv = np.where(nf > 0, 1.0 / np.maximum(nf, 1.0), 0.0).astype(np.float32)
    means = [blocks(np.where(fl, v, 0.0)) * inv for v in (rho, ux, uy)]
    return 1.0 - nf / (F * F), nf, means


def input_from_pooled(solid, nf, means, tau: float, u_in: float, s_rho: float = S_RHO,
                      s_u: float = S_U) -> np.ndarray:
    has = nf > 0
    x = np.zeros(solid.shape + (NET_IN,), np.float32)
    x[..., 0] = solid
    x[..., 1] = np.where(has, (means[0] - 1.0) / s_rho, 0.0)
    x[..., 2] = means[1] / s_u
    x[..., 3] = means[2] / s_u
    x[..., 4] = (tau - 0.5) / 0.2
    x[..., 5] = u_in / s_u
    return x


def target_from_pooled(nf, m0, m1, d_rho: float = D_RHO, d_u: float = D_U) -> np.ndarray:
    has = nf > 0
    d = [np.where(has, (m1[i] - m0[i]) / s, 0.0) for i, s in enumerate((d_rho, d_u, d_u))]
    return np.ascontiguousarray(np.stack(d, axis=-1), np.float32)


def fields_to_input(mask, rho, ux, uy, tau: float, u_in: float, factor: int = 1, s_rho: float = S_RHO,
                    s_u: float = S_U) -> np.ndarray:
    """(..., H, W) fields -> (..., H/F, W/F, NET_IN) network input."""
    solid, nf, means = pool_fields(mask, rho, ux, uy, factor)
    return input_from_pooled(solid, nf, means, tau, u_in, s_rho, s_u)


def target_from(mask, f0, f1, factor: int = 1, d_rho: float = D_RHO, d_u: float = D_U) -> np.ndarray:
    """Change of the pooled (rho, ux, uy) between two field triples, in output units."""
    _, nf, m0 = pool_fields(mask, *f0, factor)
    _, _, m1 = pool_fields(mask, *f1, factor)
    return target_from_pooled(nf, m0, m1, d_rho, d_u)


def loss_weights(mask, factor: int = 1) -> np.ndarray:
    """Per-pixel loss weight: the fluid fraction of the pixel."""
    solid, _, _ = pool_fields(mask, mask, mask, mask, factor)
    return (1.0 - solid).astype(np.float32)


def save_params(path, params) -> None:
    """Parameters as bfloat16 bit patterns (round to nearest even), .npy."""
    a = np.ascontiguousarray(params, np.float32).view(np.uint32)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16).astype(np.uint16)
    np.save(path, r)


def load_params(path) -> np.ndarray:
    """Parameters saved by save_params (or a plain float32 .npy)."""
    a = np.load(path, allow_pickle=False)
    if a.dtype == np.uint16:
        return (a.astype(np.uint32) << 16).view(np.float32)
    return a.astype(np.float32)


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
        x = self._shape(x, NET_IN)
        y = np.empty((self.batch, self.height, self.width, 3), np.float32)
        N.check_unet(self._lib.lbm_unet_forward(self._h, _fp(x), _fp(y)))
        return y

    def loss_grad(self, x, target, weights):
        x, t = self._shape(x, NET_IN), self._shape(target, 3)
        w = np.ascontiguousarray(weights, np.float32)
        loss = ctypes.c_float()
        g = np.empty(self.num_params, np.float32)
        N.check_unet(self._lib.lbm_unet_loss_grad(self._h, _fp(x), _fp(t), _fp(w), ctypes.byref(loss), _fp(g)))
        return loss.value, g

    def train_step(self, x, target, weights, lr: float) -> float:
        x, t = self._shape(x, NET_IN), self._shape(target, 3)
        w = np.ascontiguousarray(weights, np.float32)
        loss = ctypes.c_float()
        N.check_unet(self._lib.lbm_unet_train_step(self._h, _fp(x), _fp(t), _fp(w), lr, ctypes.byref(loss)))
        return loss.value

    def set_graph(self, enable: bool):
        """Replay the forward pass of surrogate jumps from a hipGraph (default) or not."""
        N.check_unet(self._lib.lbm_unet_set_graph(self._h, int(enable)))

    def time_forward(self, reps: int = 5) -> float:
        ms = ctypes.c_float()
        N.check_unet(self._lib.lbm_unet_time_forward(self._h, reps, ctypes.byref(ms)))
        return ms.value
