"""CPU oracle for the surrogate's convolution primitives — TEST INFRASTRUCTURE ONLY.

Plain numpy (float64) 3x3 convolution with zero padding on NHWC tensors and its
two gradients, written from the definitions:
    out[n, y, x, co] = b[co] + sum_{dy, dx, ci} in[n, y + dy - 1, x + dx - 1, ci] w[dy * 3 + dx, co, ci]
    d in  = conv of d out with w flipped in space and transposed in channels
    d w[t, co, ci] = sum_{n, y, x} in[n, y + dy - 1, x + dx - 1, ci] d out[n, y, x, co]
Pinned in tests/test_surrogate.py by central finite differences of a scalar
loss and by a brute-force loop on a tiny shape.
"""
import numpy as np


def pad1(a):
    return np.pad(a, ((0, 0), (1, 1), (1, 1), (0, 0)))


def conv3x3(x, w, b=None):
    """x (N,H,W,Ci), w (9,Co,Ci) -> (N,H,W,Co)."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    N, H, W, _ = x.shape
    xp = pad1(x)
    out = np.zeros((N, H, W, w.shape[1]))
    for t in range(9):
        dy, dx = divmod(t, 3)
        out += xp[:, dy:dy + H, dx:dx + W, :] @ w[t].T
    if b is not None:
        out += np.asarray(b, np.float64)
    return out


def conv3x3_dgrad(g, w):
    """g (N,H,W,Co) -> d in (N,H,W,Ci)."""
    g = np.asarray(g, np.float64)
    w = np.asarray(w, np.float64)
    N, H, W, _ = g.shape
    out = np.zeros((N, H, W, w.shape[2]))
    gp = pad1(g)
    for t in range(9):
        dy, dx = divmod(t, 3)
        # in[p + off(t)] received g[p] w[t]  <=>  d in[q] += g[q - off(t)] w[t]
        out += gp[:, 2 - dy:2 - dy + H, 2 - dx:2 - dx + W, :] @ w[t]
    return out


def conv3x3_wgrad(x, g):
    """x (N,H,W,Ci), g (N,H,W,Co) -> d w (9,Co,Ci)."""
    x = np.asarray(x, np.float64)
    g = np.asarray(g, np.float64)
    N, H, W, Ci = x.shape
    Co = g.shape[3]
    xp = pad1(x)
    dw = np.zeros((9, Co, Ci))
    for t in range(9):
        dy, dx = divmod(t, 3)
        xs = xp[:, dy:dy + H, dx:dx + W, :].reshape(-1, Ci)
        dw[t] = g.reshape(-1, Co).T @ xs
    return dw


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)
