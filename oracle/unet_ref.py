"""CPU reference of the surrogate U-Net — TEST INFRASTRUCTURE ONLY.

The same layer graph as lbm_amd/csrc/unet.hip (struct lbm_unet), restated in
numpy float64 on top of conv_ref.  ``round_bf16=True`` rounds the input, the
weights and every layer output to bfloat16 where the device does, so outputs
can be compared at the bf16 level; ``round_bf16=False`` is the exact model used
for finite-difference gradients.
"""
import numpy as np

from .conv_ref import bf16, conv3x3

# (kernel size, cin, cout, relu), in parameter order
SPECS = [(3, 32, 32, 1), (3, 32, 32, 1), (3, 32, 64, 1), (3, 64, 64, 1), (3, 64, 128, 1), (3, 128, 128, 1),
         (3, 128, 256, 1), (3, 256, 256, 1), (1, 256, 128, 1), (3, 256, 128, 1), (3, 128, 128, 1),
         (1, 128, 64, 1), (3, 128, 64, 1), (3, 64, 64, 1), (1, 64, 32, 1), (3, 64, 32, 1), (3, 32, 32, 1),
         (1, 32, 32, 0)]


def unpack(params):
    out, off = [], 0
    for ks, ci, co, relu in SPECS:
        n = ks * ks * ci * co
        w = params[off:off + n].reshape(ks * ks, co, ci)
        off += n
        b = params[off:off + co]
        off += co
        out.append((ks, w, b, relu))
    assert off == params.size
    return out


def _conv(x, layer, rnd):
    ks, w, b, relu = layer
    w = bf16(w) if rnd else np.asarray(w, np.float64)
    if ks == 3:
        y = conv3x3(x, w, b)
    else:
        y = x @ w[0].T + np.asarray(b, np.float64)
    if relu:
        y = np.maximum(y, 0.0)
    return y


def forward(params, xin, round_bf16=True, mutate=None):
    """xin (N, H, W, c <= 32) -> (N, H, W, 3); the input is zero-padded to 32 channels.

    ``mutate`` plants one wiring mistake, for tests that must tell a correct
    network from a miswired one: "skip_order" (skip tensor first in each
    concatenation), "skip_shift" (skip channels rotated by one), "up_shift"
    (upsampled tensors shifted by one column), "tap_flip" (the 3x3 taps of
    every convolution in reverse order)."""
    assert mutate in (None, "skip_order", "skip_shift", "up_shift", "tap_flip")
    L = unpack(np.asarray(params, np.float64))
    if mutate == "tap_flip":
        L = [(ks, w[::-1] if ks == 3 else w, b, relu) for ks, w, b, relu in L]
    r = (lambda a: bf16(a)) if round_bf16 else (lambda a: a)
    N, H, W, c = xin.shape
    x = np.zeros((N, H, W, 32))
    x[..., :c] = xin
    x = r(x)

    def c(i, t):
        return r(_conv(t, L[i], round_bf16))

    def pool(t):
        return r(0.25 * (t[:, 0::2, 0::2] + t[:, 0::2, 1::2] + t[:, 1::2, 0::2] + t[:, 1::2, 1::2]))

    def up(t):
        u = t.repeat(2, axis=1).repeat(2, axis=2)
        return np.roll(u, 1, axis=2) if mutate == "up_shift" else u

    def cat(a, skip):
        if mutate == "skip_shift":
            skip = np.roll(skip, 1, axis=-1)
        return np.concatenate([skip, a] if mutate == "skip_order" else [a, skip], axis=-1)

    e1 = c(1, c(0, x))
    e2 = c(3, c(2, pool(e1)))
    e3 = c(5, c(4, pool(e2)))
    e4 = c(7, c(6, pool(e3)))
    d3 = c(10, c(9, cat(c(8, up(e4)), e3)))
    d2 = c(13, c(12, cat(c(11, up(d3)), e2)))
    d1 = c(16, c(15, cat(c(14, up(d2)), e1)))
    out = _conv(d1, L[17], round_bf16)
    return out[..., :3]


def loss(params, xin, target, weights, round_bf16=False):
    y = forward(params, xin, round_bf16)
    w = np.asarray(weights, np.float64)
    return float((w[..., None] * (y - target) ** 2).sum() / w.sum())
