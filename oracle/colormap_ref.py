"""Colour maps of the reference renderer, restated in numpy f32 — TEST
INFRASTRUCTURE ONLY.

lbm-wgpu rewritten_shaders/color_map/jet.wgsl:13-47, viridis.wgsl:13-39 and
inferno.wgsl:13-39: c = clamp(s * v, -n, n) with (s, n) = (20, 4) for jet and
(15, 2) for the others; the block b = floor(c) picks control colours b + n and
b + n + 1, blended with right_weight = c - b as (1 - w) * lo + w * hi; c = n
takes the last colour; a solid cell (barrier == 1) is black.  Each operation is
an f32 operation in the shader's order.
"""
import numpy as np

INFERNO, VIRIDIS, JET = 0, 1, 2   # lbm.rs:20-24

TABLES = {
    INFERNO: [(0.98828125, 1.0, 0.64453125), (0.97265625, 0.55859375, 0.0390625),
              (0.73828125, 0.21875, 0.33203125), (0.34375, 0.06640625, 0.43359375), (0.0, 0.0, 0.01853125)],
    VIRIDIS: [(0.9921875, 0.90625, 0.1484375), (0.3671875, 0.7890625, 0.3828125),
              (0.1328125, 0.56640625, 0.55078125), (0.23046875, 0.32421875, 0.546875),
              (0.265625, 0.0078125, 0.33203125)],
    JET: [(0.0, 0.0, 0.5), (0.0, 0.0, 1.0), (0.0, 0.5, 1.0), (0.0, 1.0, 1.0), (0.5, 1.0, 0.5), (1.0, 1.0, 0.0),
          (1.0, 0.5, 0.0), (1.0, 0.0, 0.0), (0.5, 0.0, 0.0)],
}


def colormap(which, value, solid=None):
    """value (...,) f32 -> (..., 3) f32 colours."""
    f = np.float32
    tab = np.array(TABLES[which], np.float32)
    s, n = (f(20.0), 4) if which == JET else (f(15.0), 2)
    v = np.asarray(value, np.float32)
    c = np.minimum(np.maximum(s * v, f(-n)), f(n)).astype(np.float32)
    b = np.floor(c).astype(np.int64)
    w = (np.float32(-1.0) * b.astype(np.float32) + c).astype(np.float32)   # right_weight = -b + c
    top = b >= n
    lo = tab[np.clip(b + n, 0, 2 * n)]
    hi = tab[np.clip(b + n + 1, 0, 2 * n)]
    out = ((f(1.0) - w)[..., None] * lo + w[..., None] * hi).astype(np.float32)
    out[top] = tab[2 * n]
    if solid is not None:
        out[np.asarray(solid) == 1] = 0.0
    return out
