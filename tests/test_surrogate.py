"""Surrogate primitives: the CPU convolution oracle is pinned by finite
differences and a brute-force loop (CPU); the MFMA kernels are checked against
it (GPU)."""
import ctypes

import numpy as np
import pytest

from oracle import conv_ref as cr


def test_conv_oracle_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 4, 5, 3))
    w = rng.standard_normal((9, 2, 3))
    b = rng.standard_normal(2)
    out = cr.conv3x3(x, w, b)
    ref = np.zeros_like(out)
    for y in range(4):
        for xx in range(5):
            for co in range(2):
                s = b[co]
                for t in range(9):
                    dy, dx = divmod(t, 3)
                    sy, sx = y + dy - 1, xx + dx - 1
                    if 0 <= sy < 4 and 0 <= sx < 5:
                        s += np.dot(x[0, sy, sx], w[t, co])
                ref[0, y, xx, co] = s
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_conv_oracle_gradients_by_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 4, 3))
    w = rng.standard_normal((9, 4, 3))
    r = rng.standard_normal((2, 5, 4, 4))  # loss = sum(r * conv(x, w))
    gx = cr.conv3x3_dgrad(r, w)
    gw = cr.conv3x3_wgrad(x, r)
    h = 1e-6
    for idx in [(0, 0, 0, 0), (1, 4, 3, 2), (0, 2, 1, 1), (1, 0, 3, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = ((r * cr.conv3x3(xp, w)).sum() - (r * cr.conv3x3(xm, w)).sum()) / (2 * h)
        assert abs(fd - gx[idx]) < 1e-6 * max(1, abs(fd))
    for idx in [(0, 0, 0), (4, 3, 2), (8, 1, 1), (5, 2, 0)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        fd = ((r * cr.conv3x3(x, wp)).sum() - (r * cr.conv3x3(x, wm)).sum()) / (2 * h)
        assert abs(fd - gw[idx]) < 1e-6 * max(1, abs(fd))


def test_bf16_rounding_helper():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 1e-3], np.float32)
    r = cr.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12


def _run(mode, N, H, W, Ci, Co, x, g, w, b):
    from lbm_amd import _native as Nn
    lib = Nn.load()
    fp = lambda a: None if a is None else np.ascontiguousarray(a, np.float32).ctypes.data_as(  # noqa: E731
        ctypes.POINTER(ctypes.c_float))
    size = {0: N * H * W * Co, 1: N * H * W * Ci, 2: 9 * Co * Ci}[mode]
    out = np.zeros(size, np.float32)
    keep = [np.ascontiguousarray(a, np.float32) if a is not None else None for a in (x, g, w, b)]
    rc = lib.lbm_conv3x3_test(mode, N, H, W, Ci, Co, *[fp(a) for a in keep],
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 0)
    assert rc == 0, lib.lbm_unet_last_error()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("N,H,W,Ci,Co", [(1, 4, 32, 32, 32), (2, 8, 64, 64, 32), (1, 6, 32, 32, 128),
                                          (2, 16, 32, 96, 64)])
def test_mfma_conv_matches_oracle(N, H, W, Ci, Co):
    print(f"case: conv {N}x{H}x{W} {Ci}->{Co}", flush=True)
    rng = np.random.default_rng(N * 1000 + H + Ci)
    x = cr.bf16(rng.standard_normal((N, H, W, Ci)))
    w = cr.bf16(0.1 * rng.standard_normal((9, Co, Ci)))
    b = rng.standard_normal(Co).astype(np.float32)
    g = cr.bf16(rng.standard_normal((N, H, W, Co)))
    # forward: bf16 inputs are exact, accumulation in f32
    got = _run(0, N, H, W, Ci, Co, x, None, w, b).reshape(N, H, W, Co)
    ref = cr.conv3x3(x, w, b)
    scale = np.sqrt(9 * Ci) * 0.1
    assert np.abs(got - ref).max() < 1e-5 * scale * 10
    # input gradient
    got = _run(1, N, H, W, Ci, Co, g, None, w, None).reshape(N, H, W, Ci)
    ref = cr.conv3x3_dgrad(g, w)
    assert np.abs(got - ref).max() < 1e-4 * np.sqrt(9 * Co)
    # weight gradient
    got = _run(2, N, H, W, Ci, Co, x, g, None, None).reshape(9, Co, Ci)
    ref = cr.conv3x3_wgrad(x, g)
    assert np.abs(got - ref).max() < 1e-4 * np.sqrt(N * H * W)
