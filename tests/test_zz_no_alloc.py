"""lbm_conservation_static: the FP64 sums without device allocation, through
the C ABI, against lbm_conservation on the same grids (same bits), on two calls
in a row with the second on a smaller grid, and its size limit.  No reference
counterpart."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sums(lib, g, name):
    out = np.zeros(3, np.float64)
    rc = getattr(lib, name)(g._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


def _grid(X, Y, precision, seed):
    from lbm_amd import LBM, gen_circles
    g = LBM(1.7, X, Y, precision=precision)
    g.mask = gen_circles(X, Y, seed, 6, 3, 9)
    g.custom_speed(0.07)
    return g


@pytest.mark.parametrize("precision", [0, 1])
def test_static_sums_equal_allocating_sums(precision):
    print(f"case: static conservation, precision {precision}", flush=True)
    from lbm_amd import _native as N
    lib = N.load()
    # a grid that has never allocated its reduction buffers: every device
    # allocation fails, the static entry point must not need one
    big = _grid(900, 400, precision, 11)
    big.step(13)
    os.environ["LBM_TEST_ALLOC_LIMIT"] = "0"
    try:
        rc, s1 = _sums(lib, big, "lbm_conservation_static")
        assert rc == N.LBM_OK, big._lib.lbm_last_error(big._h)
        rc_alloc, _ = _sums(lib, big, "lbm_conservation")
        assert rc_alloc == N.LBM_ENOMEM  # it would have allocated
    finally:
        os.environ.pop("LBM_TEST_ALLOC_LIMIT", None)
    rc, e1 = _sums(lib, big, "lbm_conservation")
    assert rc == N.LBM_OK
    assert np.array_equal(s1.view(np.uint64), e1.view(np.uint64))
    # second call: a smaller grid (fewer partial sums than the first call left)
    small = _grid(300, 90, precision, 12)
    small.step(8)
    rc, s2 = _sums(lib, small, "lbm_conservation_static")
    assert rc == N.LBM_OK
    rc, e2 = _sums(lib, small, "lbm_conservation")
    assert rc == N.LBM_OK
    assert np.array_equal(s2.view(np.uint64), e2.view(np.uint64))
    # the big grid again (S-form after a reset as well)
    big.step(4)
    rc, s3 = _sums(lib, big, "lbm_conservation_static")
    _, e3 = _sums(lib, big, "lbm_conservation")
    assert rc == N.LBM_OK and np.array_equal(s3.view(np.uint64), e3.view(np.uint64))
    big.custom_speed(0.05)
    rc, s4 = _sums(lib, big, "lbm_conservation_static")
    _, e4 = _sums(lib, big, "lbm_conservation")
    assert rc == N.LBM_OK and np.array_equal(s4.view(np.uint64), e4.view(np.uint64))
    big.close()
    small.close()


def test_static_sums_reject_grids_beyond_the_scratch():
    """A 1280 x 52429 f32 grid needs 5 * 52429 = 262145 partial sums, one more
    than the static array holds: LBM_EINVAL with a message naming the limit;
    the allocating entry point still works on it."""
    print("case: static conservation size limit", flush=True)
    from lbm_amd import LBM
    from lbm_amd import _native as N
    lib = N.load()
    g = LBM(1.5, 1280, 52429)
    g.step(2)
    rc, _ = _sums(lib, g, "lbm_conservation_static")
    assert rc == N.LBM_EINVAL
    msg = lib.lbm_last_error(g._h)
    msg = msg.decode() if isinstance(msg, bytes) else str(msg)
    assert "static scratch" in msg, msg
    rc, s = _sums(lib, g, "lbm_conservation")
    assert rc == N.LBM_OK and np.isfinite(s).all()
    g.close()
