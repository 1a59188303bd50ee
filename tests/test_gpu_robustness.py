"""Error paths of the C ABI on the GPU: a failed allocation and a corrupt
checkpoint leave the grid stepping exactly as before.  The failures are forced
from the host (an allocation cap read by the library, a header field set out of
range); every kernel launched here runs on valid buffers."""
import os

import numpy as np
import pytest

from test_gpu_parity import circle_mask, same_bits

pytestmark = pytest.mark.gpu


class alloc_cap:
    """LBM_TEST_ALLOC_LIMIT: device requests above `limit` bytes fail the way a
    real out-of-memory request does."""

    def __init__(self, limit):
        self.limit = int(limit)

    def __enter__(self):
        os.environ["LBM_TEST_ALLOC_LIMIT"] = str(self.limit)

    def __exit__(self, *exc):
        os.environ.pop("LBM_TEST_ALLOC_LIMIT", None)


def _pair(X, Y, mask, precision=0):
    from lbm_amd import LBM
    out = []
    for _ in range(2):
        g = LBM(1.7, X, Y, precision=precision)
        g.mask = mask
        g.custom_speed(0.08)
        out.append(g)
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_failed_two_step_allocation_changes_nothing(precision):
    """set_variant(8) cannot allocate the two extra sets: it reports ENOMEM,
    keeps the variant, and the next steps succeed and equal a grid that never
    tried, bit for bit (a failed hipMalloc must not surface again as the error of
    a later launch)."""
    print(f"case: failed two-step allocation, precision {precision}", flush=True)
    from lbm_amd._native import LbmError, LBM_ENOMEM as ENOMEM
    X, Y = 700, 90
    mask = circle_mask(X, Y, [(120, 45, 11), (400, 30, 7)])
    a, b = _pair(X, Y, mask, precision)
    a.step(5)
    b.step(5)
    with alloc_cap(1 << 16):
        with pytest.raises(LbmError) as ei:
            a.variant = 8
        assert ei.value.code == ENOMEM
        assert a.variant == 0
        a.step(7)  # must not report the failed allocation
    b.step(7)
    assert same_bits(a.read_f(), b.read_f())
    assert same_bits(a.moments(True)[2], b.moments(True)[2])
    # the autotuner cannot snapshot the state under the cap: it fails, changing nothing
    with alloc_cap(1 << 16):
        with pytest.raises(LbmError) as ei:
            a.autotune(4)
        assert ei.value.code == ENOMEM
    assert a.variant == 0 and a.get_compute_num() == b.get_compute_num()
    a.step(3)
    b.step(3)
    assert same_bits(a.read_f(), b.read_f())
    # without the cap the two-step variant is set and stays exact
    a.variant = 8
    a.step(9)
    b.step(9)
    assert same_bits(a.read_f(), b.read_f())
    a.close()
    b.close()


def test_autotune_with_few_steps_skips_two_step_variants():
    """A tuning call of fewer than three steps cannot launch the two-step kernel,
    so those variants are not candidates (+inf) and their sets are not kept;
    the state is restored exactly."""
    print("case: autotune with 2 steps", flush=True)
    X, Y = 500, 60
    mask = circle_mask(X, Y, [(100, 30, 8)])
    a, b = _pair(X, Y, mask)
    a.step(4)
    b.step(4)
    best, ms = a.autotune(2)
    assert best < 8
    assert all(np.isinf(ms[v]) for v in range(8, len(ms)))
    assert np.all(np.isfinite(ms[:8]))
    b.variant = best
    a.step(6)
    b.step(6)
    assert same_bits(a.read_f(), b.read_f())
    a.close()
    b.close()


def test_corrupt_checkpoint_header_is_rejected():
    """A state image whose header names set 7 as current, or a non-finite
    omega, is LBM_EINVAL and leaves the grid untouched; a good image round-trips
    and carries the inflow speed of the grid it came from."""
    print("case: corrupt checkpoint header", flush=True)
    from lbm_amd import LBM
    from lbm_amd._native import LbmError, LBM_EINVAL as EINVAL
    X, Y = 300, 50
    mask = circle_mask(X, Y, [(60, 25, 6)])
    a, b = _pair(X, Y, mask)
    a.step(11)
    b.step(11)
    blob = bytearray(a.save_state())
    cur_off = 8 + 4 + 4 + 4 + 4 + 4  # magic, X, Y, precision, boundary, omega
    bad = bytearray(blob)
    bad[cur_off:cur_off + 4] = np.int32(7).tobytes()
    with pytest.raises(LbmError) as ei:
        a.load_state(bytes(bad))
    assert ei.value.code == EINVAL
    bad = bytearray(blob)
    bad[cur_off - 4:cur_off] = np.float32(np.nan).tobytes()
    with pytest.raises(LbmError) as ei:
        a.load_state(bytes(bad))
    assert ei.value.code == EINVAL
    a.step(6)
    b.step(6)
    assert same_bits(a.read_f(), b.read_f())
    # a fresh grid created at another inflow speed continues bit for bit
    c = LBM(1.7, X, Y, u_init=0.1)
    c.load_state(a.save_state())
    a.step(5)
    c.step(5)
    assert same_bits(a.read_f(), c.read_f())
    for g in (a, b, c):
        g.close()


def test_checkpoint_carries_the_inflow_speed():
    """The surrogate's input holds the grid's inflow speed: a jump on a grid
    restored from a checkpoint of a u_in = 0.06 flow into a grid created at
    u_in = 0.1 equals the jump on the original grid bit for bit."""
    print("case: checkpoint inflow speed", flush=True)
    from lbm_amd import LBM
    from lbm_amd.surrogate import UNet
    X, Y = 512, 64
    mask = circle_mask(X, Y, [(100, 30, 9)])
    a = LBM(1.6, X, Y)
    a.mask = mask
    a.custom_speed(0.06)
    a.step(20)
    c = LBM(1.6, X, Y, u_init=0.1)
    c.load_state(a.save_state())
    net = UNet(1, Y, X, seed=5)
    a.surrogate_jump(net, 64)
    c.surrogate_jump(net, 64)
    assert same_bits(a.read_f(), c.read_f())
    net.close()
    a.close()
    c.close()


def test_autotune_cache_selects_the_stored_variant(tmp_path):
    """lbm_autotune_cached / lbm_batch_autotune_cached: a miss tunes (state
    untouched), stores one key per (device, shape class, precision, boundary)
    and a grid of the same class then reads its variant without timing; the
    stored variant, not a fresh tuning, is what a hit selects (an entry edited
    by hand is obeyed); a two-step variant that cannot get its spare sets is a
    miss; the C++ driver reads and writes the same file."""
    import json
    import subprocess
    from lbm_amd import LBM, Batch
    print("case: autotune cache", flush=True)
    path = tmp_path / "tune.json"
    mask = circle_mask(400, 160, [(80, 80, 16)])
    g, ref = _pair(400, 160, mask)
    g.step(7)
    ref.step(7)
    best, hit = g.autotune_cached(str(path), 6)
    assert not hit and g.variant == best
    assert same_bits(g.read_f(), ref.read_f()) and g.get_compute_num() == 7
    data = json.loads(path.read_text())
    (key, entry), = data.items()
    assert "/grid/512x256/f32/ref/v10" in key and entry["variant"] == best and len(entry["ms"]) == 10
    # same class (300 x 200 rounds to 512 x 256): read, not timed
    h = LBM(1.7, 300, 200)
    assert h.autotune_cached(str(path), 6) == (best, True)
    # the file decides: an edited entry is what a hit selects
    other = 3 if best != 3 else 1
    data[key]["variant"] = other
    path.write_text(json.dumps(data))
    assert h.autotune_cached(str(path), 6) == (other, True) and h.variant == other
    # a two-step variant whose spare sets cannot be allocated is a miss: the
    # call tunes afresh, which under the same cap cannot snapshot the state
    # either, so it fails and changes nothing
    from lbm_amd._native import LbmError, LBM_ENOMEM as ENOMEM
    data[key]["variant"] = 8
    path.write_text(json.dumps(data))
    with alloc_cap(1 << 16):
        with pytest.raises(LbmError) as ei:
            h.autotune_cached(str(path), 6)
        assert ei.value.code == ENOMEM
    assert h.variant == other and json.loads(path.read_text())[key]["variant"] == 8
    assert h.autotune_cached(str(path), 6) == (8, True)
    # another precision is another key
    p2 = LBM(1.7, 400, 160, precision=2)
    assert not p2.autotune_cached(str(path), 6)[1]
    assert len(json.loads(path.read_text())) == 2
    for x in (g, ref, h, p2):
        x.close()
    # batches
    doms = [(128, 64, 1.6, 0.05), (256, 128, 1.7, 0.08), (200, 100, 1.8, 0.1)]
    b = Batch(doms)
    bb, bhit = b.autotune_cached(str(path), 4)
    assert not bhit and b.variant == bb
    b.close()
    b = Batch(doms)
    assert b.autotune_cached(str(path), 4) == (bb, True)
    b.close()
    assert any("/batch/4d/1Mcells/f32/ref/v8" in k for k in json.loads(path.read_text()))
    # the C++ driver: tuned and stored, then read
    cli = os.path.join(os.path.dirname(__file__), "..", "lbm_amd", "lbm_cli")
    cpath = tmp_path / "cli.json"
    for expect in ("tuned and stored in", "read from"):
        r = subprocess.run([cli, "--nx", "400", "--ny", "160", "--frames", "1", "--steps-per-frame", "10",
                            "--tune-cache", str(cpath)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0 and expect in r.stderr, r.stderr
