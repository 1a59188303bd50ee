"""Batched domains: generators (CPU) and per-domain bit-exactness against the
single-grid engine and the oracle (GPU)."""
import numpy as np
import pytest

import oracle

M64 = (1 << 64) - 1


def _sm(state):
    state = (state + 0x9E3779B97F4A7C15) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


class _Rng:
    def __init__(self, seed):
        self.s = seed

    def next(self):
        self.s, v = _sm(self.s)
        return v

    def below(self, m):
        return ((self.next() >> 32) * m) >> 32

    def unit(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def test_gen_batch_matches_restatement():
    from lbm_amd import gen_batch
    got = gen_batch(2024, 40)
    r = _Rng(2024)
    for nx, ny, om, u in got:
        enx = 128 + 32 * r.below(29)
        eny = 64 + 32 * r.below(15)
        tau = 0.52 + 0.28 * r.unit()
        eu = 0.02 + 0.08 * r.unit()
        assert (nx, ny) == (enx, eny)
        assert om == np.float32(1.0 / tau)
        assert u == np.float32(eu)
        assert 128 <= nx <= 1024 and nx % 32 == 0 and 64 <= ny <= 512 and ny % 32 == 0
        assert 1 / 0.8 - 1e-6 <= om <= 1 / 0.52 + 1e-6 and 0.02 <= u <= 0.1


def _py_shapes(nx, ny, seed, n_min, n_max):
    r = _Rng(seed)
    m = np.zeros((ny, nx), np.uint8)
    m[0, :] = m[-1, :] = 1
    yy, xx = np.mgrid[0:ny, 0:nx]
    count = n_min + r.below(n_max - n_min + 1)
    for _ in range(count):
        if r.below(2) == 0:
            rad = 3 + r.below(max(1, min(nx, ny) // 8))
            cx = rad + 1 + r.below(max(1, nx - rad - 1))
            cy = r.below(ny)
            m[((xx - cx) ** 2 + (yy - cy) ** 2 <= rad * rad) & (xx >= 1)] = 1
        else:
            w = 2 + r.below(max(1, nx // 8))
            h = 2 + r.below(max(1, ny // 4))
            x0 = 1 + r.below(max(1, nx - w - 1))
            y0 = r.below(max(1, ny - h))
            m[y0:y0 + h, x0:x0 + w] = 1
    return m


@pytest.mark.parametrize("nx,ny,seed", [(128, 64, 1), (1024, 512, 99), (352, 96, 7)])
def test_gen_shapes_matches_restatement(nx, ny, seed):
    from lbm_amd import gen_shapes
    got = gen_shapes(nx, ny, seed, 1, 12)
    assert np.array_equal(got, _py_shapes(nx, ny, seed, 1, 12))
    assert got[1:-1, 0].sum() == 0


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _small_batch():
    doms = [(128, 64, 1.9, 0.05), (203, 77, 1.6, 0.08), (64, 33, 1.3, 0.1), (257, 40, 1.85, 0.03),
            (96, 96, 1.5, 0.06), (130, 70, 1.95, 0.09)]
    from lbm_amd import gen_shapes
    masks = [gen_shapes(nx, ny, 500 + i, 1, 6) if min(nx, ny) >= 16 else None
             for i, (nx, ny, _, _) in enumerate(doms)]
    return doms, masks


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("variant", [0, 3, 5, 6, 7])
def test_batch_domains_equal_single_grids(precision, variant):
    """Every domain of the batch equals a single grid of its size bit for bit
    (and one equals the oracle) after calls of 1, 37 and 2 steps: one-step
    variants, and the two-step variants 6 and 7 (8- and 16-row chunks; domains
    of 33..257 cells across, so strips and chunks end inside and at the edges)."""
    print(f"case: batch vs single grids precision {precision} variant {variant}", flush=True)
    from lbm_amd import LBM, Batch
    doms, masks = _small_batch()
    b = Batch(doms, precision=precision)
    b.variant = variant
    for i, m in enumerate(masks):
        b.set_mask(i, m)
    b.step(1)
    b.step(37)   # graph replay for most of these
    b.step(2)
    singles = []
    for (nx, ny, om, u), m in zip(doms, masks):
        g = LBM(om, nx, ny, precision=precision, u_init=u)
        g.mask = m
        g.step(40)
        singles.append(g)
    for i, g in enumerate(singles):
        assert np.array_equal(bits(b.read_f(i)), bits(g.read_f())), f"domain {i}"
        for a, c in zip(b.moments(i), g.moments(lagged=True)):
            assert np.array_equal(bits(a), bits(c)), f"moments of domain {i}"
    if precision == 0:
        nx, ny, om, u = doms[1]
        r = oracle.RefCPU(nx, ny, om)
        r.set_mask(masks[1].astype(np.uint32))
        r.reset_equilibrium(1.0, u, 0.0)
        r.step(40)
        assert np.array_equal(bits(b.read_f(1)), bits(r.state()))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 2])
@pytest.mark.parametrize("variant", [0, 6, 7])
def test_batch_zero_gradient_domains_equal_single_grids(precision, variant):
    """Zero-gradient outlet (the outlet column's east pulls clamp instead of
    wrapping): every domain equals a single grid of that mode bit for bit, for a
    one-step variant and both two-step variants, after calls of 1, 37 and 2
    steps."""
    print(f"case: batch zero-gradient precision {precision} variant {variant}", flush=True)
    from lbm_amd import BOUNDARY_ZERO_GRADIENT, LBM, Batch
    doms, masks = _small_batch()
    b = Batch(doms, precision=precision, boundary=BOUNDARY_ZERO_GRADIENT)
    b.variant = variant
    for i, m in enumerate(masks):
        b.set_mask(i, m)
    for n in (1, 37, 2):
        b.step(n)
    for i, ((nx, ny, om, u), m) in enumerate(zip(doms, masks)):
        g = LBM(om, nx, ny, precision=precision, u_init=u, boundary=BOUNDARY_ZERO_GRADIENT)
        g.mask = m
        g.step(40)
        assert np.array_equal(bits(b.read_f(i)), bits(g.read_f())), f"domain {i}"
        for a, c in zip(b.moments(i), g.moments(lagged=True)):
            assert np.array_equal(bits(a), bits(c)), f"moments of domain {i}"
        g.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 2])
@pytest.mark.parametrize("variant", [6, 7])
def test_batch_open_rows_equal_single_grids(precision, variant):
    """Domains without wall rows (their outlets' pulls wrap between rows at the
    top and bottom, the last domain's into the padding past the set) equal
    single grids bit for bit in the two-step variants."""
    print(f"case: batch open rows precision {precision} variant {variant}", flush=True)
    from lbm_amd import LBM, Batch, gen_shapes
    doms, _ = _small_batch()
    masks = [gen_shapes(nx, ny, 900 + i, 1, 6, walls=False) for i, (nx, ny, _, _) in enumerate(doms)]
    b = Batch(doms, precision=precision)
    b.variant = variant
    for i, m in enumerate(masks):
        b.set_mask(i, m)
    for n in (1, 37, 2):
        b.step(n)
    for i, ((nx, ny, om, u), m) in enumerate(zip(doms, masks)):
        g = LBM(om, nx, ny, precision=precision, u_init=u)
        g.mask = m
        g.step(40)
        assert np.array_equal(bits(b.read_f(i)), bits(g.read_f())), f"domain {i}"
        g.close()
    b.close()


@pytest.mark.gpu
def test_batch_graph_variants_and_mask_change():
    from lbm_amd import Batch
    doms, masks = _small_batch()
    runs = []
    # two-step variants: a call of 25 then of 17 steps ends each call with an odd number of
    # two-step launches, so the second call replays graphs of the other set arrangement
    for graph, variant in [(True, 0), (False, 0), (True, 3), (True, 5), (False, 4), (True, 6), (False, 7),
                           (True, 7)]:
        b = Batch(doms)
        for i, m in enumerate(masks):
            b.set_mask(i, m)
        b.set_graph(graph)
        b.variant = variant
        b.step(25)
        m2 = masks[2].copy()
        m2[10:14, 20:24] = 1
        b.set_mask(2, m2)   # mid-run: flushes the pending stream first
        b.step(17)
        runs.append([b.read_f(i) for i in range(len(doms))])
        assert b.step_count == 42
    for other in runs[1:]:
        for a, c in zip(runs[0], other):
            assert np.array_equal(bits(a), bits(c))


@pytest.mark.gpu
def test_batch_autotune_restores_state():
    from lbm_amd import Batch
    doms, masks = _small_batch()
    a = Batch(doms)
    c = Batch(doms)
    for i, m in enumerate(masks):
        a.set_mask(i, m)
        c.set_mask(i, m)
    a.step(9)
    c.step(9)
    best, ms = a.autotune(4)
    assert 0 <= best < 8 and len(ms) == 8 and all(np.isfinite(ms))
    a.step(11)
    c.step(11)
    for i in range(len(doms)):
        assert np.array_equal(bits(a.read_f(i)), bits(c.read_f(i)))
