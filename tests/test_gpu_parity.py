"""GPU parity: the HIP step (through the C ABI) against the oracle.

FP32 storage must equal refcpu (the faithful f32 restatement of the reference's
eight passes) BIT FOR BIT on every cell of the state, on the renderer's fields
and on the summary planes.  FP16-shifted storage and the zero-gradient outlet
are compared with oracle64 (FP64) at stated tolerances.
"""
import numpy as np
import pytest

import oracle
from oracle.np_ref import NpRef  # noqa: F401  (imported for parity with the CPU tests)

pytestmark = pytest.mark.gpu

X1, Y1 = 400, 160
OMEGA1 = 1.0 / (3 * 0.02 + 0.5)  # config 1: nu = 0.02 (tau 0.56)
U1 = 0.1


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def same_bits(a, b):
    return np.array_equal(bits(a), bits(b))


def circle_mask(X, Y, circles):
    yy, xx = np.mgrid[0:Y, 0:X]
    m = np.zeros((Y, X), np.uint8)
    m[0, :] = m[-1, :] = 1
    for cx, cy, r in circles:
        m[(xx - cx) ** 2 + (yy - cy) ** 2 <= r * r] = 1
    return m


def make_pair(X, Y, omega, mask=None, u=U1, boundary=0, precision=0):
    from lbm_amd import LBM
    g = LBM(omega, X, Y, precision=precision, boundary=boundary)
    r = oracle.RefCPU(X, Y, omega)
    if mask is not None:
        g.mask = mask
        r.set_mask(mask.astype(np.uint32))
    g.custom_speed(u)
    r.reset_equilibrium(1.0, u, 0.0)
    return g, r


def assert_state_equal(g, r, what=""):
    gs = g.read_f()
    rs = r.state()
    if not same_bits(gs, rs):
        d = np.argwhere(bits(gs) != bits(rs))
        pytest.fail(f"{what}: {len(d)} differing values, first at (k,y,x)={tuple(d[0])}: "
                    f"{gs[tuple(d[0])]!r} vs {rs[tuple(d[0])]!r}")


def assert_lagged_equal(g, r, what=""):
    jx, jy, rho = g.moments(lagged=True)
    rjx, rjy, rrho = r.density()
    assert same_bits(jx, rjx), what
    assert same_bits(jy, rjy), what
    assert same_bits(rho, rrho), what


def test_config1_bit_exact_2000_steps():
    """Config 1 (400x160, u 0.1, nu 0.02, circle r=16 at (80,80)), 2000 steps."""
    print("case: config1 400x160 2000 steps f32", flush=True)
    mask = circle_mask(X1, Y1, [(80, 80, 16)])
    g, r = make_pair(X1, Y1, OMEGA1, mask)
    for n in (1, 9, 90, 900, 1000):
        g.step(n)
        r.step(n)
        assert g.get_compute_num() == r.compute_step
        assert_lagged_equal(g, r, f"lagged fields after {r.compute_step}")
        assert_state_equal(g, r, f"state after {r.compute_step}")
    for kind in range(5):
        assert same_bits(g.calculate_summary(kind), r.summary(kind)), kind
    # and against FP64: the error is reported, at the f32 floor
    o = oracle.Oracle64(X1, Y1, OMEGA1, U1, 0)
    o.set_solid(mask)
    o.step(2000)
    rho, ux, uy = o.macro()
    grho, gux, guy, _ = g.fields()
    fl = mask == 0
    e_u = np.sqrt((((gux - ux) ** 2 + (guy - uy) ** 2)[fl]).sum() / ((ux ** 2 + uy ** 2)[fl]).sum())
    print(f"config1 f32 vs fp64: rel-L2(u) = {e_u:.3e}")
    assert e_u < 1e-4


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_geometry_paint_and_pulses(seed):
    """Random small grids: random obstacles (including mask values other than 0/1),
    painting between steps, omega changes, pulses, both sets differing."""
    rng = np.random.default_rng(100 + seed)
    X, Y = int(rng.integers(5, 90)), int(rng.integers(4, 50))
    print(f"case: random seed={seed} {X}x{Y}", flush=True)
    omega = float(rng.uniform(0.6, 1.95))
    circles = [(int(rng.integers(1, X)), int(rng.integers(0, Y)), int(rng.integers(1, 5))) for _ in range(3)]
    mask = circle_mask(X, Y, circles)
    g, r = make_pair(X, Y, omega, mask, u=float(rng.uniform(0.0, 0.1)))
    assert_lagged_equal(g, r, "after reset (pre-collision moments without rest)")
    for it in range(6):
        n = int(rng.integers(1, 25))
        g.step(n)
        r.step(n)
        assert_state_equal(g, r, f"it {it}")
        assert_lagged_equal(g, r, f"it {it}")
        # paint a few cells, with repeats (last writer wins) and values 0/1/2
        cells = rng.integers(0, X * Y, 6)
        vals = rng.integers(0, 3, 6)
        pairs = np.stack([np.concatenate([cells, cells[:2]]), np.concatenate([vals, 1 - np.minimum(vals[:2], 1)])],
                         axis=1).ravel().astype(np.uint32)
        g.paint(pairs)
        assert r.paint(pairs) == 0
        assert np.array_equal(g.mask, (r.mask() == 1).astype(np.uint8))
        if it % 2:
            omega = float(rng.uniform(0.6, 1.95))
            g.update_omega_buffer(omega)
            r.set_omega(omega)
    g.step(5)
    r.step(5)
    assert_state_equal(g, r, "final")
    for kind in range(5):
        assert same_bits(g.calculate_summary(kind), r.summary(kind)), kind


@pytest.mark.parametrize("k", list(range(9)))
def test_single_cell_presets(k):
    """single_cell(k) pulses (lbm.rs:1486-1496) then steps; the fields are not
    reset by single_cell (no pre-collision pass)."""
    print(f"case: single_cell k={k}", flush=True)
    X, Y = 48, 20
    g, r = make_pair(X, Y, 1.25, None)
    g.step(3)
    r.step(3)
    g.single_cell(k)
    r.single_cell(k)
    assert g.get_compute_num() == 0
    assert_lagged_equal(g, r, "fields survive single_cell")
    assert_state_equal(g, r, "after single_cell")
    g.step(17)
    r.step(17)
    assert_state_equal(g, r, "after 17 steps")
    assert_lagged_equal(g, r, "after 17 steps")


def test_pure_streaming_stream_demo():
    """StreamDemo (tutorial.ts:129-134): nu huge -> omega ~ 0; an N pulse of 4.0 at
    (3X/4, Y-2) travels north one row per step."""
    X, Y = 40, 24
    nu = 1e7
    omega = np.float32(1.0) / (np.float32(3.0) * np.float32(nu) + np.float32(0.5))
    g, r = make_pair(X, Y, float(omega), None)
    g.update_viscosity(nu)
    assert abs(g.omega - omega) == 0
    g.single_cell(1)
    r.single_cell(1)
    r.set_omega(float(omega))
    for s in range(1, 8):
        g.step(1)
        r.step(1)
        assert_state_equal(g, r, f"stream demo step {s}")
    f = g.read_f()
    assert abs(f[1, Y - 2 - 7, 3 * X // 4] - 4.0) < 1e-5


def test_zero_omega_is_a_permutation():
    X, Y = 23, 11
    rng = np.random.default_rng(5)
    mask = circle_mask(X, Y, [(10, 5, 2)])
    g, r = make_pair(X, Y, 0.0, mask)
    f = (oracle.host_feq(0.05, 0.0, 1.0)[:, None, None] * (1 + 0.3 * rng.random((9, Y, X)))).astype(np.float32)
    g.write_f(f)
    r.set_set(0, f)
    r.set_set(1, f)
    r.compute_step = 0
    g_count = g.get_compute_num()
    assert g_count == 0
    for s in range(12):
        g.step(1)
        r.step(1)
    assert_state_equal(g, r, "omega 0")


def test_step_splitting_and_checkpoint_continuation():
    """N+N steps == 2N steps; save after N, load into a fresh grid, N more: same bits."""
    print("case: checkpoint 200x80", flush=True)
    from lbm_amd import LBM
    X, Y = 200, 80
    mask = circle_mask(X, Y, [(40, 40, 9), (120, 30, 6)])
    a, _ = make_pair(X, Y, 1.9, mask)
    b, _ = make_pair(X, Y, 1.9, mask)
    a.step(150)
    for _ in range(150):
        b.step(1)
    assert same_bits(a.read_f(), b.read_f())
    blob = a.save_state()
    a.step(150)
    c = LBM(1.9, X, Y)
    c.load_state(blob)
    assert c.get_compute_num() == 150
    c.step(150)
    assert same_bits(a.read_f(), c.read_f())
    assert same_bits(a.moments(True)[2], c.moments(True)[2])
    # read_f between steps does not perturb the trajectory
    d, _ = make_pair(X, Y, 1.9, mask)
    for _ in range(3):
        d.step(100)
        d.read_f()
    assert same_bits(a.read_f(), d.read_f())


def test_conservation_reduction_matches_cpu_sum():
    X, Y = 301, 77
    rng = np.random.default_rng(9)
    mask = circle_mask(X, Y, [(100, 40, 10)])
    g, _ = make_pair(X, Y, 1.7, mask)
    g.step(123)
    jx, jy, rho = g.moments(lagged=False)
    fl = g.mask != 1
    ref = np.array([rho[fl].astype(np.float64).sum(), jx[fl].astype(np.float64).sum(),
                    jy[fl].astype(np.float64).sum()])
    got = g.conservation()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-9)
    # non-lagged moments are those of the current state (moments of read_f)
    mjx, mjy, mrho = oracle.moments_ref_order(g.read_f())
    assert same_bits(mrho, rho) and same_bits(mjx, jx) and same_bits(mjy, jy)
    del rng


def test_nonfinite_state_is_reported():
    from lbm_amd import LbmError
    g, _ = make_pair(64, 32, 1.2, None)
    g.step(3)
    g.set_pulse(5, 10, 10, float("nan"))
    g.step(2)
    with pytest.raises(LbmError) as e:
        g.conservation()
    assert e.value.code == -4


def test_paint_errors_change_nothing():
    from lbm_amd import LbmError
    g, _ = make_pair(32, 16, 1.2, None)
    m0 = g.mask.copy()
    with pytest.raises(LbmError):
        g.paint(np.array([5, 1, 7], np.uint32))  # odd length
    with pytest.raises(LbmError):
        g.paint(np.array([5, 1, 32 * 16, 1], np.uint32))  # out of range
    g.paint(np.array([], np.uint32))  # empty: no-op
    assert np.array_equal(g.mask, m0)


def test_zero_gradient_outlet_vs_fp64():
    print("case: zero-gradient 300x100", flush=True)
    from lbm_amd import LBM, BOUNDARY_ZERO_GRADIENT
    X, Y, omega, u = 300, 100, 1.0 / 0.56, 0.08
    mask = circle_mask(X, Y, [(60, 50, 10)])
    g = LBM(omega, X, Y, boundary=BOUNDARY_ZERO_GRADIENT)
    g.mask = mask
    g.custom_speed(u)
    g.step(1500)
    o = oracle.Oracle64(X, Y, omega, u, 1)
    o.set_solid(mask)
    o.set_equilibrium(1.0, u, 0.0)
    o.step(1500)
    rho, ux, uy = o.macro()
    grho, gux, guy, _ = g.fields()
    fl = mask == 0
    e_u = np.sqrt((((gux - ux) ** 2 + (guy - uy) ** 2)[fl]).sum() / ((ux ** 2 + uy ** 2)[fl]).sum())
    print(f"zero-gradient f32 vs fp64 rel-L2(u) = {e_u:.3e}")
    assert e_u < 1e-4


def test_f16s_codec_round_trip():
    from lbm_amd import LBM, PRECISION_F16S
    X, Y = 70, 30
    rng = np.random.default_rng(1)
    g = LBM(1.2, X, Y, precision=PRECISION_F16S)
    w = np.array([1 / 36, 1 / 9, 1 / 36, 1 / 9, 4 / 9, 1 / 9, 1 / 36, 1 / 9, 1 / 36], np.float32)
    f = (w[:, None, None] * (1 + 0.2 * rng.standard_normal((9, Y, X)))).astype(np.float32)
    g.write_f(f)
    got = g.read_f()
    exp = (f - w[:, None, None]).astype(np.float16).astype(np.float32) + w[:, None, None]
    assert same_bits(got, exp)


@pytest.mark.parametrize("u_in", [0.0, 0.08, -0.05])
def test_f16e_codec_round_trip(u_in):
    """F16E stores binary16 of f - feq(1, (u_in, 0)): per population the shift
    d_k = feq_k(1, u_in) - w_k (computed in double, rounded to a multiple of
    2^-20) is taken off
    in f32 after w_k, then rounded to nearest even; decoding adds d_k, then
    w_k.  At u_in = 0 it is the F16S format bit for bit."""
    from lbm_amd import LBM, PRECISION_F16E
    X, Y = 70, 30
    rng = np.random.default_rng(2)
    g = LBM(1.2, X, Y, precision=PRECISION_F16E, u_init=u_in)
    w64 = np.array([1 / 36, 1 / 9, 1 / 36, 1 / 9, 4 / 9, 1 / 9, 1 / 36, 1 / 9, 1 / 36])
    cx = np.array([-1, 0, 1, -1, 0, 1, -1, 0, 1], np.float64)
    us = float(np.float32(u_in))  # the grid holds the inflow speed as f32
    cu = cx * us
    d = w64 * (3 * cu + 4.5 * cu * cu - 1.5 * us * us)
    d = (np.rint(d * 2.0 ** 20) / 2.0 ** 20).astype(np.float32)[:, None, None]  # multiples of 2^-20
    w = w64.astype(np.float32)[:, None, None]
    f = (w * (1 + 0.2 * rng.standard_normal((9, Y, X)))).astype(np.float32)
    g.write_f(f)
    got = g.read_f()
    exp = (((f - w) - d).astype(np.float16).astype(np.float32) + d) + w
    assert same_bits(got, exp), f"{int((bits(got) != bits(exp)).sum())} values differ"
    if u_in == 0.0:
        assert same_bits(got, (f - w).astype(np.float16).astype(np.float32) + w)


@pytest.mark.parametrize("steps", [1000])
def test_f16s_error_vs_fp64_config1(steps):
    """FP16-shifted storage, FP32 collide: velocity error vs FP64 on config 1."""
    from lbm_amd import LBM, PRECISION_F16S
    mask = circle_mask(X1, Y1, [(80, 80, 16)])
    g = LBM(OMEGA1, X1, Y1, precision=PRECISION_F16S)
    g.mask = mask
    g.custom_speed(U1)
    g.step(steps)
    o = oracle.Oracle64(X1, Y1, OMEGA1, U1, 0)
    o.set_solid(mask)
    o.step(steps)
    rho, ux, uy = o.macro()
    grho, gux, guy, _ = g.fields()
    fl = mask == 0
    e_u = np.sqrt((((gux - ux) ** 2 + (guy - uy) ** 2)[fl]).sum() / ((ux ** 2 + uy ** 2)[fl]).sum())
    print(f"config1 f16s vs fp64 after {steps}: rel-L2(u) = {e_u:.3e}")
    assert e_u < 1e-3


def test_full_size_properties_8192x4096():
    """Config 2 at its full size: properties that do not need the CPU oracle.
    The inlet column stays the exact f32 inflow equilibrium, one lbm_step(n)
    equals n lbm_step(1) calls bit for bit (checked through conservation sums and
    a sampled window of the state), and mass stays finite."""
    print("case: full size 8192x4096", flush=True)
    from lbm_amd import LBM, gen_circles
    X, Y, omega, u = 8192, 4096, 1.0 / 0.53, 0.08
    mask = gen_circles(X, Y, 2024, 200, 8, 64)
    a = LBM(omega, X, Y)
    a.mask = mask
    a.custom_speed(u)
    a.step(40)
    sa = a.conservation()
    b = LBM(omega, X, Y)
    b.mask = mask
    b.custom_speed(u)
    for _ in range(40):
        b.step(1)
    sb = b.conservation()
    assert np.array_equal(sa, sb)
    fa = a.read_f()
    fb = b.read_f()
    assert same_bits(fa[:, ::97, ::89], fb[:, ::97, ::89])
    feq = oracle.host_feq(u, 0.0, 1.0)
    col = fa[:, 1:-1, 0]
    assert same_bits(col, np.broadcast_to(feq[:, None], col.shape))
    assert np.isfinite(sa).all()


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("variant", list(range(10)))
def test_every_kernel_variant_gives_the_same_bits(precision, variant):
    """The step-kernel variants move memory differently and must not change a bit.
    Odd width: exercises partial vector groups, the outlet column and the slow path."""
    print(f"case: variant {variant} precision {precision}", flush=True)
    from lbm_amd import LBM
    X, Y = 203, 61
    mask = circle_mask(X, Y, [(50, 30, 8), (140, 20, 5), (202, 40, 3)])
    ref = LBM(1.8, X, Y, precision=precision)
    ref.mask = mask
    ref.custom_speed(0.09)
    ref.variant = 0
    ref.step(120)
    g = LBM(1.8, X, Y, precision=precision)
    g.mask = mask
    g.custom_speed(0.09)
    g.variant = variant
    g.step(120)
    assert same_bits(g.read_f(), ref.read_f())
    if precision == 0:
        r = oracle.RefCPU(X, Y, 1.8)
        r.set_mask(mask.astype(np.uint32))
        r.reset_equilibrium(1.0, 0.09, 0.0)
        r.step(120)
        assert_state_equal(g, r, f"variant {variant}")


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_every_kernel_variant_zero_gradient_outlet(precision):
    """Zero-gradient outlet: every step-kernel variant gives variant 0's bits
    (variant 0 in this mode is pinned against the FP64 restatement above)."""
    print(f"case: zero-gradient variants precision {precision}", flush=True)
    from lbm_amd import BOUNDARY_ZERO_GRADIENT, LBM
    X, Y = 203, 61
    mask = circle_mask(X, Y, [(50, 30, 8), (140, 20, 5), (202, 40, 3)])
    ref = None
    for variant in range(10):
        g = LBM(1.8, X, Y, precision=precision, boundary=BOUNDARY_ZERO_GRADIENT)
        g.mask = mask
        g.custom_speed(0.09)
        g.variant = variant
        g.step(120)
        f = g.read_f()
        g.close()
        if ref is None:
            ref = f
        else:
            assert same_bits(f, ref), f"variant {variant}"


def both_sets(g):
    """Both parity sets of the saved state as (2, ny, 9, nx) raw elements (the
    rows' padding, which holds no state, cut off)."""
    blob = np.frombuffer(g.save_state(), np.uint8)
    es = 4 if g.precision == 0 else 2
    pitch = (g.x + 63) // 64 * 64
    ps = (pitch * g.y + 255) // 256 * 256
    set_bytes = 9 * pitch * g.y * es
    hdr = blob.size - (2 * set_bytes + 17 * ps)
    sets = blob[hdr:hdr + 2 * set_bytes].view(np.uint32 if es == 4 else np.uint16)
    return sets.reshape(2, g.y, 9, pitch)[..., :g.x]


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("variant", [8, 9])
def test_two_step_variants_whole_state_and_paint_between_calls(variant, precision, boundary):
    """Two steps per launch: after every call of odd and even length, both sets
    (save_state), the current state and the renderer's moments equal the
    one-step kernel's bit for bit; a paint between calls then reads the other
    set on newly solid cells, which the last launch of a call must have written.
    2309 x 37: ten strips of 254 columns, row blocks cut by the grid's end."""
    print(f"case: two-step variant {variant} precision {precision} boundary {boundary}", flush=True)
    from lbm_amd import LBM
    X, Y = 2309, 37
    mask = circle_mask(X, Y, [(300, 18, 6), (1270, 10, 5), (2300, 20, 4), (254, 30, 3)])
    grids = []
    for v in (0, variant):
        g = LBM(1.75, X, Y, precision=precision, boundary=boundary)
        g.mask = mask
        g.custom_speed(0.09)
        g.variant = v
        grids.append(g)
    a, b = grids
    calls = [1, 7, 10, 2, 1, 5, 4]
    # paints after calls that launched the two-step kernel (7 and 10 steps) and
    # after one that did not (2 steps): newly solid cells read the other set
    paints = {1: [1000 + 15 * X, 1, 1001 + 15 * X, 1, 1800 + 20 * X, 1],
              2: [600 + 12 * X, 1, 1000 + 15 * X, 0, 2000 + 8 * X, 1],
              3: [1400 + 25 * X, 1, 601 + 12 * X, 1]}
    for i, n in enumerate(calls):
        a.step(n)
        b.step(n)
        sa, sb = both_sets(a), both_sets(b)
        if not np.array_equal(sa, sb):
            d = np.argwhere(sa != sb)
            pytest.fail(f"call {i} ({n} steps): {len(d)} differing values, first at (set,y,k,x)={tuple(d[0])}")
        assert same_bits(a.moments(True)[2], b.moments(True)[2])
        if i in paints:
            pairs = np.array(paints[i], np.uint32)
            a.paint(pairs)
            b.paint(pairs)
    assert same_bits(a.read_f(), b.read_f())
    if precision == 0 and boundary == 0:
        r = oracle.RefCPU(X, Y, 1.75)
        r.set_mask(mask.astype(np.uint32))
        r.reset_equilibrium(1.0, 0.09, 0.0)
        for i, n in enumerate(calls):
            r.step(n)
            if i in paints:
                assert r.paint(np.array(paints[i], np.uint32)) == 0
        assert_state_equal(b, r, f"two-step variant {variant}")
    a.close()
    b.close()


@pytest.mark.parametrize("X,Y", [(3, 3), (5, 4), (64, 9), (254, 17), (255, 8), (256, 3), (257, 6), (258, 5),
                                 (509, 12), (512, 25), (513, 4), (507, 10), (508, 9), (510, 19), (1017, 7),
                                 (1018, 33), (1024, 6), (1025, 9)])
def test_two_step_variants_small_and_boundary_sizes(X, Y):
    """Grids at and around the strip width (254 columns), the row block (8 or 16
    rows) and the row pitch (a multiple of 64, where the east pull is clamped at
    the last element of the row), down to the 3 x 3 minimum, both precisions and
    outlet modes: both sets equal the one-step kernel's bit for bit."""
    print(f"case: two-step small {X}x{Y}", flush=True)
    from lbm_amd import LBM
    circles = [(X // 2, Y // 2, max(1, min(X, Y) // 4))] if min(X, Y) >= 8 else []
    mask = circle_mask(X, Y, circles)
    for precision in (0, 1, 2):
        for boundary in (0, 1):
            for variant in (8, 9):
                grids = []
                for v in (0, variant):
                    g = LBM(1.6, X, Y, precision=precision, boundary=boundary)
                    g.mask = mask
                    g.custom_speed(0.07)
                    g.variant = v
                    g.step(9)
                    g.step(6)
                    grids.append(g)
                a, b = grids
                sa, sb = both_sets(a), both_sets(b)
                assert np.array_equal(sa, sb), (precision, boundary, variant)
                assert same_bits(a.moments(True)[0], b.moments(True)[0])
                a.close()
                b.close()


@pytest.mark.parametrize("X,Y,fill", [(254, 19, 0.25), (255, 17, 0.25), (509, 33, 0.3), (64, 9, 0.2),
                                      (1018, 21, 0.25), (300, 12, 0.0), (256, 13, 0.25), (257, 11, 0.25),
                                      (513, 9, 0.3)])
def test_two_step_variants_open_rows_and_random_obstacles(X, Y, fill):
    """No wall rows: the first and last rows are fluid, so the outlet's pulls wrap
    between rows at the grid's top and bottom (row 0's south-west pull wraps to
    column 0 of row 0, the last rows' to the row past the set, flagged), and
    random solid cells touch column 0, the outlet column and each other,
    enclosing some fluid cells.  Both sets after calls of odd and even length
    equal the one-step kernel's, and (f32, reference outlet) the oracle's."""
    print(f"case: two-step open rows {X}x{Y} fill {fill}", flush=True)
    from lbm_amd import LBM
    rng = np.random.default_rng(X * 1000 + Y)
    mask = (rng.random((Y, X)) < fill).astype(np.uint8)
    for precision in (0, 1, 2):
        for boundary in (0, 1):
            for variant in (8, 9):
                grids = []
                for v in (0, variant):
                    g = LBM(1.7, X, Y, precision=precision, boundary=boundary)
                    g.mask = mask
                    g.custom_speed(0.08)
                    g.variant = v
                    grids.append(g)
                a, b = grids
                for n in (7, 4):
                    a.step(n)
                    b.step(n)
                    sa, sb = both_sets(a), both_sets(b)
                    if not np.array_equal(sa, sb):
                        d = np.argwhere(sa != sb)
                        pytest.fail(f"{(precision, boundary, variant)} after {n}: {len(d)} differ, "
                                    f"first (set,y,k,x)={tuple(d[0])}")
                if precision == 0 and boundary == 0:
                    r = oracle.RefCPU(X, Y, 1.7)
                    r.set_mask(mask.astype(np.uint32))
                    r.reset_equilibrium(1.0, 0.08, 0.0)
                    r.step(11)
                    assert_state_equal(b, r, f"two-step variant {variant}, open rows")
                a.close()
                b.close()


def test_switching_between_one_and_two_step_variants():
    """Leaving a two-step variant frees its extra sets; switching back and forth
    between calls keeps the state equal to the one-step kernel's."""
    print("case: variant switching", flush=True)
    from lbm_amd import LBM
    X, Y = 300, 70
    mask = circle_mask(X, Y, [(60, 35, 9), (200, 20, 6)])
    a, _ = make_pair(X, Y, 1.7, mask)
    b, _ = make_pair(X, Y, 1.7, mask)
    for v, n in ((8, 7), (0, 5), (9, 6), (4, 3), (8, 4)):
        a.variant = v
        a.step(n)
        b.step(n)
        assert np.array_equal(both_sets(a), both_sets(b)), (v, n)
    a.close()
    b.close()


def test_autotune_leaves_the_state_untouched():
    from lbm_amd import LBM
    X, Y = 640, 200
    mask = circle_mask(X, Y, [(100, 100, 20)])
    a, _ = make_pair(X, Y, 1.7, mask)
    b, _ = make_pair(X, Y, 1.7, mask)
    a.step(33)
    b.step(33)
    best, ms = a.autotune(5)
    from lbm_amd._native import NUM_VARIANTS
    assert 0 <= best < NUM_VARIANTS and len(ms) == NUM_VARIANTS and all(m > 0 for m in ms)
    assert a.get_compute_num() == 33
    a.step(20)
    b.step(20)
    assert same_bits(a.read_f(), b.read_f())
    assert same_bits(a.moments(True)[0], b.moments(True)[0])


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("cmap", [0, 1, 2])
def test_render_colours_equal_colormap_restatement(cmap, precision):
    """lbm_render: the summary plane mapped to colours in the same pass equals
    the CPU restatement of color_map/*.wgsl bit for bit, for every summary and
    both storage formats."""
    print(f"case: render colormap {cmap} precision {precision}", flush=True)
    from lbm_amd import LBM
    from oracle import colormap_ref
    mask = circle_mask(X1, Y1, [(80, 80, 16)])
    g = LBM(OMEGA1, X1, Y1, precision=precision)
    g.mask = mask
    g.custom_speed(U1)
    g.step(300)
    for stat in range(5):
        rgb = g.render(stat, cmap)
        plane = g.calculate_summary(stat)
        exp = colormap_ref.colormap(cmap, plane, mask)
        assert same_bits(rgb, exp), f"stat {stat}"
    g.close()


def test_driver_loop_writes_frames(tmp_path):
    """The command-line driver: frames of N steps, rendered and written as PPM."""
    from lbm_amd import LBM, run
    rc = run.main(["--nx", "128", "--ny", "64", "--circle", "30,32,6", "--frames", "3", "--steps-per-frame", "5",
                   "--summary", "speed", "--colormap", "viridis", "--barrier", "curl",
                   "--out", str(tmp_path)])
    assert rc == 0
    frames = sorted(p.name for p in tmp_path.iterdir())
    assert frames == ["frame_0000.ppm", "frame_0001.ppm", "frame_0002.ppm"]
    data = (tmp_path / "frame_0002.ppm").read_bytes()
    assert data.startswith(b"P6 128 64 255\n") and len(data) == len(b"P6 128 64 255\n") + 128 * 64 * 3
    g = LBM(1.0 / 0.56, 128, 64)
    assert g.color_map == 2  # jet, the reference's default
    g.close()


@pytest.mark.parametrize("preset", ["curl_barrier", "chaos_barrier", "welcome_barrier"])
def test_barrier_presets_paint_their_shapes(preset):
    """curl/chaos/welcome_barrier (lbm.rs:1367-1480): the painted mask equals the
    starting mask with the shape's points set, and stepping it matches refcpu."""
    print(f"case: preset {preset}", flush=True)
    from lbm_amd import shapes
    g, r = make_pair(X1, Y1, OMEGA1)
    m0 = g.mask.copy()
    getattr(g, preset)()
    exp = m0.copy()
    for x, y, v in getattr(shapes, preset)(X1, Y1).points:
        exp[y, x] = 1 if v else 0
    got = g.mask
    assert np.array_equal(got, exp)
    r.set_mask(got.astype(np.uint32))
    g.step(50)
    r.step(50)
    assert_state_equal(g, r, preset)
    g.close()


def test_interactive_session_script(tmp_path):
    """The reference's controls driven by a command script (lib.rs:108-204,
    450-535): a barrier preset, a drawn line and its undo, step mode, pause."""
    print("case: interaction script", flush=True)
    from lbm_amd import shapes
    from lbm_amd.interaction import Session
    s = Session(X1, Y1, OMEGA1)
    g = s.lbm
    walls = g.mask.copy()
    assert s.execute("preset curl") is None
    assert s.execute("frame") is not None  # preset + 15 steps
    assert g.get_compute_num() == 15
    curl = walls.copy()
    for x, y, _ in shapes.curl_barrier(X1, Y1).points:
        curl[y, x] = 1
    assert np.array_equal(g.mask, curl)
    for c in ("draw-type line", "click 20 100", "click 60 120", "frame"):
        s.execute(c)
    line = curl.copy()
    for x, y, _ in shapes.Line((20, 100), (60, 120), X1, Y1).points:
        line[y, x] = 1
    assert np.array_equal(g.mask, line) and g.get_compute_num() == 30
    s.execute("undo")
    s.execute("frame")
    assert np.array_equal(g.mask, curl)
    # step mode: a redraw steps only after "step"; paused with no change renders nothing
    s.execute("step-mode on")
    s.execute("frame")
    assert g.get_compute_num() == 45  # the undo frame stepped 15
    s.execute("step")
    assert s.execute("frame") is not None and g.get_compute_num() == 46
    s.execute("step-mode off")
    s.execute("pause")
    assert s.execute("frame") is None and g.get_compute_num() == 46
    s.execute("colormap viridis")
    rgb = s.execute("frame")
    assert rgb is not None and g.get_compute_num() == 46 and g.color_map == 1
    s.execute(f"save {tmp_path / 'f.ppm'}")
    assert (tmp_path / "f.ppm").stat().st_size == len(b"P6 400 160 255\n") + X1 * Y1 * 3
    with pytest.raises(ValueError):
        s.execute("bogus 1")
    s.close()


def test_cpp_driver_matches_python_and_oracle(tmp_path):
    """The C++ host side (include/lbm_hip.hpp, lbm_amd/lbm_cli) on config 1:
    its FP64 sums equal the Python mirror's and its final state equals refcpu
    bit for bit."""
    import json
    import os
    import subprocess
    print("case: C++ driver config 1", flush=True)
    from lbm_amd import LBM
    cli = os.path.join(os.path.dirname(__file__), "..", "lbm_amd", "lbm_cli")
    dump = tmp_path / "f.bin"
    r = subprocess.run([cli, "--nx", "400", "--ny", "160", "--viscosity", "0.02", "--speed", "0.1",
                        "--circle", "80,80,16", "--frames", "2", "--steps-per-frame", "100", "--variant", "0",
                        "--dump-f", str(dump)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    frames = [json.loads(line) for line in r.stdout.splitlines()]
    assert [f["step"] for f in frames] == [100, 200]
    mask = circle_mask(X1, Y1, [(80, 80, 16)])
    omega = np.float32(1.0) / (np.float32(3.0) * np.float32(0.02) + np.float32(0.5))
    g = LBM(float(omega), X1, Y1)
    g.mask = mask
    g.custom_speed(U1)
    for f in frames:
        g.step(100)
        c = g.conservation()
        assert f["mass"] == c[0] and f["momentum"] == [c[1], c[2]]
    f9 = np.fromfile(dump, np.float32).reshape(9, Y1, X1)
    assert same_bits(f9, g.read_f())
    rr = oracle.RefCPU(X1, Y1, float(omega))
    rr.set_mask(mask.astype(np.uint32))
    rr.reset_equilibrium(1.0, U1, 0.0)
    rr.step(200)
    assert same_bits(f9, rr.state())
    g.close()
    # a barrier preset drawn by the C++ geometry equals the Python one
    r = subprocess.run([cli, "--nx", "400", "--ny", "160", "--barrier", "welcome", "--frames", "1",
                        "--steps-per-frame", "60", "--variant", "8", "--dump-f", str(dump)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = LBM(float(omega), X1, Y1)
    g.welcome_barrier()
    g.custom_speed(U1)
    g.step(60)
    assert same_bits(np.fromfile(dump, np.float32).reshape(9, Y1, X1), g.read_f())
    g.close()
    # binary16 storage centred on the free-stream equilibrium through the C++ driver
    r = subprocess.run([cli, "--nx", "400", "--ny", "160", "--viscosity", "0.02", "--speed", "0.1",
                        "--circle", "80,80,16", "--frames", "1", "--steps-per-frame", "100", "--variant", "8",
                        "--precision", "f16e", "--dump-f", str(dump)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = LBM(float(omega), X1, Y1, precision=2)
    g.mask = mask
    g.custom_speed(U1)
    g.variant = 8
    g.step(100)
    assert json.loads(r.stdout.splitlines()[0])["mass"] == g.conservation()[0]
    assert same_bits(np.fromfile(dump, np.float32).reshape(9, Y1, X1), g.read_f())
    g.close()


def test_cpp_mirror_api(tmp_path):
    """include/lbm_hip.hpp on the GPU: the reference's surface, summary and colour
    planes, every variant continuing one state bit for bit, and errors as
    lbm::Error with the C ABI's code (tests/cpp/api_check.cpp)."""
    import os
    import subprocess
    print("case: C++ mirror API", flush=True)
    root = os.path.join(os.path.dirname(__file__), "..")
    exe = tmp_path / "api_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", "api_check.cpp"), "-L", os.path.join(root, "lbm_amd"),
                    "-l:liblbm_hip.so", "-Wl,-rpath," + os.path.abspath(os.path.join(root, "lbm_amd")), "-o", str(exe)],
                   check=True, timeout=180)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


def test_run_script_mode(tmp_path, capsys):
    """python -m lbm_amd.run --script: the control commands from a file, one
    output line per rendered frame; a bad command stops with exit code 2."""
    print("case: run --script", flush=True)
    from lbm_amd import run
    script = tmp_path / "controls.txt"
    script.write_text("# preset, a stroke, undo, then step mode\n"
                      "preset chaos\nframe\n"
                      "press 10 10\nmove 30 20\nrelease\nframe\n"
                      "undo\nframe\n"
                      "step-mode on\nstep\nframe\n"
                      f"save {tmp_path / 'last.ppm'}\n")
    rc = run.main(["--nx", "128", "--ny", "64", "--script", str(script)])
    out = [line for line in capsys.readouterr().out.splitlines() if line.startswith("line ")]
    assert rc == 0
    assert [line.split(":")[0] for line in out] == ["line 3", "line 7", "line 9", "line 12"]
    assert out[-1].startswith("line 12: step 46 ")  # 3 x 15 steps, then one in step mode
    assert (tmp_path / "last.ppm").exists()
    bad = tmp_path / "bad.txt"
    bad.write_text("frame\nexplode\n")
    assert run.main(["--nx", "64", "--ny", "32", "--script", str(bad)]) == 2
