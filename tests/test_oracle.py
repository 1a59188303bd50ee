"""Pins for the CPU oracle (refcpu, np_ref, oracle64).  CPU only.

The reference has no tests, fixtures or golden vectors for this path, so the
oracle is pinned by what the mathematics and the WGSL source fix:
  * constants and the host equilibrium (set_equil, lbm.rs:611-643) against the
    textbook D2Q9 equilibrium and its moments;
  * the f32 fixed point of collision at the inflow state;
  * omega = 0 pure streaming as a pulse-transport known-answer rule derived from
    stream/*.wgsl (where a unit pulse goes, including walls, the outlet wrap and
    the out-of-range read);
  * per-cell collision conservation and closed-box mass conservation;
  * the viscosity: decay rate of the lowest Stokes shear mode;
  * refcpu == np_ref bit for bit (two independent restatements);
  * refcpu vs oracle64 at the f32 rounding level.
"""
import numpy as np
import pytest

import oracle
from oracle.np_ref import NpRef

CX = np.array([-1, 0, 1, -1, 0, 1, -1, 0, 1])
CY = np.array([-1, -1, -1, 0, 0, 0, 1, 1, 1])
W = np.array([1 / 36, 1 / 9, 1 / 36, 1 / 9, 4 / 9, 1 / 9, 1 / 36, 1 / 9, 1 / 36])


def feq64(rho, ux, uy_north):
    uy = -uy_north  # memory rows grow downwards
    cu = CX * ux + CY * uy
    return W * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (ux * ux + uy * uy))


def test_shader_constants_round_the_same_either_way():
    # WGSL constant folding of 1.0/9.0 etc. in f64 or in f32 gives the same f32
    for a, b in [(1, 9), (1, 36), (4, 9)]:
        assert np.float32(a / b) == np.float32(a) / np.float32(b)


@pytest.mark.parametrize("ux,uy,rho", [(0.1, 0.0, 1.0), (0.08, 0.0, 1.0), (0.0, 0.0, 1.0), (0.05, -0.03, 1.1)])
def test_host_feq_matches_textbook(oracle_lib, ux, uy, rho):
    """set_equil equals the textbook feq except in plane 7 (S), where lbm.rs:639
    writes `- uy_2` for `+ 4.5 uy^2`: the plane is low by rho * uy^2.  The
    reference only ever calls it with uy = 0 (lbm.rs:740, 1077, 1091, 1484),
    where the two agree bit for bit."""
    f = oracle.host_feq(ux, uy, rho).astype(np.float64)
    ref = feq64(rho, ux, uy)
    ref[7] -= rho * uy * uy
    np.testing.assert_allclose(f, ref, rtol=2e-6, atol=1e-7)
    f[7] += rho * uy * uy
    # moments of the corrected plane set: rho, momentum (north-positive uy), second moment
    assert abs(f.sum() - rho) < 1e-6
    assert abs((CX * f).sum() - rho * ux) < 1e-6
    assert abs((-CY * f).sum() - rho * uy) < 1e-6
    pxx = (CX * CX * f).sum()
    assert abs(pxx - rho * (ux * ux + 1 / 3)) < 1e-6


@pytest.mark.parametrize("u,omega", [(0.1, 1 / 0.56), (0.08, 1 / 0.53), (0.1, 1.25)])
def test_inflow_equilibrium_is_f32_fixed_point(oracle_lib, u, omega):
    r = oracle.RefCPU(16, 8, omega)
    r.reset_equilibrium(1.0, u, 0.0)
    before = r.get_set(0).copy()
    for _ in range(50):
        r.collide()  # parity stays 0: collide only
    after = r.get_set(0)
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))


def streamed(mask, X, Y, i):
    return mask[i] != 1 and i % X != 0 and i // X < Y - 1


def pulse_destinations(mask, X, Y, s, k):
    """Where a unit pulse in population k at flat index s lands after one step at
    omega = 0, per stream/*.wgsl:31-62 read as transport rather than as a pull:
    a streamed cell t = s + c_k takes it if s is not solid; if t is solid or out
    of range, the streamed cell s takes it back as the opposite population."""
    T = X * Y
    t = s + CX[k] + CY[k] * X  # flattened: no x wrap check, like the shaders
    out = []
    t_in = 0 <= t < T
    if t_in and mask[t] != 1:
        if streamed(mask, X, Y, t) and mask[s] != 1:
            out.append((t, k))
    else:
        if streamed(mask, X, Y, s):
            out.append((s, 8 - k))
    return out


@pytest.mark.parametrize("X,Y", [(6, 5), (7, 6)])
def test_pulse_transport_known_answers(oracle_lib, X, Y):
    T = X * Y
    rng = np.random.default_rng(X * 100 + Y)
    mask = np.zeros(T, np.uint32)
    mask[:X] = 1
    mask[(Y - 1) * X:] = 1
    # one interior obstacle cell and one free cell in the top wall row
    mask[2 * X + 3] = 1
    mask[1] = 0
    bg = oracle.host_feq(0.0, 0.0, 1.0)
    for s in range(T):
        for k in [0, 1, 2, 3, 5, 6, 7, 8]:
            r = oracle.RefCPU(X, Y, 0.0)
            r.set_mask(mask)
            base = np.repeat(bg[:, None], T, axis=1).reshape(9, Y, X).astype(np.float32)
            r.set_set(1, base)
            p0 = base.copy()
            amp = np.float32(0.25 + rng.random())
            p0.reshape(9, T)[k, s] += amp
            r.set_set(0, p0)
            r.step(1)
            got = r.get_set(1).reshape(9, T) - base.reshape(9, T)
            exp = np.zeros((9, T), np.float32)
            for (t, kk) in pulse_destinations(mask, X, Y, s, k):
                exp[kk, t] += (p0.reshape(9, T)[k, s] - base.reshape(9, T)[k, s])
            assert np.array_equal(got, exp), (s, k)


def test_outlet_wrap_and_out_of_range_read(oracle_lib):
    """The NW population of cell (X-1, Y-2) reads mask[X*Y] (se_nw_stream.wgsl:51):
    out of range, defined as solid, so it bounces back from SE."""
    X, Y = 6, 5
    r = oracle.RefCPU(X, Y, 0.0)
    r.reset_equilibrium(1.0, 0.0, 0.0)
    s1 = r.get_set(1).reshape(9, -1).copy()
    p0 = r.get_set(0).reshape(9, -1).copy()
    i = (Y - 2) * X + (X - 1)
    p0[8, i] += 1.0  # SE pulse at (X-1, Y-2)
    r.set_set(0, p0.reshape(9, Y, X))
    r.step(1)
    got = r.get_set(1).reshape(9, -1) - s1
    assert got[0, i] == np.float32(1.0)  # NW at the same cell
    assert np.count_nonzero(got) == 1
    # W at (X-1, y) pulls from column 0 of row y+1
    r = oracle.RefCPU(X, Y, 0.0)
    r.reset_equilibrium(1.0, 0.0, 0.0)
    p0 = r.get_set(0).reshape(9, -1).copy()
    p0[3, 2 * X + 0] += 1.0  # W pulse at (0, 2)
    r.set_set(0, p0.reshape(9, Y, X))
    r.step(1)
    got = r.get_set(1).reshape(9, -1) - s1
    assert got[3, 1 * X + X - 1] == np.float32(1.0)  # arrives at (X-1, 1)


def test_collision_conserves_per_cell(oracle_lib):
    rng = np.random.default_rng(3)
    X, Y = 16, 8
    r = oracle.RefCPU(X, Y, 1.7)
    base = oracle.host_feq(0.05, 0.02, 1.0)
    f = (base[:, None, None] * (1 + 0.1 * rng.standard_normal((9, Y, X)))).astype(np.float32)
    r.set_set(0, f)
    r.set_set(1, f)
    m0 = oracle.moments_ref_order(f)
    r.collide()
    g = r.get_set(0)
    m1 = oracle.moments_ref_order(g)
    for a, b in zip(m0, m1):
        np.testing.assert_allclose(a, b, rtol=0, atol=2e-6)


@pytest.mark.parametrize("omega", [0.7, 1.25, 1.95])
def test_closed_box_conserves_mass(oracle_lib, omega):
    X, Y = 24, 16
    rng = np.random.default_rng(11)
    r = oracle.RefCPU(X, Y, omega)
    mask = np.zeros((Y, X), np.uint32)
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = 1
    mask[6:9, 10:13] = 1
    r.set_mask(mask)
    base = oracle.host_feq(0.0, 0.0, 1.0)
    f = (base[:, None, None] * (1 + 0.05 * rng.standard_normal((9, Y, X)))).astype(np.float32)
    r.set_set(0, f)
    r.set_set(1, f)
    fluid = mask == 0
    m0 = r.state()[:, fluid].astype(np.float64).sum()
    r.step(200)
    m1 = r.state()[:, fluid].astype(np.float64).sum()
    # f32: only the rounding drift of the collision arithmetic (~1e-8 per step)
    assert abs(m1 - m0) / m0 < 200 * 2e-8
    # FP64: exact up to double rounding, which pins streaming + bounce-back
    o = oracle.Oracle64(X, Y, omega, 0.0, 1)
    o.set_solid(mask.astype(np.uint8))
    o.set_state(f.astype(np.float64))
    a0 = o.state()[:, fluid].sum()
    o.step(200)
    assert abs(o.state()[:, fluid].sum() - a0) / a0 < 1e-12


@pytest.mark.parametrize("which", ["refcpu", "oracle64"])
def test_viscosity_from_shear_wave_decay(oracle_lib, which):
    """Lowest no-slip Stokes mode u_x = A sin(pi (y - 1/2) / H), H = Y - 2 between
    half-way bounce-back walls, decays as exp(-nu (pi/H)^2 t), nu = (1/omega - 1/2)/3."""
    X, Y, omega, A = 512, 34, 1.25, 0.01
    H = Y - 2
    nu = (1 / omega - 0.5) / 3
    y = np.arange(Y)
    ux = A * np.sin(np.pi * (y - 0.5) / H)
    f = np.zeros((9, Y, X))
    for j in range(1, Y - 1):
        f[:, j, :] = feq64(1.0, ux[j], 0.0)[:, None]
    f[:, 0, :] = f[:, -1, :] = feq64(1.0, 0.0, 0.0)[:, None]
    mask = np.zeros((Y, X), np.uint8)
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = 1
    steps = 300
    if which == "refcpu":
        r = oracle.RefCPU(X, Y, omega)
        r.set_mask(mask.astype(np.uint32))
        r.set_set(0, f.astype(np.float32))
        r.set_set(1, f.astype(np.float32))
        r.step(steps)
        jx, jy, rho = oracle.moments_ref_order(r.state())
        u = (jx / rho).astype(np.float64)
    else:
        o = oracle.Oracle64(X, Y, omega, 0.0, 1)
        o.set_solid(mask)
        o.set_state(f)
        o.step(steps)
        _, u, _ = o.macro()
    prof = u[1:-1, X // 2]
    amp = np.dot(prof, np.sin(np.pi * (y[1:-1] - 0.5) / H)) / np.dot(
        np.sin(np.pi * (y[1:-1] - 0.5) / H), np.sin(np.pi * (y[1:-1] - 0.5) / H))
    expect = A * np.exp(-nu * (np.pi / H) ** 2 * steps)
    assert abs(amp / expect - 1) < 0.01


def _random_case(rng, X, Y):
    mask = np.zeros((Y, X), np.uint32)
    mask[0, :] = mask[-1, :] = 1
    yy, xx = np.mgrid[0:Y, 0:X]
    for _ in range(3):
        cx, cy, rr = rng.integers(2, X - 1), rng.integers(1, Y - 1), rng.integers(1, 4)
        mask[(xx - cx) ** 2 + (yy - cy) ** 2 <= rr * rr] = 1
    mask[rng.integers(0, Y, 6), rng.integers(0, X, 6)] = rng.integers(0, 3, 6)  # values 0, 1, 2
    base = oracle.host_feq(0.06, 0.01, 1.0)
    f0 = (base[:, None, None] * (1 + 0.02 * rng.standard_normal((9, Y, X)))).astype(np.float32)
    f1 = (base[:, None, None] * (1 + 0.02 * rng.standard_normal((9, Y, X)))).astype(np.float32)
    return mask, f0, f1


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_refcpu_equals_numpy_restatement(oracle_lib, seed):
    rng = np.random.default_rng(seed)
    X, Y = int(rng.integers(8, 40)), int(rng.integers(5, 24))
    omega = float(rng.uniform(0.5, 1.9))
    mask, f0, f1 = _random_case(rng, X, Y)
    r = oracle.RefCPU(X, Y, omega)
    r.set_mask(mask)
    r.set_set(0, f0)
    r.set_set(1, f1)
    n = NpRef(X, Y, omega)
    n.mask = mask.ravel().copy()
    n.sets[0] = f0.reshape(9, -1)
    n.sets[1] = f1.reshape(9, -1)
    r.step(37)
    n.step(37)
    for p in (0, 1):
        assert np.array_equal(r.get_set(p).view(np.uint32), n.sets[p].reshape(9, Y, X).view(np.uint32))
    jx, jy, rho = r.density()
    assert np.array_equal(jx.ravel().view(np.uint32), n.dens[0].view(np.uint32))
    assert np.array_equal(rho.ravel().view(np.uint32), n.dens[2].view(np.uint32))


def test_refcpu_vs_oracle64_config1(oracle_lib):
    """Config 1 (400x160, u 0.1, nu 0.02, circle r=16 at (80,80)), 2000 steps:
    the f32 restatement and the FP64 solver agree at the f32 rounding level."""
    X, Y, u = 400, 160, 0.1
    omega = 1.0 / (3 * 0.02 + 0.5)
    yy, xx = np.mgrid[0:Y, 0:X]
    mask = np.zeros((Y, X), np.uint32)
    mask[0, :] = mask[-1, :] = 1
    mask[(xx - 80) ** 2 + (yy - 80) ** 2 <= 16 * 16] = 1
    r = oracle.RefCPU(X, Y, omega)
    r.set_mask(mask)
    r.reset_equilibrium(1.0, u, 0.0)
    r.step(2000)
    o = oracle.Oracle64(X, Y, omega, u, 0)
    o.set_solid(mask.astype(np.uint8))
    o.step(2000)
    rho, ux, uy = o.macro()
    jx, jy, rr = oracle.moments_ref_order(r.state())
    fl = mask == 0
    e_u = np.sqrt(((jx / rr - ux)[fl] ** 2 + (jy / rr - uy)[fl] ** 2).sum() / (ux[fl] ** 2 + uy[fl] ** 2).sum())
    e_r = np.linalg.norm((rr - rho)[fl]) / np.linalg.norm((rho - 1)[fl])
    assert e_u < 1e-4, e_u
    assert e_r < 3e-3, e_r  # rel-L2 of (rho - 1): the denominator is ~1e-3 of rho


# ---- colour maps (oracle/colormap_ref.py) ----------------------------------------

@pytest.mark.parametrize("which", [0, 1, 2])
def test_colormap_is_continuous_and_clamped(which):
    """Piecewise-linear between control colours: continuous at every block
    boundary (a flipped weight or a shifted table fails), constant beyond the
    clamp (a wrong scale moves where that starts), black for solid cells."""
    from oracle import colormap_ref as cm
    s, n = (20.0, 4) if which == cm.JET else (15.0, 2)
    tab = np.array(cm.TABLES[which], np.float32)
    for k in range(-n, n + 1):
        at = cm.colormap(which, np.float32(k / s))
        assert np.allclose(at, tab[k + n], atol=1e-6)
        below = cm.colormap(which, np.float32((k - 1e-4) / s))
        if k > -n:
            assert np.allclose(below, tab[k + n], atol=2e-4)
    far = cm.colormap(which, np.array([-10.0, 10.0], np.float32))
    assert np.array_equal(far[0], tab[0]) and np.array_equal(far[1], tab[2 * n])
    mid = cm.colormap(which, np.float32(0.5 / s))        # halfway through block 0
    assert np.allclose(mid, 0.5 * (tab[n] + tab[n + 1]), atol=1e-6)
    assert not cm.colormap(which, np.array([0.01, 0.2], np.float32), np.array([1, 1])).any()


# ---- summary pass (refcpu.c ref_summary, summary_stats/*.wgsl) ----------------------

def _state_with_moments(jx, jy, rho):
    """Populations whose reference moments (pre_collision/*.wgsl) are exactly
    (jx, jy, rho): E - W = jx, N - S = jy (north-positive), every other moving
    population 0 and the rest population C = rho - E - W - N - S.  The inputs
    are dyadic with few bits, so every sum the shaders form is exact in f32."""
    Y, X = jx.shape
    f = np.zeros((9, Y, X), np.float64)
    f[5] = np.maximum(jx, 0)   # E
    f[3] = np.maximum(-jx, 0)  # W
    f[1] = np.maximum(jy, 0)   # N
    f[7] = np.maximum(-jy, 0)  # S
    f[4] = rho - f[5] - f[3] - f[1] - f[7]
    assert np.array_equal(f.astype(np.float32).astype(np.float64), f)
    return f.astype(np.float32)


def _fields_oracle(jx, jy, rho):
    """refcpu with density planes (jx, jy, rho): set 0 holds populations with
    those moments and one collide pass (step parity 0) computes them, rest
    population included (corner_collision.wgsl:24), into the density planes."""
    Y, X = jx.shape
    r = oracle.RefCPU(X, Y, 1.0)
    f = _state_with_moments(jx, jy, rho)
    r.set_set(0, f)
    r.set_set(1, f)
    r.collide()
    djx, djy, drho = r.density()
    # the moments the pass extracted are the prescribed ones, bit for bit
    assert np.array_equal(djx, jx.astype(np.float32))
    assert np.array_equal(djy, jy.astype(np.float32))
    assert np.array_equal(drho, rho.astype(np.float32))
    return r


def _vorticity_curl(a, b, Omega, rho):
    """Closed form of the reference's curl for the momentum field
    j = (a row + Omega row, b x + Omega x) (rows grow southwards, y north-positive
    = -row): curl.wgsl:49 is 10 * (central differences over 2 cells) / rho, i.e.
    20 (d jy/dx - d jx/dy) / rho, exact for a linear field:
    d jy/dx = b + Omega, d jx/dy = -(a + Omega)."""
    return 20.0 * ((b + Omega) + (a + Omega)) / rho


@pytest.mark.parametrize("a,b,Om", [(1 / 64, 0.0, 0.0), (0.0, 1 / 32, 0.0), (1 / 64, 1 / 32, 0.0),
                                    (0.0, 0.0, 1 / 128), (-3 / 64, 1 / 16, 0.0)])
def test_summary_curl_of_linear_momentum_fields(oracle_lib, a, b, Om):
    """Linear shear (jx = a row), a jy = b x field and solid-body rotation
    (jx = Om row, jy = Om x): the curl equals its closed form on every cell the
    pass writes, except the last column, whose east neighbour is the next row's
    column 0 (flattened index, curl.wgsl:49): there jy[i + 1] = b * 0 = 0.
    Where the pass returns early (x = 0, curl.wgsl:41; the last row,
    curl.wgsl:45) the output plane keeps what the previous pass wrote."""
    X, Y = 13, 9
    yy, xx = np.mgrid[0:Y, 0:X].astype(np.float64)
    jx = (a + Om) * yy
    jy = (b + Om) * xx
    rho = 1.0 + 0.25 * (xx % 3)  # 1, 1.25, 1.5
    r = _fields_oracle(jx, jy, rho)
    disp = r.summary(3)  # fills the output plane with the display density first
    curl = r.summary(0).astype(np.float64)
    inner = (slice(1, Y - 1), slice(1, X - 1))
    np.testing.assert_allclose(curl[inner], _vorticity_curl(a, b, Om, rho)[inner], rtol=2 ** -22, atol=0)
    # last column: the east neighbour is (0, row + 1), whose jy is 0
    xl = X - 1
    rows = np.arange(1, Y - 1)
    exp_last = 10.0 * (0.0 - (b + Om) * (xl - 1) + 2 * (a + Om)) / rho[rows, xl]
    np.testing.assert_allclose(curl[rows, xl], exp_last, rtol=2 ** -22, atol=1e-12)
    # skipped cells keep the previous pass's values
    assert np.array_equal(curl[:, 0], disp[:, 0].astype(np.float64))
    assert np.array_equal(curl[Y - 1, :], disp[Y - 1, :].astype(np.float64))


def test_summary_curl_sign_convention_against_a_vortex(oracle_lib):
    """A counter-clockwise vortex seen with y north-up (rows grow downwards)
    has positive vorticity: j = Om (-(y - yc), x - xc) with y = -row gives
    jx = Om (row - rc), jy = Om (x - xc), and a positive curl everywhere."""
    X, Y, Om = 11, 11, 1 / 64
    yy, xx = np.mgrid[0:Y, 0:X].astype(np.float64)
    jx, jy = Om * (yy - 5), Om * (xx - 5)
    r = _fields_oracle(jx, jy, np.ones((Y, X)))
    curl = r.summary(0)
    assert np.all(curl[1:-1, 1:-1] == np.float32(40 * Om))


@pytest.mark.parametrize("rho,expect", [(-1.0, -0.5), (0.0, -0.5), (1.0, 0.1), (2.0, 0.7), (6.0, 3.1),
                                        (7.0, 3.5), (10.0, 3.5)])
def test_summary_display_density(oracle_lib, rho, expect):
    """rho.wgsl:24: 4 clamp(0.15 rho, 0, 1) - 0.5 on every cell, both clamp
    edges included (0.15 rho < 0 for rho < 0; > 1 for rho > 6.67)."""
    X, Y = 8, 5
    r = _fields_oracle(np.zeros((Y, X)), np.zeros((Y, X)), np.full((Y, X), rho))
    out = r.summary(3)
    np.testing.assert_allclose(out, expect, rtol=0, atol=4e-7)


@pytest.mark.parametrize("jx,jy,expect", [(0.0, 0.0, -0.5), (0.0625, 0.0, -0.1875), (0.0, -0.125, 0.125),
                                          (0.0625, 0.125, None), (0.25, 0.25, 0.5), (-1.5, 2.0, 0.5)])
def test_summary_display_speed(oracle_lib, jx, jy, expect):
    """speed.wgsl:25: clamp(5 |j|, 0, 1) - 0.5 of the momentum (not the velocity:
    rho does not enter), clamped above at |j| = 0.2."""
    X, Y = 8, 5
    rho = np.full((Y, X), 2.0)  # a non-unit density: a division by rho would show
    r = _fields_oracle(np.full((Y, X), jx), np.full((Y, X), jy), rho)
    out = r.summary(4).astype(np.float64)
    if expect is None:
        expect = min(5 * np.hypot(jx, jy), 1.0) - 0.5
    np.testing.assert_allclose(out, expect, rtol=0, atol=2e-7)


def test_summary_velocity_planes_are_momenta(oracle_lib):
    """ux.wgsl:41 / uy.wgsl:41 copy the density bind group's jx, jy (momenta,
    jy north-positive) unchanged."""
    X, Y = 9, 6
    yy, xx = np.mgrid[0:Y, 0:X].astype(np.float64)
    jx, jy = (xx - 4) / 32, (3 - yy) / 16
    r = _fields_oracle(jx, jy, 2.0 + 0 * xx)
    assert np.array_equal(r.summary(1), jx.astype(np.float32))
    assert np.array_equal(r.summary(2), jy.astype(np.float32))
