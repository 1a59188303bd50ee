"""The GPU FP64 solver against the CPU FP64 oracle (bit for bit), and the
FP16-shifted / f32 grids' error against it."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _mask(X, Y, seed):
    from lbm_amd import gen_circles
    return gen_circles(X, Y, seed, 3, 4, 12)


@pytest.mark.parametrize("mode", [0, 1])
def test_f64_solver_equals_oracle64(mode):
    """Same scheme, same operation order, IEEE double without contraction:
    the whole state must agree bit for bit after 300 steps."""
    print(f"case: f64 vs oracle64 mode {mode}", flush=True)
    from lbm_amd import LBM64
    X, Y, omega, u = 200, 80, 1.0 / 0.56, 0.1
    mask = _mask(X, Y, 5 + mode)
    g = LBM64(omega, X, Y, u, boundary=mode)
    g.set_mask(mask)
    g.step(300)
    o = oracle.Oracle64(X, Y, omega, u, mode)
    o.set_solid(mask)
    o.step(300)
    assert g.get_compute_num() == 300
    fg, fo = g.read_f(), o.state()
    assert np.array_equal(fg.view(np.uint64), fo.view(np.uint64))
    for a, b in zip(g.macro(), o.macro()):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    g.close()


def test_f16s_and_f32_error_against_f64_on_gpu():
    """Config 1 (400 x 160, one circle r 16 at (80, 80), tau 0.56, u 0.1) for
    1000 steps: the velocity error of both binary16 formats against FP64 stays
    under the 1e-3 budget and within 60x the f32 grid's (the f32 error is the
    floor); centring the storage on the free stream (F16E) at least halves the
    error of f - w (F16S)."""
    from lbm_amd import LBM, LBM64, PRECISION_F16E, PRECISION_F16S, PRECISION_F32
    X, Y, omega, u = 400, 160, 1.0 / 0.56, 0.1
    yy, xx = np.mgrid[0:Y, 0:X]
    mask = ((xx - 80) ** 2 + (yy - 80) ** 2 <= 16 * 16).astype(np.uint8)
    mask[0, :] = 1
    mask[-1, :] = 1
    ref = LBM64(omega, X, Y, u)
    ref.set_mask(mask)
    ref.step(1000)
    rho64, ux64, uy64 = ref.macro()
    fl = mask != 1
    errs = {}
    for p in (PRECISION_F32, PRECISION_F16S, PRECISION_F16E):
        g = LBM(omega, X, Y, precision=p)
        g.mask = mask
        g.custom_speed(u)
        g.step(1000)
        rho, ux, uy, _ = g.fields()
        num = ((ux - ux64) ** 2 + (uy - uy64) ** 2)[fl].sum()
        errs[p] = float(np.sqrt(num / (ux64 ** 2 + uy64 ** 2)[fl].sum()))
        g.close()
    print(f"velocity rel-L2 vs FP64 after 1000 steps: f32 {errs[PRECISION_F32]:.3e}, "
          f"f16s {errs[PRECISION_F16S]:.3e}, f16e {errs[PRECISION_F16E]:.3e}")
    assert errs[PRECISION_F32] < 1e-4
    for p in (PRECISION_F16S, PRECISION_F16E):
        assert errs[p] < 1e-3
        assert errs[p] < 60 * errs[PRECISION_F32]
    assert errs[PRECISION_F16E] < 0.5 * errs[PRECISION_F16S]


def test_accuracy_driver_small_grid():
    """The accuracy study end to end on a small grid: finite, f32 closer to FP64 than FP16-shifted."""
    from lbm_amd import accuracy
    lines = accuracy.run(256, 128, [200, 400], tau=0.6, u_in=0.05, circles=6, r_min=4, r_max=12)
    assert [ln["step"] for ln in lines] == [200, 400]
    for ln in lines:
        assert ln["fp64_finite"] and ln["f32"]["finite"] and ln["f16s"]["finite"]
        assert ln["f32"]["u_rel_l2"] < ln["f16s"]["u_rel_l2"] < 1e-2
        assert abs(ln["f32"]["mass_rel_dev"]) < 1e-4
