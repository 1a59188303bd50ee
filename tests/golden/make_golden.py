"""Writes the committed golden fixtures in tests/golden/ from the oracle.

    python tests/golden/make_golden.py

The reference ships no fixtures or golden vectors for its simulation step, so
these are produced by the oracle's faithful f32 restatement (refcpu), whose own
ground truth is the closed-form and invariant tests in tests/test_oracle.py.
They pin regressions: a later change to the oracle or to the GPU step that
alters a single bit of these states fails tests/test_golden.py.

Case: a 32 x 16 grid, walls on rows 0 and 15, one circle of radius 3 at
(9, 7), omega 1/0.56, inflow 0.1, started from feq(1, (0.1, 0)); the state S_n
(9 planes), the renderer's density planes and the curl summary after 10 and 100
steps.  Plus a single_cell(2) pulse case after 7 steps at omega 1.25.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

X, Y = 32, 16
OMEGA = 1.0 / 0.56
U = 0.1


def mask32():
    yy, xx = np.mgrid[0:Y, 0:X]
    m = np.zeros((Y, X), np.uint32)
    m[0, :] = m[-1, :] = 1
    m[(xx - 9) ** 2 + (yy - 7) ** 2 <= 9] = 1
    return m


def main():
    oracle.build()
    out = {}
    r = oracle.RefCPU(X, Y, OMEGA)
    r.set_mask(mask32())
    r.reset_equilibrium(1.0, U, 0.0)
    done = 0
    for n in (10, 100):
        r.step(n - done)
        done = n
        jx, jy, rho = r.density()
        out[f"state_{n}"] = r.state()
        out[f"jx_{n}"], out[f"jy_{n}"], out[f"rho_{n}"] = jx, jy, rho
        out[f"curl_{n}"] = r.summary(0)
    p = oracle.RefCPU(X, Y, 1.25)
    p.single_cell(2)
    p.step(7)
    out["pulse2_state_7"] = p.state()
    out["mask"] = mask32().astype(np.uint8)
    np.savez_compressed(os.path.join(HERE, "d2q9_32x16.npz"), **out)
    print("wrote", os.path.join(HERE, "d2q9_32x16.npz"))


if __name__ == "__main__":
    main()
