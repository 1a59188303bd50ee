"""Golden fixtures (tests/golden/d2q9_32x16.npz, written by make_golden.py).

CPU: the oracle reproduces them bit for bit, and a second restatement (numpy)
agrees.  GPU: the HIP step reproduces them bit for bit."""
import os

import numpy as np
import pytest

import oracle
from oracle.np_ref import NpRef

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "d2q9_32x16.npz"))
X, Y, OMEGA, U = 32, 16, 1.0 / 0.56, 0.1


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_oracle_reproduces_golden(oracle_lib):
    r = oracle.RefCPU(X, Y, OMEGA)
    r.set_mask(G["mask"].astype(np.uint32))
    r.reset_equilibrium(1.0, U, 0.0)
    r.step(10)
    assert np.array_equal(bits(r.state()), bits(G["state_10"]))
    r.step(90)
    assert np.array_equal(bits(r.state()), bits(G["state_100"]))
    jx, jy, rho = r.density()
    assert np.array_equal(bits(rho), bits(G["rho_100"]))
    assert np.array_equal(bits(r.summary(0)), bits(G["curl_100"]))


def test_numpy_restatement_reproduces_golden():
    n = NpRef(X, Y, OMEGA)
    n.mask = G["mask"].astype(np.uint32).ravel()
    n.fill(oracle.host_feq(U, 0.0, 1.0)[:, None])
    n.step(100)
    assert np.array_equal(bits(n.state()), bits(G["state_100"]))


def test_golden_sanity():
    s = G["state_100"].astype(np.float64)
    fluid = G["mask"] == 0
    rho = s.sum(0)
    assert np.all(np.abs(rho[fluid] - 1) < 0.15)  # early transient on a tiny grid
    assert np.isfinite(s).all()


@pytest.mark.gpu
def test_gpu_reproduces_golden():
    from lbm_amd import LBM
    g = LBM(OMEGA, X, Y)
    g.mask = G["mask"]
    g.custom_speed(U)
    g.step(10)
    assert np.array_equal(bits(g.read_f()), bits(G["state_10"]))
    jx, jy, rho = g.moments(lagged=True)
    assert np.array_equal(bits(rho), bits(G["rho_10"]))
    assert np.array_equal(bits(jx), bits(G["jx_10"]))
    g.step(90)
    assert np.array_equal(bits(g.read_f()), bits(G["state_100"]))
    assert np.array_equal(bits(g.calculate_summary(0)), bits(G["curl_100"]))
    p = LBM(1.25, X, Y)
    p.single_cell(2)
    p.step(7)
    assert np.array_equal(bits(p.read_f()), bits(G["pulse2_state_7"]))
