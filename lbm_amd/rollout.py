"""Surrogate + LBM-corrector rollout (no reference counterpart).

One cycle advances the simulated time by ``jump + correctors`` steps:

1. snapshot the grid's state on the device;
2. ``lbm_surrogate_jump``: the U-Net predicts rho and u ``jump`` steps ahead and
   every streamed fluid cell is re-equilibrated from the prediction;
3. mark the velocity, run ``correctors`` plain LBM steps and measure the
   relative velocity change they made (the residual: a prediction the LBM
   disagrees with relaxes hard);
4. if the residual exceeds ``threshold``, restore the snapshot and run the
   whole ``jump + correctors`` steps with the plain LBM instead (fallback).

Time left over at the end (fewer than ``jump + correctors`` steps) runs plain.
The plain LBM is the reference for the error: ``compare`` runs both from the
same state and reports the velocity error and the wall-clock speed-up.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .surrogate import D_RHO, D_U, S_RHO, S_U


@dataclass
class RolloutStats:
    steps: int = 0
    cycles: int = 0
    jumps: int = 0
    fallbacks: int = 0
    residuals: list = field(default_factory=list)


class Rollout:
    def __init__(self, grid, net, jump: int = 64, correctors: int = 8, threshold: float = 0.05,
                 s_rho: float = S_RHO, s_u: float = S_U, d_rho: float = D_RHO, d_u: float = D_U):
        self.grid, self.net = grid, net
        self.jump, self.correctors, self.threshold = jump, correctors, threshold
        self.scales = (s_rho, s_u, d_rho, d_u)

    def run(self, steps: int) -> RolloutStats:
        st = RolloutStats()
        g = self.grid
        per = self.jump + self.correctors
        while steps - st.steps >= per:
            g.snapshot()
            g.surrogate_jump(self.net, self.jump, *self.scales)
            g.velocity_mark()
            g.step(self.correctors)
            r = g.velocity_change()
            st.residuals.append(r)
            if not (r <= self.threshold):
                g.restore()
                g.step(per)
                st.fallbacks += 1
            else:
                st.jumps += 1
            st.cycles += 1
            st.steps += per
        if steps > st.steps:
            g.step(steps - st.steps)
            st.steps = steps
        return st


def velocity_error(a, b, mask) -> float:
    """rel-L2 of the velocity field a against b over fluid cells; a, b = (rho, ux, uy, curl)."""
    fl = np.asarray(mask) != 1
    num = ((a[1] - b[1]) ** 2 + (a[2] - b[2]) ** 2)[fl].sum()
    den = (b[1] ** 2 + b[2] ** 2)[fl].sum()
    return float(np.sqrt(num / den))


def compare(make_grid, net, steps: int, **kw):
    """Run the plain LBM and the rollout for `steps` from the same state; returns
    (speed-up, velocity error of the rollout against the plain LBM, stats,
    plain seconds, rollout seconds).  make_grid() builds a grid in the start
    state."""
    plain = make_grid()
    plain.sync()
    t0 = time.perf_counter()
    plain.step(steps)
    plain.sync()
    t_plain = time.perf_counter() - t0
    ref = plain.fields()
    mask = plain.mask
    plain.close()
    g = make_grid()
    g.sync()
    ro = Rollout(g, net, **kw)
    t0 = time.perf_counter()
    st = ro.run(steps)
    g.sync()
    t_ro = time.perf_counter() - t0
    err = velocity_error(g.fields(), ref, mask)
    g.close()
    return t_plain / t_ro, err, st, t_plain, t_ro


def main(argv=None) -> int:
    """Run the rollout on a seeded grid (the workload `bench.py` measures):

        python -m lbm_amd.rollout --cycles 20 [--spinup 200] [--weights PATH]
    """
    import argparse
    import os

    from . import LBM, gen_circles
    from .surrogate import UNet, load_params
    ap = argparse.ArgumentParser(description="surrogate + LBM-corrector rollout on a seeded grid")
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--circles", type=int, default=12)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--tau", type=float, default=0.6)
    ap.add_argument("--u-in", type=float, default=0.05)
    ap.add_argument("--spinup", type=int, default=200)
    ap.add_argument("--cycles", type=int, default=20)
    ap.add_argument("--threshold", type=float, default=0.05)
    ap.add_argument("--weights", default=os.path.join(os.path.dirname(__file__), "data", "surrogate_f2.npy"))
    a = ap.parse_args(argv)
    g = LBM(1.0 / a.tau, a.nx, a.ny)
    g.mask = gen_circles(a.nx, a.ny, a.seed, a.circles, 8, 32)
    g.custom_speed(a.u_in)
    g.step(a.spinup)
    net = UNet(1, a.ny // 2, a.nx // 2)
    net.set_params(load_params(a.weights))
    st = Rollout(g, net, threshold=a.threshold).run((64 + 8) * a.cycles)
    g.sync()
    print(f"cycles {st.cycles} jumps {st.jumps} fallbacks {st.fallbacks} steps {g.get_compute_num()}", flush=True)
    net.close()
    g.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
