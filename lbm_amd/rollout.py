"""Surrogate + LBM-corrector rollout (no reference counterpart).

One cycle advances the simulated time by ``jump + correctors`` steps:
``lbm_surrogate_jump`` (the U-Net predicts rho and u ``jump`` steps ahead and
every streamed fluid cell is re-equilibrated from the prediction), then
``correctors`` plain LBM steps relax the re-equilibrated state.

The plain LBM judges the error.  Every ``check_every``-th cycle (the first one
included; 0: none, every cycle jumps, for measurements) is a check: the state is snapshotted on the device, the surrogate
cycle runs, its velocity field is marked, the snapshot is restored and the
plain LBM runs the same steps; the residual is the velocity rel-L2 between
the two (FP64 sums on the device) -- the cycle's true error -- and the run
continues from the plain state.  A residual above ``threshold`` (or a
non-finite prediction) is a fallback: the network is not trusted, and the
cycles up to the next check run the plain LBM.  With ``max_check_every``
above ``check_every`` the interval backs off: a check passed with a residual
at most half the threshold doubles the interval (up to ``max_check_every``),
one passed closer to the threshold keeps it, a fallback resets it to
``check_every``.  Measured one-cycle errors on the benchmark flows: the
trained network 0.35-1.2%, the network-free method (change predicted as zero)
4.5-7.5%.

Time left over at the end (fewer than a cycle's steps) runs plain.
``compare`` runs both from the same state and reports the velocity error and
the wall-clock speed-up.  ``method_floor`` measures what the method itself
loses with a perfect prediction.
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
    jumps: int = 0       # surrogate cycles kept
    checks: int = 0      # cycles run both ways (the plain result kept)
    fallbacks: int = 0   # checks above the threshold, and the plain cycles that follow them
    residuals: list = field(default_factory=list)  # per check: forecast error / persistence error
    errors: list = field(default_factory=list)     # per check: the forecast's velocity rel-L2


class Rollout:
    def __init__(self, grid, net, jump: int = 64, correctors: int = 8, threshold: float = 0.5,
                 check_every: int = 8, max_check_every: int | None = None, s_rho: float = S_RHO,
                 s_u: float = S_U, d_rho: float = D_RHO, d_u: float = D_U):
        if check_every < 0:
            raise ValueError("check_every must be >= 0 (0: never check, every cycle jumps)")
        if max_check_every is not None and max_check_every < check_every:
            raise ValueError("max_check_every must be >= check_every")
        self.grid, self.net = grid, net
        self.jump, self.correctors, self.threshold = jump, correctors, threshold
        self.check_every = check_every
        self.max_check_every = check_every if max_check_every is None else max_check_every
        self.scales = (s_rho, s_u, d_rho, d_u)

    @property
    def cycle_steps(self) -> int:
        return self.jump + self.correctors

    def run(self, steps: int) -> RolloutStats:
        st = RolloutStats()
        g = self.grid
        per = self.cycle_steps
        trusted = self.check_every == 0
        interval, next_check = self.check_every, 0
        while steps - st.steps >= per:
            if self.check_every and st.cycles == next_check:
                g.velocity_mark(0)  # the cycle's start: the persistence forecast
                g.snapshot()
                g.surrogate_jump(self.net, self.jump, *self.scales)
                g.step(self.correctors)
                g.velocity_mark(1)  # the surrogate's forecast
                g.restore()
                g.step(per)
                r, e = g.velocity_skill()
                st.residuals.append(r)
                st.errors.append(e)
                trusted = r <= self.threshold
                st.checks += 1
                if not trusted:
                    st.fallbacks += 1
                    interval = self.check_every
                elif r <= 0.5 * self.threshold:
                    interval = min(2 * interval, self.max_check_every)
                next_check = st.cycles + interval
            elif trusted:
                g.surrogate_jump(self.net, self.jump, *self.scales)
                g.step(self.correctors)
                st.jumps += 1
            else:
                g.step(per)
                st.fallbacks += 1
            st.cycles += 1
            st.steps += per
        if steps > st.steps:
            g.step(steps - st.steps)
            st.steps = steps
        return st


def method_floor(grid, zero_net, steps: int, jump: int = 64, correctors: int = 8) -> int:
    """The rollout with a perfect prediction: per cycle the plain LBM's own
    ``jump`` steps, then re-equilibration of every
    streamed fluid cell to its own rho, u (a zero-change jump of 0 steps with
    ``zero_net``, a network whose parameters are all zero), then
    ``correctors`` plain steps.  Its error against plain LBM is what the method
    loses by discarding the non-equilibrium part at each jump.  Returns the
    cycles run."""
    per = jump + correctors
    done, cycles = 0, 0
    while steps - done >= per:
        grid.step(jump)
        grid.surrogate_jump(zero_net, 0)
        grid.step(correctors)
        done += per
        cycles += 1
    if steps > done:
        grid.step(steps - done)
    return cycles


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
    ap.add_argument("--threshold", type=float, default=0.5)
    ap.add_argument("--check-every", type=int, default=8)
    ap.add_argument("--max-check-every", type=int, default=None, help="back off up to this interval (default: none)")
    ap.add_argument("--weights", default=os.path.join(os.path.dirname(__file__), "data", "surrogate_f2.npy"))
    a = ap.parse_args(argv)
    g = LBM(1.0 / a.tau, a.nx, a.ny)
    g.mask = gen_circles(a.nx, a.ny, a.seed, a.circles, 8, 32)
    g.custom_speed(a.u_in)
    g.step(a.spinup)
    net = UNet(1, a.ny // 2, a.nx // 2)
    net.set_params(load_params(a.weights))
    ro = Rollout(g, net, threshold=a.threshold, check_every=a.check_every, max_check_every=a.max_check_every)
    st = ro.run(ro.cycle_steps * a.cycles)
    g.sync()
    print(f"cycles {st.cycles} jumps {st.jumps} checks {st.checks} fallbacks {st.fallbacks} "
          f"steps {g.get_compute_num()}", flush=True)
    net.close()
    g.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
