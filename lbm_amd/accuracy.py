"""Error of the f32 and FP16-shifted grids against the GPU FP64 solver.

    python -m lbm_amd.accuracy --nx 8192 --ny 4096 --checkpoints 1000,5000,20000

Workload: the benchmark's (200 seeded circles r 8..64, seed 2024, tau 0.53,
u_in 0.08, reference outlet) at the given size; all three solvers start from
feq(1, (u_in, 0)).  At each checkpoint one JSON line gives, per precision, the
relative L2 over fluid cells of the velocity and of rho - 1 against FP64 (f32,
FP16 of f - w, FP16 of f - feq(1, u_in)), and
the FP64 sums of mass and x/y momentum over fluid cells (the grid's own device
reduction) with their relative deviation from the FP64 solver's sums.  The
open boundaries exchange mass, so "drift" is measured against the FP64
trajectory, not against the initial mass.  No reference counterpart.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import LBM, LBM64, PRECISION_F16E, PRECISION_F16S, PRECISION_F32, LbmError, gen_circles


def rel_l2(a, b, fl):
    return float(np.sqrt(((a - b) ** 2)[fl].sum() / (b ** 2)[fl].sum()))


def run(nx, ny, checkpoints, tau=0.53, u_in=0.08, seed=2024, circles=200, r_min=8, r_max=64):
    mask = gen_circles(nx, ny, seed, circles, r_min, r_max)
    fl = mask != 1
    ref = LBM64(1.0 / tau, nx, ny, u_in)
    ref.set_mask(mask)
    grids = {}
    for name, p in (("f32", PRECISION_F32), ("f16s", PRECISION_F16S), ("f16e", PRECISION_F16E)):
        g = LBM(1.0 / tau, nx, ny, precision=p)
        g.mask = mask
        g.custom_speed(u_in)
        grids[name] = g
    done = 0
    out = []
    for cp in checkpoints:
        t0 = time.perf_counter()
        ref.step(cp - done)
        for g in grids.values():
            g.step(cp - done)
        ref.sync()
        done = cp
        rho64, ux64, uy64 = ref.macro()
        sums64 = np.array([rho64[fl].sum(), (rho64 * ux64)[fl].sum(), (rho64 * uy64)[fl].sum()])
        line = {"nx": nx, "ny": ny, "step": cp, "tau": tau, "u_in": u_in, "fp64_sums": sums64.tolist(),
                "fp64_finite": bool(np.isfinite(sums64).all())}
        for name, g in grids.items():
            rho, ux, uy, _ = g.fields()
            try:
                s = g.conservation()
            except LbmError:  # the health check refuses a non-finite state: BGK left its stable range
                line[name] = {"finite": False}
                continue
            num = ((ux - ux64) ** 2 + (uy - uy64) ** 2)[fl].sum()
            line[name] = {
                "u_rel_l2": float(np.sqrt(num / (ux64 ** 2 + uy64 ** 2)[fl].sum())),
                "rho_minus_1_rel_l2": rel_l2(rho.astype(np.float64) - 1.0, rho64 - 1.0, fl),
                "sums": [float(v) for v in s],
                "mass_rel_dev": float((s[0] - sums64[0]) / sums64[0]),
                "momentum_rel_dev": float(np.hypot(s[1] - sums64[1], s[2] - sums64[2]) / np.hypot(*sums64[1:])),
                "finite": bool(np.isfinite(s).all()),
            }
        line["seconds"] = round(time.perf_counter() - t0, 2)
        print(json.dumps(line), flush=True)
        out.append(line)
    for g in grids.values():
        g.close()
    ref.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=8192)
    ap.add_argument("--ny", type=int, default=4096)
    ap.add_argument("--checkpoints", default="1000,5000,20000")
    ap.add_argument("--tau", type=float, default=0.53)
    ap.add_argument("--u-in", type=float, default=0.08)
    a = ap.parse_args(argv)
    cps = sorted(int(c) for c in a.checkpoints.split(","))
    run(a.nx, a.ny, cps, tau=a.tau, u_in=a.u_in)


if __name__ == "__main__":
    main(sys.argv[1:])
