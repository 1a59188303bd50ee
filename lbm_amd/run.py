"""Command-line driver loop: the reference's event loop without a window.

    python -m lbm_amd.run --nx 400 --ny 160 --viscosity 0.02 --speed 0.1 \\
        --circle 80,80,16 --frames 20 --steps-per-frame 15 --summary curl \\
        --colormap jet --out frames/

Each frame runs `steps-per-frame` steps (lib.rs:108-204: COMPUTE_PER_RENDER
steps per redraw, or one step in step mode), then the summary field through
the colour map (lbm.rs:1065-1074), optionally written as a PPM image, and
prints the step count, MLUPS of the frame and the FP64 mass.

    python -m lbm_amd.run --nx 400 --ny 160 --script controls.txt
    python -m lbm_amd.run --nx 400 --ny 160 --script -     # commands from stdin

runs the reference's interactive controls instead (lbm_amd.interaction:
pointer strokes, undo, presets, pause and step mode), one command per line.  Obstacles:
circles (--circle x,y,r, repeatable), a seeded set (--circles n) and/or a
barrier preset (--barrier, lbm.rs:1367-1480); fluid
presets as the reference's (lib.rs:155-164); viscosity as the slider's
omega = 1 / (3 nu + 1/2) (lib.rs:181-186).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import LBM, PRECISION_F16S, PRECISION_F32, ColorMap, SummaryStat, gen_circles, save_ppm

FLUID_PRESETS = {  # lib.rs:155-164
    "equilibrium": lambda g, a: g.reset_to_equilibrium(),
    "single-north": lambda g, a: g.single_cell(1),
    "single-origin": lambda g, a: g.single_cell(4),
    "single-east": lambda g, a: g.single_cell(5),
    "single-north-east": lambda g, a: g.single_cell(2),
    "custom-speed": lambda g, a: g.custom_speed(a.speed),
}


def build_mask(a) -> np.ndarray:
    if a.circles:
        return gen_circles(a.nx, a.ny, a.seed, a.circles, a.r_min, a.r_max)
    mask = np.zeros((a.ny, a.nx), np.uint8)
    mask[0, :] = 1
    mask[-1, :] = 1
    yy, xx = np.mgrid[0:a.ny, 0:a.nx]
    for c in a.circle:
        x, y, r = (int(v) for v in c.split(","))
        mask[(xx - x) ** 2 + (yy - y) ** 2 <= r * r] = 1
    return mask


def run_script(a) -> int:
    from .interaction import Session
    prec = PRECISION_F32 if a.precision == "f32" else PRECISION_F16S
    s = Session(a.nx, a.ny, 1.0 / (3.0 * a.viscosity + 0.5), precision=prec, u_init=a.speed)
    fh = sys.stdin if a.script == "-" else open(a.script)
    try:
        for n, line in enumerate(fh, 1):
            try:
                frame = s.execute(line)
            except (ValueError, KeyError, IndexError, AttributeError) as e:
                print(f"line {n}: {e}", file=sys.stderr, flush=True)
                return 2
            if frame is not None:
                print(f"line {n}: step {s.lbm.get_compute_num()}  mass {s.lbm.conservation()[0]:.6f}", flush=True)
    finally:
        if fh is not sys.stdin:
            fh.close()
        s.close()
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--nx", type=int, default=400)
    ap.add_argument("--ny", type=int, default=160)
    ap.add_argument("--viscosity", type=float, default=0.02)
    ap.add_argument("--speed", type=float, default=0.1)
    ap.add_argument("--precision", choices=("f32", "f16s"), default="f32")
    ap.add_argument("--circle", action="append", default=[], help="x,y,r (repeatable)")
    ap.add_argument("--circles", type=int, default=0, help="seeded random circles instead")
    ap.add_argument("--r-min", type=int, default=8)
    ap.add_argument("--r-max", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--barrier", choices=("none", "curl", "chaos", "welcome"), default="none")
    ap.add_argument("--fluid-preset", choices=sorted(FLUID_PRESETS), default="custom-speed")
    ap.add_argument("--frames", type=int, default=10)
    ap.add_argument("--steps-per-frame", type=int, default=15, help="COMPUTE_PER_RENDER (lib.rs:17)")
    ap.add_argument("--step-mode", action="store_true", help="one step per frame")
    ap.add_argument("--summary", choices=[s.name.lower() for s in SummaryStat], default="curl")
    ap.add_argument("--colormap", choices=[c.name.lower() for c in ColorMap], default="jet")
    ap.add_argument("--out", default=None, help="directory for frame_NNNN.ppm images")
    ap.add_argument("--script", default=None, help="control commands, one per line ('-' = stdin)")
    a = ap.parse_args(argv)
    if a.script:
        return run_script(a)

    prec = PRECISION_F32 if a.precision == "f32" else PRECISION_F16S
    g = LBM(1.0 / (3.0 * a.viscosity + 0.5), a.nx, a.ny, precision=prec, u_init=a.speed)
    g.mask = build_mask(a)
    if a.barrier != "none":
        getattr(g, f"{a.barrier}_barrier")()
    FLUID_PRESETS[a.fluid_preset](g, a)
    g.set_summary(SummaryStat[a.summary.capitalize()])
    g.color_map = ColorMap[a.colormap.capitalize()]
    if a.out:
        os.makedirs(a.out, exist_ok=True)
    n = 1 if a.step_mode else a.steps_per_frame
    for frame in range(a.frames):
        g.sync()
        t0 = time.perf_counter()
        g.step(n)
        g.sync()
        dt = time.perf_counter() - t0
        rgb = g.render()
        if a.out:
            save_ppm(os.path.join(a.out, f"frame_{frame:04d}.ppm"), rgb)
        mass = g.conservation()[0]
        print(f"frame {frame:4d}  step {g.get_compute_num():7d}  {a.nx * a.ny * n / dt / 1e6:9.1f} MLUPS  "
              f"mass {mass:.6f}", flush=True)
    g.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
