"""Benchmark of the D2Q9 step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the largest single-GPU configuration the headline metric is quoted
on): D2Q9 BGK at 8192 x 4096, tau 0.53, inflow 0.08, 200 seeded circles with
radius 8..64 (seed 2024), walls on rows 0 and 4095, the reference's outlet.
A "step" is one lattice update of the whole grid.  The state is resident in HBM
before the timed region; K steps run between two stream synchronisations.

Printed: ONE JSON line with the FP32 MLUPS as `value`, the roofline of the
fused step kernel (72 algorithmic bytes per cell update times the updates one
launch makes -- two for the two-step variants, which read and write the state
once per pair of steps -- over the average launch duration of that kernel
from HIP events recorded on the grid's own stream; `frac_hbm` the same with
the launch's measured HBM bytes, calibrated FETCH_SIZE + WRITE_SIZE), the same
for FP16-shifted storage (36 B per cell update; f - w, and centred on the
free-stream equilibrium), both again at 4096 x 2048 and 16384 x 8192
(grid_leg), the batched configuration's MLUPS (batched_leg, 5000 steps), the
surrogate rollout's speed-up and velocity error against plain FP32 LBM
(surrogate_leg), and the CPU baseline: the oracle's faithful f32 restatement
of the reference's eight passes built for the host, timed on a bounded sample
of the same workload on all host cores and on one.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "MLUPS & %HBM roofline, FP32 vs FP16-store; surrogate speedup at ≤1% u-L2 err"
HBM_PEAK_GBS = 8000.0
SEED = 2024


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per launch of the step kernel from a committed rocprofv3 summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), flags=re.M)
    return float(m.group(1)) if m else None


KERNEL_NAMES = {1: "k_cell (one step per launch)",
                2: "k_pair (two steps per launch: the state read and written once per pair of steps; its last "
                   "blocks step the cells that read only themselves)"}


def updates_per_launch(variant):
    """Cell updates per cell in one launch of the step kernel: variants 8 and 9
    run two steps per launch (the state is read and written once per pair)."""
    return 2 if variant is not None and variant >= 8 else 1


def run_gpu(precision, X, Y, mask, steps, warmup, omega, u_in, variant=None):
    from lbm_amd import LBM, LbmError
    g = LBM(omega, X, Y, precision=precision)
    g.mask = mask
    g.custom_speed(u_in)
    tuned = None
    if variant is not None:
        g.variant = variant
        tuned = {"variant": variant, "us_per_step_by_variant": None}
    else:
        best, ms = g.autotune(100)  # state is restored afterwards; 100 steps keep the pick stable
        tuned = {"variant": best, "us_per_step_by_variant": [round(m * 1e3, 2) for m in ms]}
        log(f"autotune: variant {best}, us/step {tuned['us_per_step_by_variant']}")
    g.step(warmup)
    g.sync()
    t0 = time.perf_counter()
    dev_ms = g.step_timed(steps)  # waits for the last launch
    wall = time.perf_counter() - t0
    try:
        cons = g.conservation()
    except LbmError as e:  # the flow left BGK's stable range: timing stands, sums do not
        log(f"conservation: {e}")
        cons = None
    tuned["kernel_us"] = launch_us(g, tuned["variant"])
    g.close()
    return wall, dev_ms, cons, tuned


def launch_us(g, variant, launches=40):
    """Average duration of one launch of the step kernel, from HIP events on the
    grid's stream.  A call of n steps with a two-step variant ends with one-step
    launches, so the two-step launch is timed as the difference of a call of
    2 + 2m steps and one of 2 steps (m two-step launches)."""
    if updates_per_launch(variant) == 2:
        g.step_timed(2)
        short = min(g.step_timed(2) for _ in range(3))
        return round((g.step_timed(2 + 2 * launches) - short) / launches * 1e3, 2)
    return round(g.step_timed(launches) / launches * 1e3, 2)


SURROGATE_WEIGHTS = os.path.join(ROOT, "lbm_amd", "data", "surrogate_f2.npy")


def surrogate_leg(steps=10000):
    """Surrogate + LBM-corrector rollout against plain FP32 LBM on 1024 x 512.

    12 seeded circles (r 8..32, seed 2024), tau 0.6, u_in 0.05, 2000 plain
    spin-up steps; then `steps` steps of plain LBM and of the rollout (jump 64
    with the half-resolution network, 8 corrector steps) from the same state.
    Gated: a check runs the cycle both ways and keeps the plain result; a
    forecast error above half that of predicting no change distrusts the
    network until the next check (fallback to plain).  Checks start every 8th
    cycle and back off to every 64th while the network passes them with a
    clear margin (a fallback resets the interval).  Reported: the wall-clock
    speed-up (both sides end with a stream sync) gated, gated at a fixed
    interval of 8 cycles, and ungated, the velocity rel-L2
    against plain LBM at the end, the gate's residuals and decisions for the
    trained network and for the network-free ablation (the change predicted as
    zero: re-equilibration + correctors only), the short-horizon errors of
    both ungated, and the method's own floor (a perfect prediction: the plain
    LBM's 64 steps, then re-equilibration, then the correctors).  Both sides
    step with the autotuned step kernel."""
    import numpy as np

    from lbm_amd import LBM, gen_circles
    from lbm_amd.rollout import Rollout, compare, method_floor, velocity_error
    from lbm_amd.surrogate import UNet, load_params
    if not os.path.exists(SURROGATE_WEIGHTS):
        return {"skipped": f"no trained weights at {os.path.relpath(SURROGATE_WEIGHTS, ROOT)}"}
    X, Y, tau, u = 1024, 512, 0.6, 0.05
    threshold, check_every, max_check_every = 0.5, 8, 64
    mask = gen_circles(X, Y, SEED, 12, 8, 32)

    def make():
        g = LBM(1.0 / tau, X, Y)
        g.mask = mask
        g.custom_speed(u)
        g.step(2000)
        g.autotune(60)  # the fastest step kernel for both sides (state restored)
        return g
    net = UNet(1, Y // 2, X // 2, seed=1)
    params = load_params(SURROGATE_WEIGHTS)
    zero = np.zeros_like(params)
    out = {"workload": f"{X}x{Y}, 12 circles r 8-32 (seed {SEED}), tau {tau}, u_in {u}, after 2000 steps; "
                       f"jump 64 with a {X // 2}x{Y // 2} U-Net, 8 corrector steps; gate: a check every "
                       f"{check_every} cycles backing off to every {max_check_every}, threshold {threshold} of "
                       f"the persistence forecast's error",
           "steps": steps}
    short = {}
    horizons = (1, 2, 3, 5, 10)
    for label, p in (("trained", params), ("zero_change", zero)):
        net.set_params(p)
        errs = {}
        for cycles in horizons:
            a, b = make(), make()
            Rollout(a, net, check_every=0).run(72 * cycles)
            b.step(72 * cycles)
            errs[cycles] = round(velocity_error(a.fields(), b.fields(), mask), 5)
            a.close()
            b.close()
        short[label] = {f"u_rel_l2_after_{c}_cycles": e for c, e in errs.items()}
        short[label]["cycles_within_1pct"] = max([c for c in horizons if errs[c] <= 0.01], default=0)
    out["short_horizon_ungated"] = short
    net.set_params(params)
    sp, err, st, t_plain, t_roll = compare(make, net, steps, threshold=threshold, check_every=check_every,
                                           max_check_every=max_check_every)
    out.update({"speedup": round(sp, 3), "u_rel_l2_at_end": round(err, 5), "jumps": st.jumps, "checks": st.checks,
                "fallbacks": st.fallbacks, "residuals": [round(r, 4) for r in st.residuals],
                "check_errors": [round(e, 5) for e in st.errors],
                "plain_ms": round(t_plain * 1e3, 2), "rollout_ms": round(t_roll * 1e3, 2)})
    spf, errf, stf, _, tf = compare(make, net, steps, threshold=threshold, check_every=check_every)
    out["gated_every_8"] = {"speedup": round(spf, 3), "u_rel_l2_at_end": round(errf, 5), "jumps": stf.jumps,
                            "checks": stf.checks, "fallbacks": stf.fallbacks, "rollout_ms": round(tf * 1e3, 2)}
    sp0, err0, st0, _, t0 = compare(make, net, steps, check_every=0)
    out["ungated"] = {"speedup": round(sp0, 3), "u_rel_l2_at_end": round(err0, 5), "jumps": st0.jumps,
                      "rollout_ms": round(t0 * 1e3, 2)}
    net.set_params(zero)
    spz, errz, stz, _, _ = compare(make, net, steps, threshold=threshold, check_every=check_every,
                                   max_check_every=max_check_every)
    out["ablation_gated"] = {"speedup": round(spz, 3), "u_rel_l2_at_end": round(errz, 5), "jumps": stz.jumps,
                             "checks": stz.checks, "fallbacks": stz.fallbacks,
                             "residuals": [round(r, 4) for r in stz.residuals[:4]],
                             "check_errors": [round(e, 5) for e in stz.errors[:4]]}
    floor = {}
    a, b = make(), make()
    done = 0
    for cycles in (1, 10, steps // 72):
        method_floor(a, net, 72 * cycles - done)  # zero parameters: re-equilibration only
        b.step(72 * cycles - done)
        done = 72 * cycles
        floor[f"u_rel_l2_after_{cycles}_cycles"] = round(velocity_error(a.fields(), b.fields(), mask), 5)
    a.close()
    b.close()
    out["method_floor"] = floor
    net.set_params(params)
    out["unet_forward_ms"] = round(net.time_forward(10), 4)
    net.close()
    return out


def grid_leg(X, Y, steps):
    """FP32 and FP16-shifted MLUPS at another grid size (4096 x 2048, the small
    grid of the plain FP32 configuration; 16384 x 8192, the large grid of the
    FP16 configuration), same obstacle generator and parameters as the main
    line, variants autotuned."""
    from lbm_amd import PRECISION_F16S, PRECISION_F32, gen_circles
    mask = gen_circles(X, Y, SEED, 200, 8, 64)
    out = {"workload": f"{X}x{Y}, 200 circles r 8-64 (seed {SEED}), tau 0.53, u_in 0.08"}
    for name, prec, bpc in (("f32", PRECISION_F32, 72), ("f16s", PRECISION_F16S, 36)):
        wall, dev_ms, cons, tuned = run_gpu(prec, X, Y, mask, steps, 10, 1.0 / 0.53, 0.08)
        k = dev_ms / 1e3 / steps
        out[name] = {"mlups": round(X * Y / k / 1e6, 1), "us_per_step": round(k * 1e6, 2),
                     "achieved_gbs": round(bpc * X * Y / k / 1e9, 1), "variant": tuned["variant"],
                     "steps_per_launch": updates_per_launch(tuned["variant"]), "steps": steps,
                     "state_finite": cons is not None}
    out["f16s_speedup_vs_f32"] = round(out["f16s"]["mlups"] / out["f32"]["mlups"], 3)
    return out


def batched_leg(n_domains=512, steps=5000):
    """The batched configuration: 512 independent domains (sizes 128x64 ..
    1024x512 in multiples of 32, 1-12 circles/rectangles, tau in [0.52, 0.8],
    u_in in [0.02, 0.1], all from seed 2024) stepped together; tile variant
    autotuned, hipGraph replay on.  Whole-batch MLUPS per precision."""
    from lbm_amd import PRECISION_F16S, PRECISION_F32, Batch, gen_batch, gen_shapes
    doms = gen_batch(SEED, n_domains)
    masks = [gen_shapes(nx, ny, 10_000 + i, 1, 12) for i, (nx, ny, _, _) in enumerate(doms)]
    out = {"workload": f"{n_domains} domains from seed {SEED} (128x64..1024x512, 1-12 shapes, tau 0.52-0.8, "
                       f"u_in 0.02-0.1), {steps} timed steps, hipGraph replay"}
    for name, prec, bpc in (("f32", PRECISION_F32, 72), ("f16s", PRECISION_F16S, 36)):
        b = Batch(doms, precision=prec)
        for i, m in enumerate(masks):
            b.set_mask(i, m)
        b.step(4)
        cells = b.info()["cells"]
        best, _ = b.autotune(10)
        b.set_graph(True)
        b.step(20)
        b.sync()
        ms = b.step_timed(steps)
        s = ms / 1e3
        out[name] = {"mlups": round(cells * steps / s / 1e6, 1), "cells": cells, "variant": best,
                     "achieved_gbs": round(bpc * cells * steps / s / 1e9, 1), "us_per_step": round(s / steps * 1e6, 2)}
        b.close()
    return out


def _host_cpu():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(X, Y, mask, omega, u_in, budget_s):
    """The oracle's faithful f32 restatement of the reference's eight passes,
    built for this host (-O3 -march=native, no contraction; same bits as the
    oracle), timed on the same workload on all host cores (the value) and on
    one core; about budget_s seconds each after two warm-up steps."""
    import tempfile

    import numpy as np

    import oracle
    lib = oracle.build_native(tempfile.mkdtemp(prefix="lbm_cpu_baseline_"))
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()

    def sample(threads, budget):
        r = oracle.RefCPU(X, Y, omega, threads=threads, lib_path=lib)
        r.set_mask(mask.astype(np.uint32))
        r.reset_equilibrium(1.0, u_in, 0.0)
        r.step(1)
        t0 = time.perf_counter()
        r.step(1)
        t1 = time.perf_counter() - t0
        n = max(2, min(400, int(budget / max(t1, 1e-6))))
        t0 = time.perf_counter()
        r.step(n)
        t = time.perf_counter() - t0
        del r
        return X * Y * n / t / 1e6, n, t

    v_all, n_all, t_all = sample(cores, budget_s)
    v_one, n_one, t_one = sample(1, budget_s / 2)
    return {
        "value": round(v_all, 2),
        "unit": "MLUPS",
        "cores": cores,
        "kind": "port",
        "single_thread": round(v_one, 2),
        "host": f"{_host_cpu()}, {os.cpu_count()} CPUs visible, {cores} in this process's affinity",
        "build": "gcc -O3 -march=native -fopenmp -ffp-contract=off (bit for bit the oracle's build)",
        "sample": f"oracle refcpu (f32 restatement of the reference's 8 WGSL passes, OpenMP, one parallel loop "
                  f"per pass) on the same {X}x{Y} workload after 2 warm-up steps: {n_all} steps on {cores} "
                  f"threads in {t_all:.1f} s; {n_one} steps on 1 thread in {t_one:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--nx", type=int, default=8192)
    ap.add_argument("--ny", type=int, default=4096)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-f16", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=5000, help="timed steps of the batched leg")
    ap.add_argument("--variant", type=int, default=None, help="FP32 kernel variant (default: autotune)")
    ap.add_argument("--f16-variant", type=int, default=None, help="FP16-shifted kernel variant (default: autotune)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain copy of 1 GiB each way (y = x * 1.0 over 2^28 f32, three times): the "
                         "known-size kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE to HBM bytes")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.ones(2 ** 28, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print("copy calibration: 3 x (1 GiB read + 1 GiB written)", flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the D2Q9 step here runs on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build).", file=sys.stderr)
        sys.exit(2)

    from lbm_amd import PRECISION_F16S, PRECISION_F32, gen_circles

    X, Y = args.nx, args.ny
    omega, u_in = 1.0 / 0.53, 0.08
    cells = X * Y
    log(f"obstacles: 200 circles r 8..64 seed {SEED} on {X}x{Y}")
    mask = gen_circles(X, Y, SEED, 200, 8, 64)
    solid_frac = float(mask[1:-1].mean())

    log(f"f32: warmup {args.warmup}, timed {args.steps}")
    wall, dev_ms, cons, tuned = run_gpu(PRECISION_F32, X, Y, mask, args.steps, args.warmup, omega, u_in,
                                       args.variant)
    mlups = cells * args.steps / wall / 1e6
    upl = updates_per_launch(tuned["variant"])
    kern_s = tuned["kernel_us"] / 1e6 / upl  # per cell update pass, from the dominant kernel's launches
    achieved = 72.0 * cells / kern_s / 1e9
    log(f"f32: {mlups:.0f} MLUPS, {kern_s * 1e6:.1f} us per step, {upl} step(s) per launch, "
        f"{achieved:.0f} GB/s algorithmic")
    traffic = None
    if (X, Y) == (8192, 4096):
        traffic = read_traffic("traffic_f32_pair_8192x4096.txt" if upl == 2 else "traffic_f32_8192x4096.txt")

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"D2Q9 BGK FP32 {X}x{Y}, 200 seeded circles r 8-64 (seed {SEED}), tau 0.53, "
                        f"u_in 0.08, reference outlet",
            "nx": X, "ny": Y, "tau": 0.53, "u_in": u_in, "solid_fraction": round(solid_frac, 4),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": KERNEL_NAMES[upl],
            "algorithmic_bytes_per_launch": 72 * cells * upl,
            "cell_updates_per_launch": cells * upl,
            "kernel_avg_us": tuned["kernel_us"],
            # measured HBM bytes (calibrated FETCH_SIZE + WRITE_SIZE) per launch over the launch time
            "achieved_hbm": round(traffic / (tuned["kernel_us"] * 1e-6) / 1e9, 1) if traffic else None,
            "frac_hbm": round(traffic / (tuned["kernel_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
        },
        "fp64_mass_momentum": [float(c) for c in cons] if cons is not None else None,
        "autotune": tuned,
    }

    if not args.no_f16:
        log("f16s: same workload, FP16-shifted storage")
        w16, d16, c16, tuned16 = run_gpu(PRECISION_F16S, X, Y, mask, args.steps, args.warmup, omega,
                                        u_in, args.f16_variant)
        m16 = cells * args.steps / w16 / 1e6
        u16 = updates_per_launch(tuned16["variant"])
        k16 = tuned16["kernel_us"] / 1e6 / u16
        a16 = 36.0 * cells / k16 / 1e9
        log(f"f16s: {m16:.0f} MLUPS, {k16 * 1e6:.1f} us per step, {u16} step(s) per launch, "
            f"{a16:.0f} GB/s algorithmic")
        t16 = (read_traffic("traffic_f16s_8192x4096.txt" if u16 == 1 else "traffic_f16s_pair_8192x4096.txt")
               if (X, Y) == (8192, 4096) else None)
        out["f16s"] = {
            "mlups": round(m16, 1),
            "speedup_vs_f32": round(m16 / mlups, 3),
            "roofline": {"bound": "hbm", "achieved": round(a16, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(a16 / HBM_PEAK_GBS, 4), "kernel": KERNEL_NAMES[u16],
                         "algorithmic_bytes_per_launch": 36 * cells * u16,
                         "cell_updates_per_launch": cells * u16, "kernel_avg_us": tuned16["kernel_us"],
                         "traffic": t16,
                         "achieved_hbm": round(t16 / (tuned16["kernel_us"] * 1e-6) / 1e9, 1) if t16 else None,
                         "frac_hbm": round(t16 / (tuned16["kernel_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
                         if t16 else None},
            "fp64_mass_momentum": [float(c) for c in c16] if c16 is not None else None,
            "autotune": tuned16,
        }
        # against the FP32 one-step kernel (one pass over the state per step,
        # as the binary16 kernel makes): the ratio of the storage formats alone
        f32_one = min(tuned["us_per_step_by_variant"][:8]) if tuned.get("us_per_step_by_variant") else None
        if f32_one:
            out["f16s"]["speedup_vs_f32_one_step_kernel"] = round(f32_one / (k16 * 1e6), 3)
        log("f16e: same workload, binary16 centred on the free-stream equilibrium")
        from lbm_amd import PRECISION_F16E
        w16e, _, c16e, tuned16e = run_gpu(PRECISION_F16E, X, Y, mask, args.steps, args.warmup, omega, u_in,
                                          args.f16_variant)
        m16e = cells * args.steps / w16e / 1e6
        out["f16e"] = {"mlups": round(m16e, 1), "speedup_vs_f32": round(m16e / mlups, 3),
                       "kernel_avg_us": tuned16e["kernel_us"], "variant": tuned16e["variant"],
                       "fp64_mass_momentum": [float(c) for c in c16e] if c16e is not None else None}

    if not args.no_large and not args.no_f16:
        log("4096x2048 and 16384x8192: f32 and f16s")
        out["small_grid"] = grid_leg(4096, 2048, 2000)
        log(f"small grid: {out['small_grid']}")
        out["large_grid"] = grid_leg(16384, 8192, 2000)
        log(f"large grid: {out['large_grid']}")

    if not args.no_batched:
        log("batched: 512 domains")
        out["batched"] = batched_leg(steps=args.batched_steps)
        log(f"batched: {out['batched']}")

    if not args.no_surrogate:
        log("surrogate rollout vs plain LBM (1024x512, 10k steps)")
        out["surrogate"] = surrogate_leg()
        log(f"surrogate: {out['surrogate']}")

    if not args.no_cpu_baseline:
        log(f"cpu baseline: ~{args.cpu_budget:.0f} s of refcpu")
        out["cpu_baseline"] = cpu_baseline(X, Y, mask, omega, u_in, args.cpu_budget)

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
