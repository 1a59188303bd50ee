"""Online training of the surrogate on trajectories of the batched LBM.

    python -m lbm_amd.train --iters 2000 --factor 2 --out surrogate.npy

A batch of square domains (obstacles, tau and inflow drawn like the batched
configuration: 1-12 circles/rectangles, tau in [0.52, 0.8], u_in in
[0.02, 0.1]) is spun up with the plain LBM, then snapshot pairs (t, t + 64) are
read from it continually (the renderer moments: three planes per domain);
each training step draws a minibatch from a ring of recent pairs.  The network
sees (rho, u) in units S_RHO, S_U and learns their 64-step change, pooled over
F x F cells, in units D_RHO, D_U.  Every `refresh` advances the batch is
rebuilt from the next seed (new obstacles, tau and inflow), so the pairs keep
covering start-up transients as well as developed flow.  No reference
counterpart.
"""
from __future__ import annotations

import argparse
import time

import numpy as np

from .batch import Batch, gen_batch, gen_shapes
from .surrogate import (D_RHO, D_U, S_RHO, S_U, UNet, input_from_pooled, load_params, loss_weights, pool_fields,
                        save_params, target_from_pooled)


def velocity(jx, jy, rho):
    """(ux, uy) from momenta; cells with rho <= 0 (never in fluid) read 0."""
    safe = np.where(rho > 0, rho, 1.0)
    return np.where(rho > 0, jx / safe, 0.0), np.where(rho > 0, jy / safe, 0.0)


class PairSource:
    """Snapshot pairs (t, t + jump) from a batch of square domains.  A domain
    whose state stops being finite or physical (|u| or |rho - 1| above 0.3:
    tau near 1/2 with a fast inflow can leave BGK's stable range) is dropped.
    Targets are clipped to +-TARGET_CLIP output units, which only start-up
    transients reach."""

    TARGET_CLIP = 20.0

    def __init__(self, n_domains=16, size=256, factor=1, jump=64, spinup=1000, seed=7, device=0, tau=None,
                 u_in=None):
        params = gen_batch(seed, n_domains)
        self.factor = factor
        self.doms = [(size, size, 1.0 / tau if tau else om, u_in if u_in else u) for (_, _, om, u) in params]
        self.masks = [gen_shapes(size, size, seed * 1000 + i, 1, 12) for i in range(n_domains)]
        self.weights = [loss_weights(m, factor) for m in self.masks]
        self.batch = Batch(self.doms, device=device)
        for i, m in enumerate(self.masks):
            self.batch.set_mask(i, m)
        self.jump = jump
        self.dropped = set()
        self.batch.step(spinup)
        self.prev = [self._pooled(i) for i in range(n_domains)]

    def _pooled(self, i):
        jx, jy, rho = self.batch.moments(i)
        fl = self.masks[i] != 1
        ux, uy = velocity(jx, jy, rho)
        ok = np.isfinite(jx).all() and np.isfinite(jy).all() and np.isfinite(rho).all()
        if ok:
            ok = (np.abs(rho[fl] - 1.0).max() < 0.3 and np.abs(ux[fl]).max() < 0.3 and np.abs(uy[fl]).max() < 0.3)
        if not ok:
            self.dropped.add(i)
            return None
        return pool_fields(self.masks[i], rho, ux, uy, self.factor)

    def advance(self):
        """Step every domain `jump` steps; returns one (x, y, w) per live domain."""
        self.batch.step(self.jump)
        out = []
        for i in range(len(self.masks)):
            if i in self.dropped:
                continue
            cur = self._pooled(i)
            p = self.prev[i]
            self.prev[i] = cur
            if cur is None or p is None:
                continue
            om, u_in = self.doms[i][2], self.doms[i][3]
            x = input_from_pooled(p[0], p[1], p[2], 1.0 / om, u_in)
            y = np.clip(target_from_pooled(p[1], p[2], cur[2]), -self.TARGET_CLIP, self.TARGET_CLIP)
            out.append((x, y, self.weights[i]))
        return out

    def close(self):
        self.batch.close()


def train(iters=1000, minibatch=8, lr=1e-3, n_domains=16, size=256, factor=1, ring=256, seed=7, device=0,
          log_every=100, net=None, verbose=True, tau=None, u_in=None, advance_every=4, refresh=40, spinup=300,
          time_budget=None):
    """Train for `iters` iterations or until `time_budget` seconds have passed,
    whichever comes first; the cosine learning-rate decay follows whichever
    of the two is further along."""
    def source(k):
        return PairSource(n_domains=n_domains, size=size, factor=factor, seed=seed + 101 * k, device=device,
                          tau=tau, u_in=u_in, spinup=spinup)
    src, generation, advances, dropped = source(0), 0, 0, 0
    tile = size // factor
    if net is None:
        net = UNet(minibatch, tile, tile, seed=seed, device=device)
    rng = np.random.default_rng(seed)
    pool = []
    losses = []
    t0 = time.perf_counter()
    for it in range(iters):
        elapsed = time.perf_counter() - t0
        if time_budget is not None and elapsed >= time_budget:
            break
        progress = max(it / max(1, iters), elapsed / time_budget if time_budget else 0.0)
        if it % advance_every == 0 or len(pool) < minibatch:
            if advances == refresh:
                dropped += len(src.dropped)
                src.close()
                generation += 1
                src, advances = source(generation), 0
            pool.extend(src.advance())
            advances += 1
            pool = pool[-ring:]
        pick = rng.choice(len(pool), minibatch, replace=len(pool) < minibatch)
        x = np.stack([pool[j][0] for j in pick])
        y = np.stack([pool[j][1] for j in pick])
        w = np.stack([pool[j][2] for j in pick])
        cur_lr = lr * 0.5 * (1 + np.cos(np.pi * min(1.0, progress)))  # cosine decay
        losses.append(net.train_step(x, y, w, float(cur_lr)))
        if not np.isfinite(losses[-1]):
            raise FloatingPointError(f"non-finite training loss at iteration {it}")
        if verbose and it % log_every == 0:
            print(f"[train] iter {it} loss {losses[-1]:.4e} ({time.perf_counter() - t0:.1f} s)", flush=True)
    dropped += len(src.dropped)
    if verbose:
        print(f"[train] {generation + 1} batches of {n_domains} domains, {dropped} dropped as non-finite", flush=True)
    src.close()
    return net, np.array(losses)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=1000)
    ap.add_argument("--minibatch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--domains", type=int, default=16)
    ap.add_argument("--factor", type=int, default=1, choices=(1, 2))
    ap.add_argument("--size", type=int, default=0, help="domain side (default 256 x factor)")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--time-budget", type=float, default=None, help="stop after this many seconds")
    ap.add_argument("--out", default="surrogate.npy", help="parameters, bfloat16 bit patterns (.npy)")
    ap.add_argument("--init", default=None, help="start from these parameters (.npy, as --out writes)")
    a = ap.parse_args()
    size = a.size or 256 * a.factor
    net = None
    if a.init:
        net = UNet(a.minibatch, size // a.factor, size // a.factor, seed=a.seed)
        net.set_params(load_params(a.init))
    net, losses = train(a.iters, a.minibatch, a.lr, a.domains, size, a.factor, seed=a.seed,
                        time_budget=a.time_budget, net=net)
    save_params(a.out, net.get_params())
    print(f"saved {a.out}; loss {losses[:10].mean():.4e} -> {losses[-10:].mean():.4e}; "
          f"S_RHO={S_RHO} S_U={S_U} D_RHO={D_RHO} D_U={D_U} factor={a.factor}")


if __name__ == "__main__":
    main()
